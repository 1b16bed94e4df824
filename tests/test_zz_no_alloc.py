"""fipc_batch_ccd_noalloc (the batched CCD with its device copies in a fixed
device array) against fipc_batch_ccd (buffers allocated per call), through
the C-ABI: bit-identical step bounds on two calls in a row, the second with
fewer pairs, and a clear error above the maximum."""
import ctypes as C

import numpy as np
import pytest

from fastipc_amd import _lib
from oracle import lib as O

pytestmark = pytest.mark.gpu

MAX = 16384


def _pairs(n, kind, seed):
    """Seeded CCD inputs (points, displacements, current squared distances)
    like the reference fixtures: most pairs approach within the step."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, size=(n, 4, 3))
    disp = rng.normal(size=(n, 4, 3)) * rng.uniform(0.1, 1.5, size=(n, 1, 1))
    L = O.lib()
    d2 = np.empty(n)
    for i in range(n):
        P = [O.dptr(np.ascontiguousarray(v)) for v in pts[i]]
        d2[i] = L.orc_PT_dist2(*P, O.pt_type(*pts[i])) if kind == 0 else L.orc_EE_dist2(*P, O.ee_type(*pts[i]))
    return np.ascontiguousarray(pts.reshape(n, 12)), np.ascontiguousarray(disp.reshape(n, 12)), d2


@pytest.mark.parametrize("kind", [0, 1])
def test_noalloc_batch_ccd_matches_allocating_entry_point(kind):
    so = _lib.load()
    for n, seed in ((MAX, 1 + kind), (1500, 11 + kind)):  # second call smaller
        pts, disp, d2 = _pairs(n, kind, seed)
        want = np.zeros(n)
        got = np.full(n, -7.0)
        assert so.fipc_batch_ccd(kind, n, _lib.dptr(pts), _lib.dptr(disp), 0.2, _lib.dptr(d2), _lib.dptr(want)) == 0
        assert so.fipc_batch_ccd_noalloc(kind, n, _lib.dptr(pts), _lib.dptr(disp), 0.2, _lib.dptr(d2),
                                         _lib.dptr(got)) == 0
        assert np.array_equal(got, want), (n, int(np.sum(got != want)))
        assert np.sum(want < 1.0) > n // 20  # the narrow phase finds hits
    big = np.zeros(12 * (MAX + 1))
    one = np.zeros(MAX + 1)
    assert so.fipc_batch_ccd_noalloc(kind, MAX + 1, _lib.dptr(big), _lib.dptr(big), 0.2, _lib.dptr(one),
                                     _lib.dptr(one)) == 3  # FIPC_ERR_CAPACITY
