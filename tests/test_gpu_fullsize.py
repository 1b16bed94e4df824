"""Parity at the benchmark scene sizes.  The oracle cannot run whole steps at
these sizes (its direct solve would take minutes), so the GPU runs the steps
and the oracle evaluates the GPU's state: the constraint sets must match
bit-exactly (np.unique rows), the incremental potential within 1e-10
relative (and the gradient for the small scene), and the oracle's own
collision check must find no intersection."""
import time

import numpy as np
import pytest

from fastipc_amd import scene as S
from fastipc_amd.newton import NewtonIPC
from oracle.newton import OracleIPC

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,steps", [(0, 20), (1, 8), (2, 6)])
def test_full_size_state_parity(cfg, steps):
    sc = S.config(cfg)
    t0 = time.perf_counter()
    g = NewtonIPC(sc)
    for _ in range(steps):
        st = g.step()
        assert st.converged
        assert g.check_collision() == 0
    t1 = time.perf_counter()
    o = OracleIPC(sc)
    o.x = np.ascontiguousarray(g.x)
    o.v = g.v
    o.frame = g.frame
    o.compute_xn_and_xTilde(g.frame * sc.dt)
    g.compute_xn_and_xTilde(g.frame * sc.dt)
    g.find_constraints()
    o.find_constraints()
    cg, co = g.constraints(), o.constraints()
    for k in co:
        assert np.array_equal(cg[k], co[k]), (cfg, k, len(cg[k]), len(co[k]))
    Eg, pg = g.compute_energy(parts=True)
    Eo, po = o.compute_energy(parts=True)
    assert abs(Eg - Eo) <= 1e-10 * abs(Eo), (Eg, Eo)
    if cfg == 0:
        g.compute_hessian_and_gradient()
        o.assemble()
        D = np.repeat(o.dmask, 3)
        ro = o.rhs.copy()
        ro[D] = 0.0
        assert np.abs(g.vector(2).ravel() - ro).max() <= 1e-10 * np.abs(ro).max()
    assert o.check_collision() == 0
    n = {k: len(v) for k, v in co.items()}
    print(f"config {cfg}: {sc.n_verts} verts, {sc.n_tets} tets, {steps} GPU steps in {t1 - t0:.2f} s, "
          f"contacts {n}, energy {Eg:.12e}")
