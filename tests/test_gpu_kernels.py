"""Kernel-level parity of the HIP path against the C oracle (FP64).

Each batched entry point of libfastipc.so runs the same device kernels as
the Newton step.  Tolerances: energies and gradients 1e-10 relative (the
FP64 contract); SPD-projected Hessians 1e-9 relative to their norm (two
independent eigen-solvers); CCD time of impact within 1e-6 and never above
the oracle's by more than 1e-12.
"""
import ctypes as C

import numpy as np
import pytest

from fastipc_amd import _lib
from oracle import lib as O

pytestmark = pytest.mark.gpu

NPTS = (2, 3, 4, 4, 4, 4, 4)
NDOF = (6, 9, 12, 12, 12, 12, 12)


def _cases(kind, n, rng, dhat2):
    """Point sets whose relevant distance is below dhat2 (barrier active)."""
    k = NPTS[kind]
    out = []
    while len(out) < n:
        x = rng.uniform(-1, 1, size=(k, 3)) * 0.05
        x0 = x + rng.normal(size=(k, 3)) * 0.01
        if kind >= 4:  # nearly parallel edges so the mollifier is active
            d = rng.normal(size=3)
            x[1] = x[0] + d * 0.05
            x[3] = x[2] + (d + rng.normal(size=3) * 0.02) * 0.05
            x0 = x.copy()
            x0[3] = x0[2] + (d + rng.normal(size=3) * 0.3) * 0.05
        E = O.constraint_E(kind, x, x0, dhat2, 1.0)
        if E > 0:
            out.append((x, x0))
    return out


@pytest.mark.parametrize("kind", range(7))
def test_constraint_gH_matches_oracle(kind):
    rng = np.random.default_rng(100 + kind)
    dhat2 = 1e-3
    cases = _cases(kind, 64, rng, dhat2)
    n = len(cases)
    k = NPTS[kind]
    nd = NDOF[kind]
    X = np.ascontiguousarray(np.stack([c[0] for c in cases]), np.float64)
    X0 = np.ascontiguousarray(np.stack([c[1] for c in cases]), np.float64)
    for project in (0, 1):
        E = np.zeros(n)
        g = np.zeros((n, nd))
        H = np.zeros((n, nd, nd))
        _lib.call("fipc_batch_constraint", kind, n, _lib.dptr(X), _lib.dptr(X0), dhat2, 1.0, project,
                  _lib.dptr(E), _lib.dptr(g), _lib.dptr(H))
        for i in range(n):
            Eo, go, Ho = O.constraint_gH(kind, X[i], X0[i], dhat2, 1.0, project=bool(project))
            assert abs(E[i] - Eo) <= 1e-10 * abs(Eo), (kind, i, E[i], Eo)
            assert np.abs(g[i] - go).max() <= 1e-10 * np.abs(go).max(), (kind, i)
            scale = np.abs(Ho).max()
            assert np.abs(H[i] - Ho).max() <= 1e-9 * scale, (kind, project, i, np.abs(H[i] - Ho).max() / scale)


def test_tet_matches_oracle():
    rng = np.random.default_rng(7)
    n = 200
    # deformation gradients with distinct singular values: at repeated singular
    # values (F = I) the reference's 2x2 make_pd (math_tools.py:39-64) is
    # discontinuous and the result depends on the last bit of rounding
    F = np.eye(3)[None] + rng.normal(size=(n, 3, 3)) * 0.3
    bad = np.linalg.det(F) < 0.2
    F[bad] = np.diag([1.3, 1.0, 0.8])
    sv = np.linalg.svd(F, compute_uv=False)
    assert np.min(np.diff(-sv, axis=1)) > 1e-6
    Dm = np.eye(3)[None] * 0.1 + rng.normal(size=(n, 3, 3)) * 0.01
    IB = np.linalg.inv(Dm)
    la = rng.uniform(1e4, 1e6, size=n)
    mu = rng.uniform(1e4, 1e6, size=n)
    scale = rng.uniform(1e-6, 1e-3, size=n)
    psi = np.zeros(n)
    P = np.zeros((n, 3, 3))
    H = np.zeros((n, 12, 12))
    args = [np.ascontiguousarray(a, np.float64) for a in (F, IB, scale, la, mu)]
    _lib.call("fipc_batch_tet", n, *[_lib.dptr(a) for a in args], _lib.dptr(psi), _lib.dptr(P), _lib.dptr(H))
    L = O.lib()
    for i in range(n):
        Fi = np.ascontiguousarray(F[i])
        U, s, V = O.svd3(Fi)
        psio = L.orc_nh_energy(O.dptr(np.ascontiguousarray(s)), la[i], mu[i])
        assert abs(psi[i] - psio) <= 1e-10 * max(1.0, abs(psio)) * mu[i], i
        Po = O.nh_P(Fi, la[i], mu[i])
        assert np.abs(P[i] - Po).max() <= 1e-10 * np.abs(Po).max()
        # oracle tet Hessian through the reference's intermediate contraction:
        # build a rest shape with D_m = Dm[i] and deformed vertices giving F
        x = np.zeros((4, 3))
        x[1:] = (Fi @ Dm[i]).T
        restT = np.ascontiguousarray(Dm[i])
        vol = np.linalg.det(restT) / 6.0
        dt2 = scale[i] / vol
        g12 = np.zeros(12)
        H12 = np.zeros(144)
        L.orc_tet_gH(O.dptr(np.ascontiguousarray(x)), O.iptr(np.arange(4, dtype=np.int32)), O.dptr(restT),
                     la[i], mu[i], np.sqrt(dt2), O.dptr(g12), O.dptr(H12))
        Ho = H12.reshape(12, 12)
        assert np.abs(H[i] - Ho).max() <= 1e-9 * np.abs(Ho).max(), (i, np.abs(H[i] - Ho).max() / np.abs(Ho).max())


@pytest.mark.parametrize("kind", [0, 1])
def test_ccd_matches_oracle(kind):
    rng = np.random.default_rng(11 + kind)
    n = 4000
    pts = rng.uniform(0, 1, size=(n, 4, 3))
    disp = rng.normal(size=(n, 4, 3)) * rng.uniform(0.1, 1.5, size=(n, 1, 1))
    d2 = np.zeros(n)
    Lb = O.lib()
    for i in range(n):
        P = [O.dptr(np.ascontiguousarray(v)) for v in pts[i]]
        if kind == 0:
            d2[i] = Lb.orc_PT_dist2(*P, O.pt_type(*pts[i]))
        else:
            d2[i] = Lb.orc_EE_dist2(*P, O.ee_type(*pts[i]))
    toc = np.zeros(n)
    pa = np.ascontiguousarray(pts.reshape(n, 12))
    da = np.ascontiguousarray(disp.reshape(n, 12))
    _lib.call("fipc_batch_ccd", kind, n, _lib.dptr(pa), _lib.dptr(da), 0.2, _lib.dptr(d2), _lib.dptr(toc))
    ref = np.array([(O.pt_ccd if kind == 0 else O.ee_ccd)(*pts[i], *disp[i], 0.2, d2[i]) for i in range(n)])
    assert (ref < 1).sum() > 100
    assert np.abs(toc - ref).max() <= 1e-6
    assert np.all(toc <= ref + 1e-12)


def test_ccd_scene_pairs_match_oracle_and_reference():
    """The narrow phase on the pairs real scene states produce (golden
    reference_ccd_scenes.npz, the reference's compiled CCD): equal to the
    oracle (both the smaller of the strict and fused-multiply-add
    evaluations), never later than the reference except its proven false
    hits (test_oracle_pins.test_ccd_scene_pairs_vs_reference)."""
    import os
    F = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_ccd_scenes.npz"))
    for kind in (0, 1):
        m = F["kind"] == kind
        pts, disp, d2, ref = F["pts"][m], F["disp"][m], np.ascontiguousarray(F["d2"][m]), F["toc"][m]
        n = len(ref)
        toc = np.zeros(n)
        _lib.call("fipc_batch_ccd", kind, n, _lib.dptr(np.ascontiguousarray(pts.reshape(n, 12))),
                  _lib.dptr(np.ascontiguousarray(disp.reshape(n, 12))), 0.2, _lib.dptr(d2), _lib.dptr(toc))
        orc = np.array([(O.pt_ccd if kind == 0 else O.ee_ccd)(*pts[i], *disp[i], 0.2, d2[i]) for i in range(n)])
        assert np.abs(toc - orc).max() <= 1e-12, int(np.sum(np.abs(toc - orc) > 1e-12))
        later = np.flatnonzero(toc > ref + 1e-9)
        assert np.all(orc[later] > ref[later] + 1e-9) and len(later) <= 2
        print(f"kind {kind}: {n} pairs, {int(np.sum(np.abs(toc - ref) > 1e-6))} outside 1e-6 of the reference, "
              f"{len(later)} later (false hits of the reference)")


def test_inversion_bound_matches_oracle():
    """The inversion part of the line-search filter (NEWTON_IPC.py:443-453)
    equals the oracle's minimum over every tet, for displacements that invert
    some tets and for small ones that invert none."""
    from fastipc_amd import scene as S
    from fastipc_amd.newton import NewtonIPC
    sc = S.two_cubes(4)
    sim = NewtonIPC(sc)
    sim.compute_xn_and_xTilde(0.0)
    rng = np.random.default_rng(5)
    x = sim.x
    for scale in (0.3, 0.05, 1e-3):
        p = rng.normal(size=x.shape) * scale * sc.bbox_diag
        sim.set_vector(1, p)
        a = C.c_double()
        parts = np.zeros(3)
        _lib.call("fipc_ccd_alpha", sim.ctx, sc.eta, 1, C.byref(a), _lib.dptr(parts))
        ref = 1.0
        for t in sc.tets:
            ref = min(ref, O.tet_inversion_alpha([x[i] for i in t], [p[i] for i in t], 0.2))
        assert parts[2] == pytest.approx(ref, rel=1e-9, abs=1e-12), (scale, parts[2], ref)
        print(f"scale {scale}: inversion bound {parts[2]}")


def test_begin_step_rows_matches_full_target():
    """fipc_begin_step_rows (Dirichlet rows only) gives the same xTilde as
    fipc_begin_step with the full target array; bad rows are rejected."""
    from fastipc_amd import scene as S
    from fastipc_amd.newton import NewtonIPC
    sc = S.compressed_slab(6, 8, 6)
    a, b = NewtonIPC(sc), NewtonIPC(sc)
    for sim in (a, b):
        sim.step()
    t = 0.37
    full = np.ascontiguousarray(sc.dirichlet(t), np.float64)
    _lib.call("fipc_begin_step", a.ctx, _lib.dptr(full))
    rows, xyz = sc.dirichlet_rows(t)
    _lib.call("fipc_begin_step_rows", b.ctx, len(rows), _lib.iptr(rows), _lib.dptr(xyz))
    assert np.array_equal(a.vector(3), b.vector(3))
    bad = np.array([a.nv], np.int32)
    with pytest.raises(_lib.FipcError):
        _lib.call("fipc_begin_step_rows", b.ctx, 1, _lib.iptr(bad), _lib.dptr(xyz[:1].copy()))
