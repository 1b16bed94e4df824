"""Newton-step parity of the HIP path against the oracle on the same seeded
scenes: constraint sets bit-exact (as np.unique rows), energies and
gradients within 1e-10 relative, the assembled Hessian through products with
random vectors (on deformed states, away from the rest-state discontinuity of
the reference's 2x2 make_pd), the step filters, and positions after whole
time steps.

Position parity is tested with the Newton tolerance tightened to 1e-4 of its
default so that both runs converge to the same minimiser (with the default
tolerance two correct solvers may stop anywhere within the tolerance); the
default-tolerance GPU step is checked by evaluating the oracle's own Newton
criterion at the GPU's final state."""
import numpy as np
import pytest

from fastipc_amd import _lib
from fastipc_amd import scene as S
from fastipc_amd.newton import NewtonIPC
from oracle.newton import OracleIPC

pytestmark = pytest.mark.gpu


def _deformed(sc, amp=0.02, seed=0):
    """Smooth, seeded deformation of the free vertices (every tet strained)."""
    rng = np.random.default_rng(seed)
    x = sc.x.copy()
    k = rng.normal(size=(3, 3)) * 3.0
    ph = rng.uniform(0, 2 * np.pi, size=3)
    free = ~sc.dirichlet_mask
    x[free] += amp * np.sin(x[free] @ k + ph)
    return x


def _sync(g, o, x, v=None, frame=0):
    v = np.zeros_like(x) if v is None else v
    o.x = np.ascontiguousarray(x.copy())
    o.v = v.copy()
    g.set_state(x, v)
    g.compute_xn_and_xTilde(frame * o.dt)
    o.compute_xn_and_xTilde(frame * o.dt)


def _spmv(g, v):
    vv = np.ascontiguousarray(v.reshape(-1, 3))
    yy = np.zeros_like(vv)
    _lib.call("fipc_spmv", g.ctx, _lib.dptr(vv), _lib.dptr(yy))
    return yy.ravel()


@pytest.mark.parametrize("scene_fn,steps", [(S.tiny_pair, 2), (lambda: S.two_cubes(5), 7)])
def test_constraints_energy_gradient_hessian(scene_fn, steps):
    sc = scene_fn()
    o = OracleIPC(sc)
    for _ in range(steps):
        o.step()
    g = NewtonIPC(sc)
    # 1) contact state reached by the oracle: sets, energy, gradient
    _sync(g, o, o.x.copy(), o.v.copy(), o.frame)
    g.find_constraints()
    o.find_constraints()
    cg, co = g.constraints(), o.constraints()
    assert sum(len(v) for v in co.values()) > 0
    for k in co:
        assert np.array_equal(cg[k], co[k]), k
    Eg, pg = g.compute_energy(parts=True)
    Eo, po = o.compute_energy(parts=True)
    assert abs(Eg - Eo) <= 1e-10 * abs(Eo)
    for a, b in zip(pg, po):
        assert abs(a - b) <= 1e-10 * abs(b) or abs(a - b) < 1e-14 * abs(Eo)
    g.compute_hessian_and_gradient()
    o.assemble()
    D = np.repeat(o.dmask, 3)
    ro = o.rhs.copy()
    ro[D] = 0.0  # rhs[D] = 0 (NEWTON_IPC.py:772); the GPU does not assemble fixed rows
    assert np.abs(g.vector(2).ravel() - ro).max() <= 1e-10 * np.abs(ro).max()
    # 2) deformed copy of that state: every tet strained, Hessian products
    x = o.x.copy()
    free = ~sc.dirichlet_mask
    rng = np.random.default_rng(1)
    x[free] += 1e-3 * np.sin(x[free] @ rng.normal(size=(3, 3)) * 5.0)
    _sync(g, o, x, o.v.copy(), o.frame)
    g.find_constraints()
    o.find_constraints()
    g.compute_hessian_and_gradient()
    o.assemble()
    ro = o.rhs.copy()
    ro[D] = 0.0
    assert np.abs(g.vector(2).ravel() - ro).max() <= 1e-10 * np.abs(ro).max()
    A = o.matrix()
    for _ in range(3):
        v = rng.normal(size=3 * o.nv)
        v[D] = 0.0
        yo = A @ v
        yo[D] = 0.0
        yg = _spmv(g, v)
        assert np.abs(yg - yo).max() <= 1e-9 * np.abs(yo).max()


def test_filters_match_oracle():
    sc = S.two_cubes(5)
    o = OracleIPC(sc)
    for _ in range(6):
        o.step()
    g = NewtonIPC(sc)
    _sync(g, o, o.x.copy(), o.v.copy(), 6)
    o.find_constraints()
    o.assemble()
    o.solve_system()
    p = o.p.copy()
    rng = np.random.default_rng(5)
    for scale in (1.0, 30.0, 300.0):
        pp = p * scale + rng.normal(size=p.shape) * 1e-4 * scale
        o.p = pp
        g.set_vector(1, pp)
        ag, partg = g.compute_intersection_free_step_size(parts=True)
        ao, parto = o.compute_intersection_free_step_size(parts=True)
        assert abs(ag - ao) <= 1e-6, (scale, partg, parto)
        assert ag <= ao + 1e-12
    assert g.check_collision() == o.check_collision() == 0
    x2 = o.x.copy()
    x2[~o.dmask, 1] -= 0.5  # pushes the cubes through the ground
    o.x = x2
    g.set_state(x2)
    assert g.check_collision() == o.check_collision() == 1


def _ccd_brute_minimum(x, p, bp, be, bt, dHat, eta):
    """Least time of impact of fipc_batch_ccd over every point-triangle and
    edge-edge pair that passes the CCD pass's box and lower-bound filters
    (compute_filter_3D_PT/EE, NEWTON_IPC.py:377-440), enumerated by brute
    force: (point-triangle minimum, edge-edge minimum), 1.0 for none."""
    def n2(a):
        return a[..., 0] * a[..., 0] + a[..., 1] * a[..., 1] + a[..., 2] * a[..., 2]

    def boxes(v):  # (n, k) vertex ids -> swept boxes
        a, b = x[v], x[v] + p[v]
        return np.minimum(a, b).min(axis=1), np.maximum(a, b).max(axis=1)

    out = []
    for kind in (0, 1):
        if kind == 0:
            va, vb = bp[:, None], bt
        else:
            va, vb = be, be
        la, ha = boxes(va)
        lb, hb = boxes(vb)
        ii, jj = np.nonzero(np.all((la[:, None] < hb[None] + dHat) & (lb[None] - dHat < ha[:, None]), axis=2))
        if kind == 1:
            keep = ii < jj
            ii, jj = ii[keep], jj[keep]
        shared = (va[ii][:, :, None] == vb[jj][:, None, :]).any(axis=(1, 2))
        ii, jj = ii[~shared], jj[~shared]
        ids = np.concatenate([va[ii], vb[jj]], axis=1)  # (m, 4): P T0 T1 T2 / A0 A1 B0 B1
        dn = np.sqrt(n2(p[ids]))
        if kind == 0:
            md = dn[:, 0] + np.maximum(np.maximum(dn[:, 1], dn[:, 2]), dn[:, 3])
        else:
            md = np.maximum(dn[:, 0], dn[:, 1]) + np.maximum(dn[:, 2], dn[:, 3])
        ids, md = ids[md != 0], md[md != 0]
        tf, df = ("PT_type", "PT_dist2") if kind == 0 else ("EE_type", "EE_dist2")
        d2 = np.array([_lib.scalar(df, *x[q].ravel(), _lib.scalar(tf, *x[q].ravel())[0])[0] for q in ids])
        keep = (1 - eta) * np.sqrt(d2) / md < 1
        ids, d2 = ids[keep], np.ascontiguousarray(d2[keep])
        n = len(ids)
        if n == 0:
            out.append(1.0)
            continue
        toc = np.zeros(n)
        _lib.call("fipc_batch_ccd", kind, n, _lib.dptr(np.ascontiguousarray(x[ids].reshape(n, 12))),
                  _lib.dptr(np.ascontiguousarray(p[ids].reshape(n, 12))), eta, _lib.dptr(d2), _lib.dptr(toc))
        out.append(min(1.0, float(toc.min())))
    return out


def test_ccd_pass_equals_batch_minimum():
    """The CCD pass (swept-box grid queries, filters, narrow phase with the
    strict and fused roundings in separate wavefronts) gives, bit for bit,
    the least fipc_batch_ccd time of impact over the brute-force pair set:
    stacked cubes in contact, along the Newton direction scaled 10 and 30
    times, whose pairs reach the root finding."""
    sc = S.two_cubes(5)
    o = OracleIPC(sc)
    for _ in range(6):
        o.step()
    g = NewtonIPC(sc)
    _sync(g, o, o.x.copy(), o.v.copy(), 6)
    o.find_constraints()
    o.assemble()
    o.solve_system()
    p0 = o.p.copy()
    reached = 0
    for scale in (10.0, 30.0):
        p = np.ascontiguousarray(p0 * scale)
        g.set_vector(1, p)
        _, parts = g.compute_intersection_free_step_size(parts=True)
        pairs = np.zeros(2, np.int32)
        _lib.call("fipc_ccd_counts", g.ctx, _lib.iptr(pairs))
        ref = _ccd_brute_minimum(o.x, p, o.bp.astype(np.int64), o.be.astype(np.int64), o.bt.astype(np.int64),
                                 o.dHat, o.eta)
        print(f"scale {scale}: pairs {pairs.tolist()}, parts {list(parts[:2])}, brute force {ref}")
        assert parts[0] == ref[0] and parts[1] == ref[1]
        reached += min(ref) < 1.0
    assert reached >= 1


@pytest.mark.parametrize("scene_fn,steps,amp", [(S.tiny_pair, 4, 0.001), (lambda: S.two_cubes(4), 12, 0.004)])
def test_steps_match_oracle(scene_fn, steps, amp):
    """Whole time steps from the same pre-strained state (every tet away from
    F = I, where the reference's 2x2 make_pd jumps) with the Newton tolerance
    tightened 1e3 times: positions agree within 1e-6 of the bbox diagonal."""
    sc = scene_fn()
    sc.tol_rel *= 1e-3
    g = NewtonIPC(sc, pcg_tol=1e-12)
    o = OracleIPC(sc)
    x = _deformed(sc, amp=amp, seed=2)
    o.x = np.ascontiguousarray(x.copy())
    assert o.check_collision() == 0
    g.set_state(x)
    for f in range(steps):
        sg = g.step()
        so = o.step()
        assert sg.converged
        assert g.check_collision() == 0
        err = np.abs(g.x - o.x).max()
        assert err <= 1e-6 * sc.bbox_diag, (f, err, sg.newton_iters, so["newton_iters"])


@pytest.mark.parametrize("scene_fn,steps", [(S.tiny_pair, 3), (lambda: S.two_cubes(5), 10)])
def test_step_converged_per_oracle(scene_fn, steps):
    """Default tolerance: after each GPU step the oracle, evaluated at the GPU's
    state, finds the same constraint sets, the same energy and a Newton
    direction below the stopping tolerance."""
    sc = scene_fn()
    g = NewtonIPC(sc)
    o = OracleIPC(sc)
    for f in range(steps):
        xn = g.x
        vn = g.v
        st = g.step()
        assert st.converged and g.check_collision() == 0
        # oracle at the GPU's end state of this step
        o.x, o.v = xn.copy(), vn.copy()
        o.compute_xn_and_xTilde(f * sc.dt)
        o.x = np.ascontiguousarray(g.x)
        o.find_constraints()
        cg = g.constraints()
        for k, rows in o.constraints().items():
            assert np.array_equal(cg[k], rows), (f, k)
        o.assemble()
        o.solve_system(f * sc.dt)
        assert o.output_residual() < sc.newton_tol * 1.5, (f, o.output_residual(), sc.newton_tol)


def test_products_need_the_current_assembly():
    """A product or solve after constraints were found again, without an
    assembly, is refused (FIPC_ERR_STATE): its contact part would pair the
    last assembly's Hessians with the new rows.  After the assembly both
    work again."""
    sc = S.tiny_pair()
    g = NewtonIPC(sc)
    g.compute_xn_and_xTilde(0.0)
    g.find_constraints()
    g.compute_hessian_and_gradient()
    g.solve_system()
    g.find_constraints()
    with pytest.raises(_lib.FipcError) as e:
        g.solve_system()
    assert e.value.code == 6
    v = np.zeros((sc.n_verts, 3))
    y = np.zeros((sc.n_verts, 3))
    assert g.so.fipc_spmv(g.ctx, _lib.dptr(v), _lib.dptr(y)) == 6
    g.compute_hessian_and_gradient()
    it, rel = g.solve_system()
    assert rel <= g.pcg_tol * 1.0001
