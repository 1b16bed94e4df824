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
