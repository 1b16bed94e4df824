"""FP32 Hessian mode against FP64: the elastic blocks are stored and multiplied
in FP32 while PCG vectors, contacts, energies and CCD stay FP64; the PCG is
restarted from the FP64 residual b - A x until that residual meets the
tolerance.  Checked here: the refined direction meets the relative tolerance
of the system FP32 mode solves (FP32-stored elastic blocks, FP64 contact
rows) and, at the benchmark tolerance, the one measured with the FP64
Hessian; it agrees with the FP64-mode direction, and
whole time steps agree with the oracle within 1e-6 of the bbox diagonal at a
tightened Newton tolerance."""
import ctypes as C

import numpy as np
import pytest

from fastipc_amd import _lib
from fastipc_amd import scene as S
from fastipc_amd.newton import NewtonIPC
from oracle.newton import OracleIPC

pytestmark = pytest.mark.gpu


def _deformed(sc, amp, seed):
    rng = np.random.default_rng(seed)
    x = sc.x.copy()
    k = rng.normal(size=(3, 3)) * 3.0
    ph = rng.uniform(0, 2 * np.pi, size=3)
    free = ~sc.dirichlet_mask
    x[free] += amp * np.sin(x[free] @ k + ph)
    return x


def _spmv(g, v):
    vv = np.ascontiguousarray(v.reshape(-1, 3))
    yy = np.zeros_like(vv)
    _lib.call("fipc_spmv", g.ctx, _lib.dptr(vv), _lib.dptr(yy))
    return yy.ravel()


@pytest.mark.parametrize("tol", [1e-6, 1e-10])
def test_fp32_solve_meets_fp64_residual(tol):
    sc = S.two_cubes(5)
    x = _deformed(sc, 0.004, 3)
    ctx = {}
    for prec in (0, 1):
        g = NewtonIPC(sc, precision=prec, pcg_tol=tol)
        g.set_state(x)
        g.compute_xn_and_xTilde(0.0)
        g.find_constraints()
        g.compute_hessian_and_gradient()
        it, rel = g.solve_system()
        ctx[prec] = (g, it, rel)
    g64 = ctx[0][0]
    b = g64.vector(2).ravel()
    D = np.repeat(sc.dirichlet_mask, 3)
    p32 = ctx[1][0].vector(1).ravel()
    p64 = g64.vector(1).ravel()
    assert np.all(p32[D] == 0.0)  # static Dirichlet vertices: zero increment
    # the system FP32 mode solves: its FP32-stored elastic blocks (products in
    # FP64 arithmetic) and the FP64 contact rows; the reported residual is
    # that system's, recomputed in FP64
    g32 = ctx[1][0]
    rs = b - _spmv(g32, p32)
    rs[D] = 0.0
    rels = np.linalg.norm(rs) / np.linalg.norm(b)
    assert rels <= tol * 1.0001, (rels, ctx[1][1:])
    assert ctx[1][2] == pytest.approx(rels, rel=1e-3, abs=1e-3 * tol)
    # against the FP64 Hessian the residual differs by the blocks' rounding
    # applied to the direction, (A32 - A64) p32, computed here exactly as
    # the difference of the two products: at the benchmark tolerance (1e-6)
    # that floor is far below tol, so both modes meet the same FP64 residual
    r32 = b - _spmv(g64, p32)
    r32[D] = 0.0
    rel32 = np.linalg.norm(r32) / np.linalg.norm(b)
    dA = _spmv(g32, p32) - _spmv(g64, p32)
    dA[D] = 0.0
    floor = np.linalg.norm(dA) / np.linalg.norm(b)
    assert rel32 <= (tol + floor) * 1.0001, (rel32, tol, floor)
    assert floor <= 1e-6  # a few FP32 units of the blocks, amplified by |p|/|b|
    if tol >= 1e-6:
        assert rel32 <= tol * 1.0001, (rel32, floor)
    # both directions solve the same system to tol: they differ by at most cond * tol
    err = np.abs(p32 - p64).max() / np.abs(p64).max()
    print(f"tol {tol:g}: FP32 iterations {ctx[1][1]} vs FP64 {ctx[0][1]}, residual {rels:.3e} "
          f"(against the FP64 Hessian {rel32:.3e}, rounding floor {floor:.3e}), direction difference {err:.3e}")
    assert err <= 1e4 * max(tol, floor)


def test_fp32_hessian_blocks_are_rounded_fp64():
    """FP32 mode stores each elastic block as the FP32 rounding of an FP64 sum
    of chunk partials, each an FP32 sum of FP32 per-tet contributions: every
    entry within a few FP32 units of the FP64 block (relative to the
    block's largest entry), and the two modes share the pattern."""
    sc = S.two_cubes(4)
    x = _deformed(sc, 0.004, 4)
    vals = {}
    for prec in (0, 1):
        g = NewtonIPC(sc, precision=prec)
        g.set_state(x)
        g.compute_xn_and_xTilde(0.0)
        g.find_constraints()
        g.compute_hessian_and_gradient()
        import ctypes as C
        n = C.c_int()
        _lib.call("fipc_get_bsr", g.ctx, C.byref(n), None, None, None)
        v = np.zeros(n.value * 9)
        _lib.call("fipc_get_bsr", g.ctx, C.byref(n), None, None, _lib.dptr(v))
        vals[prec] = v
    v0, v1 = vals[0].reshape(-1, 9), vals[1].reshape(-1, 9)
    assert v0.shape == v1.shape
    scale = np.abs(v0).max(axis=1, keepdims=True)
    err = np.abs(v1 - v0) / scale
    print(f"FP32 block entries: max error {err.max():.3e} of the block's largest entry")
    assert err.max() <= 2.0 ** -18
    assert np.array_equal(v1, v1.astype(np.float32).astype(np.float64))


def _assembled(sc, x, prec):
    g = NewtonIPC(sc, precision=prec)
    g.set_state(x)
    g.compute_xn_and_xTilde(0.0)
    g.find_constraints()
    out = []
    for _ in range(2):
        g.compute_hessian_and_gradient()
        n = C.c_int()
        _lib.call("fipc_get_bsr", g.ctx, C.byref(n), None, None, None)
        v = np.zeros(n.value * 9)
        _lib.call("fipc_get_bsr", g.ctx, C.byref(n), None, None, _lib.dptr(v))
        out.append((g.vector(2).ravel().copy(), v.reshape(-1, 9)))
    return out


@pytest.mark.parametrize("shuffle", [False, True])
def test_fp32_tiled_assembly_matches_fp64(shuffle):
    """FP32 mode's tiled tet pass (a wavefront sums up to 64 consecutive tets'
    blocks and corner gradients in LDS and writes one partial per distinct
    block and vertex) against FP64 mode's per-tet records, on a deformed
    10x16x10 compressed slab: the gradient (FP64 in both modes; only the
    summation order differs) within 1e-12 of its largest entry, the blocks
    within 2^-18 of each block's largest entry, and a second assembly gives
    the same bits.  With the tets shuffled, consecutive tets share few
    vertices and the chunks close on the LDS limit of 256 distinct blocks
    instead of at 64 tets."""
    import dataclasses
    sc = S.compressed_slab(10, 16, 10)
    if shuffle:
        perm = np.random.default_rng(7).permutation(len(sc.tets))
        sc = dataclasses.replace(sc, tets=np.ascontiguousarray(sc.tets[perm]), density=sc.density[perm],
                                 youngs=sc.youngs[perm], poisson=sc.poisson[perm])
    x = _deformed(sc, 0.02, 5)
    (r64, v64), _ = _assembled(sc, x, 0)
    (r32, v32), (r32b, v32b) = _assembled(sc, x, 1)
    gerr = np.abs(r32 - r64).max() / np.abs(r64).max()
    scale = np.abs(v64).max(axis=1, keepdims=True)
    berr = (np.abs(v32 - v64) / scale).max()
    print(f"shuffle={shuffle}: gradient error {gerr:.3e} of its largest entry, block error {berr:.3e}")
    assert gerr <= 1e-12
    assert berr <= 2.0 ** -18
    assert np.array_equal(r32, r32b) and np.array_equal(v32, v32b)


@pytest.mark.parametrize("scene_fn,steps,amp", [(S.tiny_pair, 4, 0.001), (lambda: S.two_cubes(4), 12, 0.004)])
def test_fp32_steps_match_oracle(scene_fn, steps, amp):
    """Whole steps in FP32 mode from a pre-strained state, Newton tolerance
    tightened 1e3 times: positions within 1e-6 of the bbox diagonal of the
    FP64 oracle's, and the per-step energy error is printed."""
    sc = scene_fn()
    sc.tol_rel *= 1e-3
    g = NewtonIPC(sc, precision=1, pcg_tol=1e-12)
    o = OracleIPC(sc)
    x = _deformed(sc, amp, 2)
    o.x = np.ascontiguousarray(x.copy())
    g.set_state(x)
    for f in range(steps):
        sg = g.step()
        so = o.step()
        assert sg.converged
        assert g.check_collision() == 0
        err = np.abs(g.x - o.x).max()
        print(f"step {f}: newton {sg.newton_iters} vs {so['newton_iters']}, position error {err / sc.bbox_diag:.3e} bbox")
        assert err <= 1e-6 * sc.bbox_diag, (f, err)


@pytest.mark.parametrize("prec", [0, 1])
def test_solve_refine_continues_to_tight_tolerance(prec):
    """fipc_solve at a loose tolerance, then fipc_solve_refine to a tight one:
    the continued solution meets the tight tolerance in the FP64 residual
    (relative to the same |b|), agrees with a direct tight solve, and the
    continuation is refused once the system has been reassembled."""
    sc = S.two_cubes(5)
    x = _deformed(sc, 0.004, 5)
    loose, tight = 1e-2, 1e-9
    g = NewtonIPC(sc, precision=prec, pcg_tol=tight, pcg_tol_search=loose)
    ref = NewtonIPC(sc, precision=prec, pcg_tol=tight)
    for s in (g, ref):
        s.set_state(x)
        s.compute_xn_and_xTilde(0.0)
        s.find_constraints()
        s.compute_hessian_and_gradient()
    it0, rel0 = g.solve_system(tol=loose)
    it1, rel1 = g.refine_solve()
    itr, relr = ref.solve_system()
    b = ref.vector(2).ravel()
    D = np.repeat(sc.dirichlet_mask, 3)
    p = g.vector(1).ravel()
    pr = ref.vector(1).ravel()
    r = b - _spmv(ref, p)
    r[D] = 0.0
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"precision {prec}: loose {it0} it ({rel0:.2e}), continued +{it1} it ({rel1:.2e}, FP64 {rel:.2e}); "
          f"direct {itr} it ({relr:.2e})")
    assert rel0 <= loose * 1.0001 and it0 > 0 and it1 > 0
    assert rel <= tight * 1.0001
    assert np.abs(p - pr).max() <= 1e4 * tight * np.abs(pr).max()
    # a continuation needs the system the last solve solved
    g.compute_hessian_and_gradient()
    it_, rel_ = C.c_int(), C.c_double()
    assert g.so.fipc_solve_refine(g.ctx, tight, 100, C.byref(it_), C.byref(rel_)) == 6  # FIPC_ERR_STATE
    with pytest.raises(ValueError):
        NewtonIPC(sc, pcg_tol=1e-6, pcg_tol_search=1e-8)


@pytest.mark.parametrize("prec", [0, 1])
def test_inexact_search_directions_steps(prec):
    """Whole steps with search directions at 1e-2 and the deciding direction
    continued to 1e-12, against steps of the tight solver from the same state,
    Newton tolerance tightened 1e3 times: every step converges
    intersection-free and the positions agree within 1e-6 of the bbox
    diagonal."""
    sc = S.two_cubes(4)
    sc.tol_rel *= 1e-3
    g = NewtonIPC(sc, precision=prec, pcg_tol=1e-12, pcg_tol_search=1e-2)
    ref = NewtonIPC(sc, precision=prec, pcg_tol=1e-12)
    x = _deformed(sc, 0.004, 6)
    g.set_state(x)
    ref.set_state(x)
    for f in range(6):
        x0, v0 = g.x.copy(), g.v.copy()
        ref.set_state(x0, v0)
        st = g.step()
        sr = ref.step()
        assert st.converged and sr.converged and g.check_collision() == 0
        err = np.abs(g.x - ref.x).max()
        print(f"step {f}: newton {st.newton_iters} vs {sr.newton_iters}, pcg {sum(st.pcg_iters)} vs "
              f"{sum(sr.pcg_iters)}, position difference {err / sc.bbox_diag:.3e} bbox")
        assert err <= 1e-6 * sc.bbox_diag, (f, err)


def test_certified_newton_test_keeps_the_trajectory():
    """With inexact search directions in FP32 mode, the deciding direction's
    continuation may stop once max|p*| < tol is certified for the exact
    direction (max|p| + |b - A p|_2 / (m_min / 2), A >= M).  A certified
    direction is never applied (the step ends), so on a 10x16x10 compressed
    slab with moving plates, 12 steps with and without the certificate give
    the same Newton iterations and the same positions bit for bit, with at
    least one certified test and fewer PCG iterations."""
    sc = S.compressed_slab(10, 16, 10, steps=12)
    runs = {}
    for cert in (False, True):
        g = NewtonIPC(sc, precision=1, pcg_tol_search=2e-2, certify=cert)
        newton, pcg = [], 0
        for _ in range(12):
            st = g.step()
            assert st.converged
            newton.append(st.newton_iters)
            pcg += sum(st.pcg_iters)
        runs[cert] = (g.x.copy(), newton, pcg, g.certified_solves)
    (x0, n0, p0, _), (x1, n1, p1, c1) = runs[False], runs[True]
    print(f"newton {n1}, pcg {p0} -> {p1}, certified tests {c1}")
    assert n0 == n1
    assert np.array_equal(x0, x1)
    assert c1 >= 1 and p1 < p0


@pytest.mark.parametrize("prec", [0, 1])
def test_residual_is_max_abs_direction(prec):
    """output_residual (max |p|, NEWTON_IPC.py:805-812) equals the maximum of
    the direction read back, bit for bit: after a solve (measured by the PCG
    iterations), after a continuation, and after the direction is overwritten
    from the host (measured by its own reduction)."""
    sc = S.two_cubes(5)
    g = NewtonIPC(sc, precision=prec, pcg_tol=1e-9, pcg_tol_search=1e-2)
    g.set_state(_deformed(sc, 0.004, 7))
    g.compute_xn_and_xTilde(0.0)
    g.find_constraints()
    g.compute_hessian_and_gradient()
    g.solve_system(tol=1e-2)
    assert g.output_residual() == np.abs(g.vector(1)).max()
    g.refine_solve()
    assert g.output_residual() == np.abs(g.vector(1)).max()
    p = np.random.default_rng(8).normal(size=(g.nv, 3))
    g.set_vector(1, p)
    assert g.output_residual() == np.abs(p).max()


def test_energy_reuse_matches_full_evaluation():
    """After xTilde moves (compute_xn_and_xTilde) at an unchanged x and
    constraint set, compute_energy re-evaluates only the inertia term; the
    result equals a full evaluation (forced by rewriting the same x) bit for
    bit, and so do its parts."""
    sc = S.two_cubes(4)
    g = NewtonIPC(sc)
    x = _deformed(sc, 0.003, 9)
    g.set_state(x, np.full_like(x, 0.01))
    g.compute_xn_and_xTilde(0.0)
    g.find_constraints()
    e1 = g.compute_energy(parts=True)
    g.set_state(v=np.full_like(x, -0.02))  # new velocity only: x version unchanged
    g.compute_xn_and_xTilde(sc.dt)      # new xTilde, same x and constraints
    e2 = g.compute_energy(parts=True)
    g.set_state(g.x)                     # same x, new x version: full evaluation
    g.find_constraints()
    e3 = g.compute_energy(parts=True)
    assert e2[0] == e3[0] and np.array_equal(e2[1], e3[1])
    assert e1[1][1] == e3[1][1] and e1[1][2] == e3[1][2] and e1[0] != e3[0]


def test_fp32_per_step_error_against_fp64_on_cube_column():
    """The FP32-Hessian form on the 200-cube column (343k tets)
    against FP64, step by step from the same state (the FP64 run's): after
    each step the FP32 step's positions, its incremental potential and the
    norm of its gradient (both evaluated by the FP64 oracle at the end
    state) are compared with the FP64 step's.  Both steps stop anywhere
    within the Newton tolerance of the same minimiser, so positions differ by
    at most about twice the tolerance; the printed table is the per-step
    error the north star asks to state."""
    sc = S.config(2)
    g64 = NewtonIPC(sc)
    g32 = NewtonIPC(sc, precision=1)
    o = OracleIPC(sc)
    tol = sc.newton_tol
    rows = []
    for f in range(40):
        x0, v0 = g64.x, g64.v
        g32.set_state(x0, v0)
        g32.frame = g64.frame
        s64 = g64.step()
        s32 = g32.step()
        assert (s64.converged or s64.stalled) and (s32.converged or s32.stalled)
        assert g32.check_collision() == 0
        if f < 18 or f % 4:
            continue  # contact begins near step 15; the oracle evaluates every fourth step after it
        x64, x32 = g64.x, g32.x
        out = []
        for x in (x64, x32):
            o.x, o.v = np.ascontiguousarray(x0), v0.copy()
            o.compute_xn_and_xTilde(f * sc.dt)  # x_n and xTilde of this step
            o.x = np.ascontiguousarray(x)
            o.find_constraints()
            E = o.compute_energy()
            o.assemble()
            D = np.repeat(o.dmask, 3)
            gvec = o.rhs.copy()
            gvec[D] = 0.0
            out.append((E, np.linalg.norm(gvec)))
        dx = np.abs(x32 - x64).max()
        rows.append((f, s64.newton_iters, s32.newton_iters, dx / tol, abs(out[1][0] - out[0][0]) / abs(out[0][0]),
                     out[0][1], out[1][1], sum(o.constraints()[k].shape[0] for k in ("PT", "EE"))))
        assert dx <= 2.5 * tol, (f, dx, tol)
        assert abs(out[1][0] - out[0][0]) <= 1e-6 * abs(out[0][0])
    print("step  newton64 newton32  max|x32-x64|/tol  |E32-E64|/|E64|  |g64|  |g32|  PT+EE rows")
    for r in rows:
        print(f"{r[0]:4d} {r[1]:8d} {r[2]:8d} {r[3]:16.3e} {r[4]:16.3e} {r[5]:.3e} {r[6]:.3e} {r[7]}")
    assert rows and rows[-1][7] > 0


@pytest.mark.parametrize("prec", [0, 1])
def test_iteration_budget_edges(prec):
    """A solve whose budget ends on the iteration that meets the tolerance is
    converged (the test of that last residual runs after the loop); one
    iteration fewer, or a small budget, returns not-converged (status 5)
    with the residual above the tolerance."""
    sc = S.two_cubes(4)
    x = _deformed(sc, 0.004, 7)
    g = NewtonIPC(sc, precision=prec)
    g.set_state(x)
    g.compute_xn_and_xTilde(0.0)
    g.find_constraints()
    g.compute_hessian_and_gradient()
    n, rel = g.solve_system()
    assert g.last_solve_converged and rel <= g.pcg_tol * 1.0001 and n > 3
    it, r = C.c_int(), C.c_double()
    assert g.so.fipc_solve(g.ctx, g.pcg_tol, n, C.byref(it), C.byref(r)) == 0
    assert it.value == n and r.value <= g.pcg_tol * 1.0001
    assert g.so.fipc_solve(g.ctx, g.pcg_tol, 3, C.byref(it), C.byref(r)) == 5
    assert it.value == 3 and r.value > g.pcg_tol
    if prec == 0:
        assert g.so.fipc_solve(g.ctx, g.pcg_tol, n - 1, C.byref(it), C.byref(r)) == 5
        assert it.value == n - 1 and r.value > g.pcg_tol
    print(f"precision {prec}: {n} iterations to converge; budgets n, 3{', n - 1' if prec == 0 else ''} checked")
