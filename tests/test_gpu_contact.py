"""Parity in the contact regime, at the benchmark sizes and with moving
Dirichlet plates.

* Whole steps of a small compressed slab whose plates close on it, GPU
  against the oracle: the prescribed plate increments enter the solve as
  Dirichlet increments (the PCG start x0 and r = b - A x0) and the CCD, the
  path the benchmark scene runs (NEWTON_IPC.py:505-511, :955-1002).
* The benchmark scenes run on the GPU into contact; the oracle then
  evaluates the GPU's state: constraint sets bit-exact (np.unique rows),
  incremental potential and gradient within 1e-10 relative, no intersection
  by the oracle's own test, and contact counts asserted non-zero.  The oracle
  cannot run whole steps at these sizes (its direct solve takes minutes)."""
import time

import numpy as np
import pytest

from fastipc_amd import scene as S
from fastipc_amd.newton import NewtonIPC
from oracle.newton import OracleIPC

pytestmark = pytest.mark.gpu

NAMES = ("PP", "PE", "PT", "EE", "EEM", "PPM", "PEM")


def _state_parity(g, o, sc, gradient=True):
    """The oracle at the GPU's state (x, v, frame): sets, energy, gradient,
    intersection-free.  Returns the constraint rows."""
    o.x = np.ascontiguousarray(g.x)
    o.v = g.v
    o.frame = g.frame
    o.compute_xn_and_xTilde(g.frame * sc.dt)
    g.compute_xn_and_xTilde(g.frame * sc.dt)
    g.find_constraints()
    o.find_constraints()
    cg, co = g.constraints(), o.constraints()
    for k in co:
        assert np.array_equal(cg[k], co[k]), (k, len(cg[k]), len(co[k]))
    Eg, pg = g.compute_energy(parts=True)
    Eo, po = o.compute_energy(parts=True)
    assert abs(Eg - Eo) <= 1e-10 * abs(Eo), (Eg, Eo)
    if gradient:
        g.compute_hessian_and_gradient()
        o.assemble()
        D = np.repeat(o.dmask, 3)
        ro = o.rhs.copy()
        ro[D] = 0.0  # rhs[D] = 0 (NEWTON_IPC.py:772)
        err = np.abs(g.vector(2).ravel() - ro).max() / np.abs(ro).max()
        assert err <= 1e-10, err
        o.coo = None
    assert o.check_collision() == 0
    return co


def _deformed(sc, amp, seed):
    rng = np.random.default_rng(seed)
    x = sc.x.copy()
    k = rng.normal(size=(3, 3)) * 3.0
    ph = rng.uniform(0, 2 * np.pi, size=3)
    free = ~sc.dirichlet_mask
    x[free] += amp * np.sin(x[free] @ k + ph)
    return x


@pytest.mark.parametrize("precision", [0, 1])
def test_moving_plates_steps_match_oracle(precision):
    """A 6x10x6 slab between plates closing at 2 slab heights per second with
    the 30 degree twist: both runs start from the same pre-strained state
    (every tet away from F = I, where the reference's 2x2 make_pd jumps) and
    take whole steps with the Newton tolerance tightened 1e3 times; positions
    agree within 1e-6 of the bbox diagonal after every step, and the plates
    are in contact (PT rows with plate vertices) for most of them."""
    sc = S.compressed_slab(6, 10, 6, close_rate=2.0)
    sc.tol_rel *= 1e-3
    plates = np.flatnonzero(sc.dirichlet_mask)
    g = NewtonIPC(sc, precision=precision, pcg_tol=1e-12)
    o = OracleIPC(sc)
    x = _deformed(sc, 0.002, 3)
    o.x = np.ascontiguousarray(x.copy())
    assert o.check_collision() == 0
    g.set_state(x)
    touching = 0
    for f in range(10):
        sg = g.step()
        so = o.step()
        assert sg.converged
        assert g.check_collision() == 0
        err = np.abs(g.x - o.x).max()
        rows = o.constraints()
        plate_rows = sum(int(np.isin(r, plates).any(axis=1).sum()) for r in rows.values() if len(r))
        touching += plate_rows > 0
        print(f"step {f}: newton {sg.newton_iters} vs {so['newton_iters']}, plate rows {plate_rows}, "
              f"position error {err / sc.bbox_diag:.3e} bbox")
        assert err <= 1e-6 * sc.bbox_diag, (f, err)
        # the plates reached the positions prescribed for this step to
        # within the Newton tolerance (the increment is part of the direction)
        assert np.abs(g.x[plates] - sc.dirichlet((g.frame - 1) * sc.dt)[plates]).max() <= sc.newton_tol
    assert touching >= 6, touching


@pytest.mark.parametrize("variant", ["plain", "chosen+inexact"])
def test_compressed_slab_full_size_into_compression(variant):
    """The benchmark slab (50x80x50 vertices, 948k tets) with its moving
    plates, 120 steps: both plates press on the slab from about step 80.
    Every step is intersection-free; at the end the oracle evaluates the
    GPU's state.  Slab-slab rows are counted and printed (the 20%
    compression of this scene does not fold the slab onto itself)."""
    import os
    from fastipc_amd import warmstart as W
    sc = S.compressed_slab(50, 80, 50)
    hook = None
    if variant == "plain":
        g = NewtonIPC(sc)
    else:
        g = NewtonIPC(sc, precision=1, pcg_tol_search=2e-2)
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        net = W.WarmStartMLP.load(os.path.join(root, "fastipc_amd", "weights", "ws_slab.npz"))
        hook = W.LearnedWarmStart(net, W.scale_for(sc))
    t0 = time.perf_counter()
    for _ in range(120):
        st = g.step(warm_start=hook)
        assert st.converged
        assert g.check_collision() == 0
    t1 = time.perf_counter()
    o = OracleIPC(sc)
    g.search_stats(True)
    co = _state_parity(g, o, sc)
    n = {k: len(v) for k, v in co.items()}
    nslab = sc.n_verts - 16
    self_rows = {k: int(np.all(v < nslab, axis=1).sum()) for k, v in co.items() if len(v)}
    print(f"{variant}: 120 steps in {t1 - t0:.2f} s, contacts {n}, slab-slab rows {self_rows}, "
          f"candidates {g.search_counts()}")
    # both plates: each plate face carries a point-triangle row per slab face vertex
    assert n["PT"] >= 2 * 50 * 50 and n["EE"] > 0


@pytest.mark.parametrize("cfg,steps", [(0, 50), (1, 100), (2, 150)])
def test_benchmark_scenes_into_contact(cfg, steps):
    """The stacked cubes, the cube grid and the cube column at their sizes and
    lengths (the column: 150 steps), run into
    contact: after the GPU steps the oracle finds the same constraint sets,
    energy and gradient, and no intersection; rows between two different
    bodies (not the static ground or walls) are asserted for the cube
    scenes.  Steps of the cube grid and the column may end stalled: an edge pair's
    closest-feature switch adds a barrier term that the np.unique merge of
    point-edge rows had counted once, so the energy jumps along the
    direction and the reference's line search would halve alpha without end
    (the oracle with its exact direct solve stalls the same way on a smaller
    cube grid, test_oracle_pins.test_reference_line_search_stall); such a
    step ends at its last iterate, which is still intersection-free."""
    sc = S.config(cfg)
    g = NewtonIPC(sc)
    t0 = time.perf_counter()
    stalled = 0
    for f in range(steps):
        st = g.step()
        assert st.converged or st.stalled, f
        stalled += st.stalled
        assert g.check_collision() == 0, f
    t1 = time.perf_counter()
    o = OracleIPC(sc)
    g.search_stats(True)
    co = _state_parity(g, o, sc)
    n = {k: len(v) for k, v in co.items()}
    # body id of every vertex: the scenes add bodies one after another
    body = np.searchsorted(np.cumsum(_body_sizes(sc)), np.arange(sc.n_verts), side="right")
    fixed = sc.dirichlet_mask
    between = 0
    for v in co.values():
        if len(v):
            b = body[v]
            between += int(((b != b[:, :1]).any(axis=1) & ~fixed[v].any(axis=1)).sum())
    cand = g.search_counts()
    print(f"config {cfg}: {sc.n_verts} verts, {sc.n_tets} tets, {steps} GPU steps in {t1 - t0:.2f} s "
          f"({stalled} stalled), contacts {n}, body-body rows {between}, broad-phase candidates {cand}")
    assert sum(n.values()) > 0
    if cfg in (1, 2):
        assert between > 0


def _body_sizes(sc):
    """Vertex counts of the bodies in the order the scene generator added
    them (connected components of the tet mesh, which are contiguous)."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    t = sc.tets
    n = sc.n_verts
    rows = np.repeat(t, 4, axis=1).ravel()
    cols = np.tile(t, (1, 4)).ravel()
    A = sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    _, lab = cg.connected_components(A, directed=False)
    # labels of contiguous bodies increase with the first vertex
    change = np.flatnonzero(np.diff(lab) != 0) + 1
    edges = np.r_[0, change, n]
    return np.diff(edges)


@pytest.mark.parametrize("cfg,steps,precision", [(1, 50, 0), (2, 60, 1)])
def test_contact_steps_are_bit_reproducible(cfg, steps, precision):
    """Two runs of the same scene into contact give the same bits: the
    contact gradient, diagonal blocks and products sum each vertex's rows
    in a fixed order (no atomics), so the Newton iterates, PCG iteration
    counts and step sizes repeat exactly."""
    sc = S.config(cfg)
    runs = []
    for _ in range(2):
        g = NewtonIPC(sc, precision=precision)
        hist = []
        for f in range(steps):
            st = g.step()
            hist.append((st.newton_iters, tuple(st.pcg_iters), tuple(st.alphas)))
        runs.append((np.array(g.x), hist, {k: len(v) for k, v in g.constraints().items()}))
        del g
    (xa, ha, ca), (xb, hb, cb) = runs
    print(f"config {cfg} precision {precision}: contacts {ca}")
    assert sum(ca.values()) > 0
    assert ha == hb
    assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
