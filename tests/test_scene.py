"""Scene generators, boundary extraction and set-up quantities (host logic)."""
import math

import numpy as np
import pytest

from fastipc_amd import scene as S
from fastipc_amd.model import SceneSetup, adaptive_kappa, barrier_H


def _signed_volumes(x, t):
    a, b, c = x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]], x[t[:, 3]] - x[t[:, 0]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)) / 6


@pytest.mark.parametrize("n", [2, 3, 5, 10])
def test_lattice_counts_and_orientation(n):
    x, t = S.lattice(n)
    assert len(x) == n ** 3 and len(t) == 5 * (n - 1) ** 3
    vol = _signed_volumes(x, t)
    assert np.all(vol > 0)
    assert vol.sum() == pytest.approx(((n - 1) * S.LATTICE_SPACING) ** 3)


def test_boundary_is_closed_and_outward():
    x, t = S.lattice(4, 5, 3)
    p, e, f = S.find_boundary(t)
    # every undirected surface edge is used by exactly two triangles, once each way
    from collections import Counter
    dir_edges = Counter()
    for a, b, c in f:
        for u, v in ((a, b), (b, c), (c, a)):
            dir_edges[(u, v)] += 1
    assert all(k == 1 for k in dir_edges.values())
    assert all(dir_edges[(v, u)] == 1 for (u, v) in dir_edges)
    assert len(e) == len(dir_edges) // 2
    assert np.all(e[:, 0] < e[:, 1])
    # outward orientation: divergence theorem gives the positive volume
    v = np.einsum("ij,ij->i", x[f[:, 0]], np.cross(x[f[:, 1]], x[f[:, 2]])).sum() / 6
    assert v == pytest.approx(3 * 4 * 2 * S.LATTICE_SPACING ** 3)
    # surface points of an n1 x n2 x n3 lattice
    assert len(p) == 4 * 5 * 3 - 2 * 3 * 1
    # deterministic order
    p2, e2, f2 = S.find_boundary(t[::-1].copy())
    assert np.array_equal(p, p2) and np.array_equal(e, e2)


def test_config_sizes_match_the_workloads():
    sc = S.config(0)
    assert sc.n_verts == 2000 + 8 and sc.n_tets == 7290 + 5
    p, e, f = sc.boundary
    assert (len(p), len(e), len(f)) == (976 + 8, 2916 + 18, 1944 + 12)
    sc = S.compressed_slab(50, 80, 50, steps=1)
    assert sc.n_verts == 200000 + 16 and sc.n_tets == 948395 + 10
    assert sc.dt == 0.005


def test_slab_plates_move_and_twist():
    sc = S.compressed_slab(4, 6, 4)
    x0 = sc.dirichlet(0.0)
    x1 = sc.dirichlet(0.25)
    fixed = sc.dirichlet_mask
    assert np.allclose(x0[fixed], sc.x[fixed])
    assert not np.allclose(x1[fixed], sc.x[fixed])
    assert np.array_equal(x1[~fixed], sc.x[~fixed])
    # plates close: bottom up, top down
    H = 5 * S.LATTICE_SPACING
    bot = fixed & (sc.x[:, 1] < 0)
    top = fixed & (sc.x[:, 1] > H)
    assert np.all(x1[bot, 1] > x0[bot, 1]) and np.all(x1[top, 1] < x0[top, 1])


def test_setup_quantities():
    sc = S.two_cubes(4)
    s = SceneSetup(sc)
    # total mass = density x total volume
    assert s.mass.sum() == pytest.approx(np.sum(sc.density * s.vol))
    # adaptive kappa (NEWTON_IPC.py:116-124, x10 at :943)
    bH = barrier_H(1e-16, s.dHat2, 1.0)
    assert s.kappa == pytest.approx(10 * 1e13 * s.mass.mean() / (4e-16 * bH))
    assert adaptive_kappa(s.mass, s.dHat2) == pytest.approx(s.kappa)
    assert s.dHat == pytest.approx(1e-3 * sc.bbox_diag)
    e = s.be
    assert 1 / s.inv_dx == pytest.approx(np.mean(np.linalg.norm(sc.x[e[:, 0]] - sc.x[e[:, 1]], axis=1)))
    la, mu = sc.lame()
    E, nu = 1e5, 0.4
    assert la[0] == pytest.approx(E * nu / ((1 + nu) * (1 - 2 * nu))) and mu[0] == pytest.approx(E / (2 * (1 + nu)))


def test_seeded_scenes_are_reproducible():
    a = S.cube_grid(4, 2, seed=3)
    b = S.cube_grid(4, 2, seed=3)
    c = S.cube_grid(4, 2, seed=4)
    assert np.array_equal(a.x, b.x) and not np.array_equal(a.x, c.x)


@pytest.mark.parametrize("cfg", [0, 1, 2, 4])
def test_config_initial_states_are_intersection_free(cfg):
    """IPC needs an intersection-free start: the oracle's collision check finds
    no edge crossing a triangle and no pair is within dHat (in particular no
    pair of fixed vertices: static boxes do not touch)."""
    from oracle.newton import OracleIPC
    sc = S.config(cfg)
    o = OracleIPC(sc)
    assert o.check_collision() == 0
    assert sum(o.find_constraints()) == 0


@pytest.mark.parametrize("make", [lambda: S.compressed_slab(6, 8, 6), lambda: S.two_cubes(3), S.tiny_pair])
def test_dirichlet_rows_match_full_targets(make):
    """The per-row prescribed positions equal the Dirichlet rows of the full
    target array, at several times."""
    sc = make()
    mask = sc.dirichlet_mask
    for t in (0.0, 0.013, 0.25):
        rows, xyz = sc.dirichlet_rows(t)
        assert rows.dtype == np.int32 and np.array_equal(rows, np.flatnonzero(mask))
        assert np.array_equal(xyz, sc.dirichlet(t)[mask])
