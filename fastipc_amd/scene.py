"""Seeded scene generators for the IPC Newton step.

The reference builds scenes in ``projects/FastIPC/reader.py`` (``add_object`` at
reader.py:30-85, ``find_boundary`` from ``common/math/graph_tools.py:4-46`` and
a Dirichlet box at reader.py:278-290).  This module provides the same pieces:

* a regular lattice with five tetrahedra per cell (alternating parity, so the
  face diagonals of neighbouring cells agree), positively oriented like the
  reference's six-tet lattice (``det[x1-x0, x2-x0, x3-x0] > 0``);
* a one-cell box used for Dirichlet grounds, walls and plates;
* ``find_boundary`` with the reference's semantics but a deterministic order
  (triangles, edges and points sorted lexicographically);
* the five benchmark scenes (two stacked cubes, 27 rotated cubes in a box,
  a 200-cube column, the compressed slab between plates) plus tiny ones for
  tests.

All arrays are numpy: positions float64 ``(V,3)``, tets int32 ``(T,4)``.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, Optional

import numpy as np

LATTICE_SPACING = 0.1  # reader.py:43-45 uses 0.1 between lattice vertices


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def _tet_orient(x: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """Swap the last two vertices of every tet whose signed volume is negative."""
    a = x[tets[:, 1]] - x[tets[:, 0]]
    b = x[tets[:, 2]] - x[tets[:, 0]]
    c = x[tets[:, 3]] - x[tets[:, 0]]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    out = tets.copy()
    neg = det < 0
    out[neg, 2], out[neg, 3] = tets[neg, 3], tets[neg, 2]
    return out


def lattice(nx: int, ny: Optional[int] = None, nz: Optional[int] = None,
            h: float = LATTICE_SPACING):
    """Vertex lattice of ``nx*ny*nz`` points with five tets per cell.

    Vertex ``(i, j, k)`` has index ``i*ny*nz + j*nz + k`` and position
    ``(i*h, j*h, k*h)``, the indexing of reader.py:37-45.  Cells with even
    ``i+j+k`` place the central tet on their even local corners, odd cells on
    their odd local corners; either way every face diagonal joins two lattice
    points of even global parity, so the mesh is conforming.
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x = np.stack([I.ravel() * h, J.ravel() * h, K.ravel() * h], axis=1).astype(np.float64)

    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()

    def corner(dx, dy, dz):
        return (ci + dx) * ny * nz + (cj + dy) * nz + (ck + dz)

    c = {}
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c[(dx, dy, dz)] = corner(dx, dy, dz)
    even = [(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    odd = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]

    def split(center, tips):
        tets = [np.stack([c[v] for v in center], axis=1)]
        for t in tips:
            nb = [v for v in center if sum(abs(a - b) for a, b in zip(v, t)) == 1]
            tets.append(np.stack([c[t]] + [c[v] for v in nb], axis=1))
        return np.stack(tets, axis=1)  # (ncell, 5, 4)

    parity = (ci + cj + ck) % 2
    t_even = split(even, odd)
    t_odd = split(odd, even)
    tets = np.where(parity[:, None, None] == 0, t_even, t_odd).reshape(-1, 4).astype(np.int32)
    tets = _tet_orient(x, tets)
    return x, tets


def box(lo, hi):
    """Eight-vertex, five-tet box spanning ``lo``..``hi`` (a Dirichlet obstacle)."""
    x, tets = lattice(2, 2, 2, h=1.0)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    return lo + x * (hi - lo), tets


def rotation_matrix(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)


# ---------------------------------------------------------------------------
# boundary extraction
# ---------------------------------------------------------------------------
def find_boundary(tets: np.ndarray):
    """Surface points, edges and triangles of a tet mesh.

    Semantics of ``common/math/graph_tools.py:26-45``: every tet contributes
    faces ``(p0,p2,p1), (p0,p3,p2), (p0,p1,p3), (p1,p2,p3)``; a face is on the
    boundary when no face with the opposite orientation exists.  An edge
    ``(a,b)`` of a boundary triangle is emitted only where ``a < b`` inside that
    triangle, so each undirected edge appears once with sorted endpoints.
    The reference iterates a Python set; here every output is sorted
    lexicographically so that runs are reproducible.
    """
    tets = np.asarray(tets, np.int64)
    p0, p1, p2, p3 = tets.T
    faces = np.concatenate([
        np.stack([p0, p2, p1], 1), np.stack([p0, p3, p2], 1),
        np.stack([p0, p1, p3], 1), np.stack([p1, p2, p3], 1)], axis=0)
    # rotate each face so its smallest vertex is first (keeps orientation)
    r = np.argmin(faces, axis=1)
    idx = (r[:, None] + np.arange(3)[None, :]) % 3
    canon = np.take_along_axis(faces, idx, axis=1)
    rev = canon[:, [0, 2, 1]]
    nb = int(max(1, int(faces.max()) + 1)).bit_length()
    assert 3 * nb <= 63, "vertex count too large for boundary keys"

    def key(f):
        return (f[:, 0] << (2 * nb)) | (f[:, 1] << nb) | f[:, 2]

    fk = key(canon)
    rk = key(rev)
    fk_sorted = np.unique(fk)
    has_rev = np.isin(rk, fk_sorted, assume_unique=False)
    bfaces = faces[~has_rev]
    # the same oriented face can be produced twice only in a non-manifold mesh
    _, first = np.unique(key(np.take_along_axis(bfaces, ((np.argmin(bfaces, 1)[:, None] + np.arange(3)[None, :]) % 3), 1)), return_index=True)
    bfaces = bfaces[np.sort(first)]
    order = np.lexsort((bfaces[:, 2], bfaces[:, 1], bfaces[:, 0]))
    tris = bfaces[order]
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]], axis=0)
    e = e[e[:, 0] < e[:, 1]]
    e = np.unique(e, axis=0)
    pts = np.unique(tris.ravel())
    return pts.astype(np.int32), e.astype(np.int32), tris.astype(np.int32)


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Scene:
    name: str
    x: np.ndarray                 # (V,3) float64 rest/initial positions
    tets: np.ndarray              # (T,4) int32
    density: np.ndarray           # (T,) float64
    youngs: np.ndarray            # (T,) float64
    poisson: np.ndarray           # (T,) float64
    dirichlet_mask: np.ndarray    # (V,) bool
    dt: float
    gravity: float = -9.8         # applied on axis 1 (NEWTON_IPC.py:502)
    steps: int = 1
    dirichlet_fn: Optional[Callable[[float], np.ndarray]] = None
    # optional: the prescribed positions of the Dirichlet rows only, in the
    # order of np.flatnonzero(dirichlet_mask), shape (n_fixed, 3)
    dirichlet_rows_fn: Optional[Callable[[float], np.ndarray]] = None
    boundary: tuple = None        # (points, edges, triangles)
    dhat_rel: float = 1e-3        # dhat = dhat_rel * bbox diagonal
    tol_rel: float = 1e-2         # Newton stop: max|p| < tol_rel * dt * tol_len
    # length of the Newton tolerance: the bbox diagonal when None; 1.0 gives
    # the reference's absolute test max|p| < 1e-2 dt (NEWTON_IPC.py:975)
    tol_len: Optional[float] = None
    eta: float = 0.2              # CCD gap fraction (NEWTON_IPC.py:405, :439)
    seed: int = 0

    def __post_init__(self):
        if self.boundary is None:
            self.boundary = find_boundary(self.tets)

    @property
    def n_verts(self):
        return len(self.x)

    @property
    def n_tets(self):
        return len(self.tets)

    @property
    def bbox_diag(self):
        return float(np.linalg.norm(self.x.max(0) - self.x.min(0)))

    @property
    def dhat(self):
        return self.dhat_rel * self.bbox_diag

    @property
    def newton_tol(self):
        return self.tol_rel * self.dt * (self.bbox_diag if self.tol_len is None else self.tol_len)

    def dirichlet(self, t: float) -> np.ndarray:
        """Prescribed positions (V,3) at time ``t``; only masked rows matter."""
        if self.dirichlet_fn is None:
            return self.x.copy()
        return self.dirichlet_fn(t)

    def dirichlet_rows(self, t: float):
        """(rows, positions): the Dirichlet vertices (int32, ascending) and
        their prescribed positions at time ``t``, shape (len(rows), 3)."""
        mask = self.dirichlet_mask
        key = (id(mask), int(np.count_nonzero(mask)))
        if getattr(self, "_drows_key", None) != key:
            self._drows = np.flatnonzero(mask).astype(np.int32)
            self._drows_key = key
        rows = self._drows
        if self.dirichlet_rows_fn is not None:
            return rows, np.ascontiguousarray(self.dirichlet_rows_fn(t), np.float64)
        if self.dirichlet_fn is None:
            return rows, np.ascontiguousarray(self.x[rows], np.float64)
        return rows, np.ascontiguousarray(self.dirichlet_fn(t)[rows], np.float64)

    def lame(self):
        E, nu = self.youngs, self.poisson
        la = E * nu / ((1 + nu) * (1 - 2 * nu))
        mu = E / (2 * (1 + nu))
        return la, mu


class _Builder:
    def __init__(self):
        self.xs, self.ts, self.rho, self.E, self.nu, self.fix = [], [], [], [], [], []
        self.nv = 0

    def add(self, x, t, rho=1000.0, E=1e5, nu=0.4, fixed=False):
        self.xs.append(np.asarray(x, np.float64))
        self.ts.append(np.asarray(t, np.int64) + self.nv)
        nt = len(t)
        self.rho.append(np.full(nt, rho))
        self.E.append(np.full(nt, E))
        self.nu.append(np.full(nt, nu))
        self.fix.append(np.full(len(x), fixed, bool))
        first = self.nv
        self.nv += len(x)
        return first

    def build(self, **kw):
        return Scene(x=np.concatenate(self.xs), tets=np.concatenate(self.ts).astype(np.int32),
                     density=np.concatenate(self.rho), youngs=np.concatenate(self.E),
                     poisson=np.concatenate(self.nu), dirichlet_mask=np.concatenate(self.fix), **kw)


def two_cubes(n: int = 10, steps: int = 50) -> Scene:
    """Two stacked ``n^3``-vertex cubes dropped onto a Dirichlet ground box.

    Each cube's gap (to the ground, and between the cubes) is 0.05 of the cube
    edge; the upper cube is shifted sideways so no two edges are exactly
    aligned.  Material E=1e5, nu=0.4, rho=1000; dt=0.01.
    """
    b = _Builder()
    xc, tc = lattice(n)
    edge = (n - 1) * LATTICE_SPACING
    gap = 0.05 * edge
    b.add(xc + [0.0, gap, 0.0], tc)
    b.add(xc + [0.31 * edge, 2 * gap + edge, 0.17 * edge], tc)
    gx, gt = box([-edge, -0.1 * edge, -edge], [2.5 * edge, 0.0, 2.5 * edge])
    b.add(gx, gt, fixed=True)
    return b.build(name=f"two_cubes_{n}", dt=0.01, steps=steps)


def cube_grid(n: int = 16, grid: int = 3, seed: int = 0, steps: int = 100) -> Scene:
    """``grid^3`` cubes of ``n^3`` vertices with seeded random rotations and jitter,
    falling into a fixed box (ground plus four walls, all Dirichlet)."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    xc, tc = lattice(n)
    edge = (n - 1) * LATTICE_SPACING
    center = xc.mean(0)
    spacing = edge * (math.sqrt(3.0) + 0.3)
    for gi in range(grid):
        for gj in range(grid):
            for gk in range(grid):
                axis = rng.normal(size=3)
                ang = rng.uniform(0, 2 * math.pi)
                R = rotation_matrix(axis, ang)
                jit = rng.uniform(-0.05, 0.05, size=3) * edge
                pos = np.array([gi, gj, gk], np.float64) * spacing + spacing / 2 + jit
                pos[1] += 0.1 * edge
                b.add((xc - center) @ R.T + pos, tc)
    L = grid * spacing
    th = 0.1 * edge
    H = L + spacing
    # ground and four walls as separate boxes, kept apart by gaps well above
    # dHat: IPC needs an intersection-free start, and touching static boxes
    # would put zero-distance pairs between fixed vertices
    gap = 0.05 * edge
    for lo, hi in [([-th, -th, -th], [L + th, 0.0, L + th]),
                   ([-th, gap, -th], [0.0, H, L + th]), ([L, gap, -th], [L + th, H, L + th]),
                   ([gap, gap, -th], [L - gap, H, 0.0]), ([gap, gap, L], [L - gap, H, L + th])]:
        gx, gt = box(lo, hi)
        b.add(gx, gt, fixed=True)
    # the reference's absolute Newton tolerance 1e-2 dt: scaled by this
    # scene's 18 m diagonal it would exceed a step's free-fall increment
    # dt^2 g = 9.8e-4, the test would pass before the first direction is
    # applied (NEWTON_IPC.py:975 breaks before the line search) and the
    # cubes would never move
    return b.build(name=f"cube_grid_{grid}x{n}", dt=0.01, steps=steps, seed=seed, tol_len=1.0)


def cube_column(n: int = 8, count: int = 200, seed: int = 0, width: float = 5.6,
                steps: int = 20) -> Scene:
    """``count`` axis-aligned cubes of ``n^3`` vertices at seeded uniform positions
    in a ``width x 4*width x width`` column above a Dirichlet ground; positions
    are rejection-sampled so that every pair of cubes is at least 0.05 apart."""
    rng = np.random.default_rng(seed)
    xc, tc = lattice(n)
    edge = (n - 1) * LATTICE_SPACING
    gapmin = 0.05
    lo_all = []
    tries = 0
    while len(lo_all) < count:
        tries += 1
        if tries > 200000:
            raise RuntimeError("cube_column: cannot place cubes")
        p = rng.uniform([0, 0.05, 0], [width - edge, 4 * width - edge, width - edge])
        ok = True
        for q in lo_all:
            if np.all(np.abs(p - q) < edge + gapmin):
                ok = False
                break
        if ok:
            lo_all.append(p)
    b = _Builder()
    for p in lo_all:
        b.add(xc + p, tc)
    gx, gt = box([-0.5, -0.2, -0.5], [width + 0.5, 0.0, width + 0.5])
    b.add(gx, gt, fixed=True)
    # the reference's absolute Newton tolerance, as for cube_grid (this
    # scene's diagonal is 24 m)
    return b.build(name=f"cube_column_{count}x{n}", dt=0.01, steps=steps, seed=seed, tol_len=1.0)


def compressed_slab(nx: int = 50, ny: int = 80, nz: int = 50, steps: int = 200,
                    close_rate: float = 0.2, twist_deg: float = 30.0,
                    omega: float = 2 * math.pi) -> Scene:
    """A ``nx x ny x nz``-vertex slab (E=1e6, nu=0.45) between two rigid plates.

    The plates are Dirichlet boxes that close on the slab at ``close_rate``
    slab heights per second (each plate moves half of it) while twisting about
    the vertical axis by ``twist_deg * sin(omega t)`` in opposite senses.
    dt = 0.005.
    """
    b = _Builder()
    xs, ts = lattice(nx, ny, nz)
    Hh = (ny - 1) * LATTICE_SPACING
    W = (nx - 1) * LATTICE_SPACING
    D = (nz - 1) * LATTICE_SPACING
    b.add(xs, ts, E=1e6, nu=0.45)
    gap = 0.004 * math.sqrt(W * W + D * D + Hh * Hh) * 1.0
    pw = 0.35 * max(W, D)
    cx, cz = W / 2, D / 2
    bot_x, bot_t = box([cx - W / 2 - pw, -gap - 0.2, cz - D / 2 - pw], [cx + W / 2 + pw, -gap, cz + D / 2 + pw])
    top_x, top_t = box([cx - W / 2 - pw, Hh + gap, cz - D / 2 - pw], [cx + W / 2 + pw, Hh + gap + 0.2, cz + D / 2 + pw])
    ib = b.add(bot_x, bot_t, E=1e6, nu=0.45, fixed=True)
    it = b.add(top_x, top_t, E=1e6, nu=0.45, fixed=True)
    scene = b.build(name=f"slab_{nx}x{ny}x{nz}", dt=0.005, steps=steps)
    x0 = scene.x.copy()
    speed = 0.5 * close_rate * Hh
    amp = math.radians(twist_deg)

    def plates(t):
        """the 16 plate vertices (bottom then top) at time t"""
        th = amp * math.sin(omega * t)
        out = []
        for first, sgn in ((ib, 1.0), (it, -1.0)):
            R = rotation_matrix([0, 1, 0], sgn * th)
            p = x0[first:first + 8] - [cx, 0.0, cz]
            q = p @ R.T + [cx, 0.0, cz]
            q[:, 1] += sgn * speed * t
            out.append(q)
        return np.concatenate(out)

    def fn(t):
        out = x0.copy()
        q = plates(t)
        out[ib:ib + 8] = q[:8]
        out[it:it + 8] = q[8:]
        return out

    scene.dirichlet_fn = fn
    # the fixed vertices are exactly the plates' (bottom block first)
    assert np.array_equal(np.flatnonzero(scene.dirichlet_mask), np.r_[ib:ib + 8, it:it + 8])
    scene.dirichlet_rows_fn = plates
    return scene


def tiny_pair(seed: int = 0) -> Scene:
    """Two small cubes (3^3 vertices) almost touching above a ground box; a quick
    scene for parity tests."""
    b = _Builder()
    xc, tc = lattice(3)
    rng = np.random.default_rng(seed)
    edge = 2 * LATTICE_SPACING
    b.add(xc + [0.0, 0.002, 0.0], tc)
    R = rotation_matrix(rng.normal(size=3), rng.uniform(0, 2 * math.pi))
    c = xc.mean(0)
    b.add((xc - c) @ R.T + c + [0.05, edge * 1.9 + 0.004, 0.03], tc)
    gx, gt = box([-0.3, -0.05, -0.3], [0.5, 0.0, 0.5])
    b.add(gx, gt, fixed=True)
    return b.build(name="tiny_pair", dt=0.01, steps=5, seed=seed)


CONFIGS = {
    0: lambda: two_cubes(10, steps=50),
    1: lambda: cube_grid(16, 3, seed=0, steps=100),
    2: lambda: cube_column(8, 200, seed=0),
    4: lambda: compressed_slab(50, 80, 50, steps=200),
}


def config(i: int) -> Scene:
    return CONFIGS[i]()
