"""CPU oracle of the Newton IPC time step (test infrastructure and CPU baseline).

Restates the frame loop of projects/FastIPC/NEWTON_IPC.py:955-1002 on the C
oracle (``libipc_oracle.so``, OpenMP) with the reference's linear-algebra
structure: COO triplets -> CSR (duplicates summed), Dirichlet rows/columns
zeroed with an identity on fixed dofs (:750-772) and a fresh direct
factorisation per Newton iteration (:773-775).  CHOLMOD (scikit-sparse) is
used when importable, else SuperLU (``scipy.sparse.linalg.splu``).

Moving Dirichlet vertices (whose prescribed position changes between steps)
are given the increment ``xTilde - x`` in the solve, the same extension the
product uses; for static Dirichlet vertices this is the reference exactly.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import scipy.sparse
import scipy.sparse.linalg

from . import lib as L

try:  # the reference's solver, when present
    from sksparse.cholmod import cholesky as _cholesky  # type: ignore
    SOLVER = "cholmod"
except Exception:  # pragma: no cover - depends on the box
    _cholesky = None
    SOLVER = "superlu"

NAMES = ("PP", "PE", "PT", "EE", "EEM", "PPM", "PEM")
WIDTH = (2, 3, 4, 4, 4, 4, 4)


def _barrier_H(d, dHat2, kappa):
    t2 = d - dHat2
    return kappa * ((math.log(d / dHat2) * -2.0 - t2 * 4.0 / d) + 1.0 / (d * d) * (t2 * t2)) if d < dHat2 else 0.0


class OracleIPC:
    def __init__(self, scene, max_newton=200, max_halvings=60, inversion=True):
        lib = L.lib()
        self.lib = lib
        self.scene = scene
        self.x = np.ascontiguousarray(scene.x, np.float64).copy()
        self.x0 = self.x.copy()                      # save_x0 (NEWTON_IPC.py:781-783)
        self.nv = len(self.x)
        self.tets = np.ascontiguousarray(scene.tets, np.int32)
        self.nt = len(self.tets)
        self.restT = np.zeros((self.nt, 9))
        self.vol = np.zeros(self.nt)
        lib.orc_tet_restT(self.nt, L.dptr(self.x), L.iptr(self.tets), L.dptr(self.restT), L.dptr(self.vol))
        # compute_restT_and_m (NEWTON_IPC.py:485-494)
        self.m = np.zeros(self.nv)
        for a in range(4):
            np.add.at(self.m, self.tets[:, a], self.vol * scene.density / 4.0)
        E, nu = np.asarray(scene.youngs, float), np.asarray(scene.poisson, float)
        self.la = np.ascontiguousarray(E * nu / ((1 + nu) * (1 - 2 * nu)))
        self.mu = np.ascontiguousarray(E / (2 * (1 + nu)))
        self.dt = scene.dt
        self.dHat = scene.dhat
        self.dHat2 = self.dHat * self.dHat
        # compute_adaptive_kappa (:116-124) times 10 (:943)
        self.kappa = 10.0 * 1.0e13 * self.m.sum() / self.nv / (4.0e-16 * _barrier_H(1.0e-16, self.dHat2, 1.0))
        bp, be, bt = scene.boundary
        self.bp, self.be, self.bt = L.i32(bp), L.i32(be), L.i32(bt)
        # compute_mean_of_boundary_edges (:127-134)
        self.inv_dx = 1.0 / float(np.mean(np.linalg.norm(self.x[self.be[:, 0]] - self.x[self.be[:, 1]], axis=1)))
        self.dmask = np.asarray(scene.dirichlet_mask, bool)
        self.v = np.zeros_like(self.x)
        self.xn = self.x.copy()
        self.xt = self.x.copy()
        self.xprev = self.x.copy()
        self.p = np.zeros_like(self.x)
        self.rows = [np.zeros((0, w), np.int32) for w in WIDTH]
        self.max_newton = max_newton
        self.max_halvings = max_halvings
        self.inversion = inversion
        self.tol = scene.newton_tol
        self.eta = scene.eta
        self.frame = 0
        self.cap = np.full(7, 4096, np.int64)
        self.timings = {}

    # -------------------------------------------------------------- helpers
    def _rows_ptrs(self):
        rows = [np.ascontiguousarray(r, np.int32) if len(r) else np.zeros((1, WIDTH[t]), np.int32)
                for t, r in enumerate(self.rows)]
        arr = (C.POINTER(C.c_int) * 7)(*[L.iptr(r) for r in rows])
        n = np.array([len(r) for r in self.rows], np.int64)
        return rows, arr, n

    def _tick(self, name, t0):
        self.timings[name] = self.timings.get(name, 0.0) + time.perf_counter() - t0

    # ------------------------------------------------- reference step functions
    def compute_xn_and_xTilde(self, t):
        self.xn = self.x.copy()
        self.xt = self.x + self.dt * self.v
        self.xt[:, 1] += self.dt * self.dt * self.scene.gravity
        target = self.scene.dirichlet(t)
        self.xt[self.dmask] = target[self.dmask]

    def find_constraints(self):
        t0 = time.perf_counter()
        while True:
            bufs = [np.zeros((int(self.cap[t]), WIDTH[t]), np.int32) for t in range(7)]
            arr = (C.POINTER(C.c_int) * 7)(*[L.iptr(b) for b in bufs])
            counts = np.zeros(7, np.int64)
            self.lib.orc_find_constraints(L.dptr(self.x), L.dptr(self.x0), len(self.bp), L.iptr(self.bp),
                                          len(self.be), L.iptr(self.be), len(self.bt), L.iptr(self.bt),
                                          self.dHat, self.inv_dx, arr, self.cap.ctypes.data_as(C.POINTER(C.c_int64)),
                                          counts.ctypes.data_as(C.POINTER(C.c_int64)))
            if np.all(counts <= self.cap):
                break
            self.cap = np.maximum(self.cap, counts * 2)
        # np.unique per type (NEWTON_IPC.py:348-356)
        self.rows = [np.unique(bufs[t][:counts[t]], axis=0) if counts[t] else np.zeros((0, WIDTH[t]), np.int32)
                     for t in range(7)]
        self._tick("find", t0)
        return [len(r) for r in self.rows]

    def constraints(self):
        return {NAMES[t]: self.rows[t] for t in range(7)}

    def assemble(self):
        """compute_hessian_and_gradient (NEWTON_IPC.py:736-747): COO + rhs."""
        t0 = time.perf_counter()
        rows, arr, n = self._rows_ptrs()
        size = self.lib.orc_coo_size(self.nv, self.nt, n.ctypes.data_as(C.POINTER(C.c_int64)))
        ci = np.zeros(size, np.int32)
        cj = np.zeros(size, np.int32)
        cv = np.zeros(size)
        rhs = np.zeros(3 * self.nv)
        self.lib.orc_assemble(self.nv, L.dptr(self.x), L.dptr(self.xt), L.dptr(self.m), L.dptr(self.x0),
                              self.nt, L.iptr(self.tets), L.dptr(self.restT), L.dptr(self.la), L.dptr(self.mu), self.dt,
                              arr, n.ctypes.data_as(C.POINTER(C.c_int64)), self.dHat2, self.kappa,
                              L.iptr(ci), L.iptr(cj), L.dptr(cv), L.dptr(rhs))
        self.coo = (ci, cj, cv)
        self.rhs = rhs
        self._tick("assemble", t0)
        return ci, cj, cv, rhs

    def matrix(self):
        ci, cj, cv = self.coo
        n = 3 * self.nv
        return scipy.sparse.csr_matrix((cv, (ci, cj)), shape=(n, n))

    def solve_system(self, t=0.0):
        """solve_system (NEWTON_IPC.py:750-777) with prescribed Dirichlet increments."""
        t0 = time.perf_counter()
        n = 3 * self.nv
        A = self.matrix()
        D = np.repeat(self.dmask, 3)
        pD = np.zeros(n)
        pD[D] = (self.xt - self.x).reshape(-1)[D]
        b = self.rhs - A @ pD
        keep = scipy.sparse.diags((~D).astype(float))
        A = keep @ A @ keep + scipy.sparse.diags(D.astype(float))
        b[D] = 0.0
        A = A.tocsc()
        if _cholesky is not None:
            sol = _cholesky(A)(b)
        else:
            # SPD system: symmetric ordering, no threshold pivoting
            sol = scipy.sparse.linalg.splu(A, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                                           options=dict(SymmetricMode=True)).solve(b)
        sol[D] = pD[D]
        self.p = sol.reshape(-1, 3)
        self._tick("solve", t0)
        return sol

    def output_residual(self):
        return float(np.abs(self.p).max())

    def compute_energy(self, parts=False):
        t0 = time.perf_counter()
        rows, arr, n = self._rows_ptrs()
        pr = np.zeros(3)
        e = self.lib.orc_energy(self.nv, L.dptr(self.x), L.dptr(self.xt), L.dptr(self.m), L.dptr(self.x0),
                                self.nt, L.iptr(self.tets), L.dptr(self.restT), L.dptr(self.la), L.dptr(self.mu), self.dt,
                                arr, n.ctypes.data_as(C.POINTER(C.c_int64)), self.dHat2, self.kappa, L.dptr(pr))
        self._tick("energy", t0)
        return (e, pr) if parts else e

    def compute_intersection_free_step_size(self, parts=False):
        t0 = time.perf_counter()
        p = np.ascontiguousarray(self.p)
        pr = np.zeros(3)
        a = self.lib.orc_ccd_alpha(L.dptr(self.x), L.dptr(p), len(self.bp), L.iptr(self.bp), len(self.be), L.iptr(self.be),
                                   len(self.bt), L.iptr(self.bt), self.nt, L.iptr(self.tets), self.dHat, self.inv_dx,
                                   self.eta, int(self.inversion), L.dptr(pr))
        self._tick("ccd", t0)
        return (a, pr) if parts else a

    def check_collision(self):
        t0 = time.perf_counter()
        h = self.lib.orc_check_collision(L.dptr(self.x), len(self.be), L.iptr(self.be), len(self.bt), L.iptr(self.bt),
                                         self.dHat, self.inv_dx)
        self._tick("collision", t0)
        return h

    def save_xPrev(self):
        self.xprev = self.x.copy()

    def apply_sol(self, alpha):
        self.x = np.ascontiguousarray(self.xprev + self.p * alpha)

    def compute_v(self):
        self.v = (self.x - self.xn) / self.dt

    # ------------------------------------------------------------ frame loop
    def step(self):
        f = self.frame
        self.compute_xn_and_xTilde(f * self.dt)
        it = 0
        alphas, residuals = [], []
        stalled = False
        while True:
            it += 1
            self.find_constraints()
            self.assemble()
            self.solve_system(f * self.dt)
            res = self.output_residual()
            residuals.append(res)
            if res < self.tol or it >= self.max_newton:
                break
            E0 = self.compute_energy()
            self.save_xPrev()
            alpha = self.compute_intersection_free_step_size()
            self.apply_sol(alpha)
            n = 0
            while self.check_collision() and n < self.max_halvings:
                alpha /= 2.0
                self.apply_sol(alpha)
                n += 1
            self.find_constraints()
            E = self.compute_energy()
            n = 0
            while E > E0 and n < self.max_halvings:
                alpha *= 0.5
                self.apply_sol(alpha)
                self.find_constraints()
                E = self.compute_energy()
                n += 1
            n = 0
            while self.check_collision() and n < self.max_halvings:
                alpha /= 2.0
                self.apply_sol(alpha)
                n += 1
            alphas.append(alpha)
            if alpha * res < 1e-3 * self.tol:
                stalled = True  # the product's stall exit, same criterion
                break
        self.compute_v()
        self.frame += 1
        return {"newton_iters": it, "alphas": alphas, "residuals": residuals, "stalled": stalled}
