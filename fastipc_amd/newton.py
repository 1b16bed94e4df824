"""Projected-Newton IPC time stepping on one MI355X.

Host driver of ``projects/FastIPC/NEWTON_IPC.py``: the module-level step
functions of the reference become methods with the same names, each calling
one batched entry point of libfastipc.so, and :meth:`NewtonIPC.step` runs the
reference's frame loop (NEWTON_IPC.py:955-1002):

* Newton iterations stop when ``max|p| < tol`` *before* the line search
  (:975), the final direction is not applied;
* the line search takes the CCD / inversion-free step (:980-982), halves while
  an edge crosses a triangle (:983-985), halves while ``E > E0`` re-running
  the constraint search (:986-992), then halves on collision once more
  (:993-995).

Differences, all documented: the linear solve is a block-Jacobi PCG instead
of CHOLMOD; Dirichlet vertices whose prescribed position moves between steps
reach it through a prescribed increment in the solve (static Dirichlet
vertices behave exactly as in the reference); safety caps bound the Newton and
halving loops, which the reference leaves unbounded, and a step whose line
search no longer moves x (alpha max|p| below 1e-3 of the Newton tolerance)
ends there and is reported as stalled, where the reference would loop
forever.

Inexact search directions (``pcg_tol_search``, off by default): the solves
whose direction is applied stop at the looser relative tolerance
``pcg_tol_search``; when such a direction passes the convergence test, the
solve is continued to ``pcg_tol`` (``fipc_solve_refine``) and the test is
repeated on the refined direction, so a step is accepted only on a direction
solved to ``pcg_tol``, as without the option.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .model import SceneSetup
from .timer import Timers

TYPE_NAMES = ("PP", "PE", "PT", "EE", "EEM", "PPM", "PEM")
TYPE_WIDTH = (2, 3, 4, 4, 4, 4, 4)


@dataclass
class StepStats:
    newton_iters: int = 0
    pcg_iters: list = field(default_factory=list)
    alphas: list = field(default_factory=list)
    residuals: list = field(default_factory=list)
    counts: list = field(default_factory=list)
    energy: float = float("nan")
    converged: bool = True
    warm_accepted: bool = False
    # solves that ended without meeting the PCG tolerance (their directions
    # are applied by the line search but never decide convergence)
    unconverged_solves: int = 0
    relres: list = field(default_factory=list)
    # the line search stopped making progress: the applied step alpha max|p|
    # fell below 1e-3 of the Newton tolerance while the test still failed
    stalled: bool = False


class NewtonIPC:
    def __init__(self, scene, precision: int = 0, pcg_tol: float = 1e-6, pcg_maxit: int = 20000,
                 max_newton: int = 200, max_halvings: int = 60, inversion: bool = True, timing: bool = False,
                 pcg_tol_search: float | None = None, certify: bool = False):
        self.so = _lib.load()
        self.scene = scene
        self.setup = s = SceneSetup(scene)
        self.precision = precision
        self.pcg_tol = pcg_tol
        if pcg_tol_search is not None and not pcg_tol_search >= pcg_tol:
            raise ValueError("pcg_tol_search must be at least pcg_tol")
        self.pcg_tol_search = pcg_tol_search
        # with inexact search directions: the deciding direction's
        # continuation stops once the exact direction provably passes the
        # Newton test (fipc_solve_certify; FP32 mode)
        self.certify = certify
        self.certified_solves = 0
        self.pcg_maxit = pcg_maxit
        self.last_pcg_iters = 0
        self.last_solve_converged = True
        self.max_newton = max_newton
        self.max_halvings = max_halvings
        self.inversion = inversion
        self.nv = len(s.x0)
        self.timers = Timers()
        ctx = C.c_void_p()
        _lib.call("fipc_create", C.byref(ctx))
        self.ctx = ctx
        _lib.call("fipc_upload_mesh", ctx, self.nv, _lib.dptr(s.x0), _lib.dptr(np.ascontiguousarray(s.mass)),
                  len(s.tets), _lib.iptr(s.tets), _lib.dptr(np.ascontiguousarray(s.IB.reshape(-1, 9))),
                  _lib.dptr(np.ascontiguousarray(s.vol)), _lib.dptr(_lib.f64(s.la)), _lib.dptr(_lib.f64(s.mu)),
                  len(s.bp), _lib.iptr(s.bp), len(s.be), _lib.iptr(s.be), len(s.bt), _lib.iptr(s.bt),
                  s.dirichlet.ctypes.data_as(C.POINTER(C.c_ubyte)))
        _lib.call("fipc_set_params", ctx, s.dt, s.dHat, s.kappa, s.inv_dx, s.gravity)
        if timing:
            _lib.call("fipc_enable_timing", ctx, 1)
        self.frame = 0
        self.counts = [0] * 7

    def close(self):
        if self.ctx:
            _lib.call("fipc_destroy", self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ state
    @property
    def x(self) -> np.ndarray:
        out = np.empty((self.nv, 3))
        _lib.call("fipc_get_state", self.ctx, _lib.dptr(out), None)
        return out

    @property
    def v(self) -> np.ndarray:
        out = np.empty((self.nv, 3))
        _lib.call("fipc_get_state", self.ctx, None, _lib.dptr(out))
        return out

    # The constraint set and the energy of the current x are kept until x (or
    # xTilde, for the energy) changes, so the frame loop does not recompute
    # what the line search has just computed at the same x: the Newton
    # iteration reuses the constraints found after the last step of the line
    # search, and the next line search starts from the energy evaluated there
    # (same x, same constraints, deterministic reduction: the same bits).
    # The collision test's result is kept the same way: the test after the
    # energy backtracking repeats the one after the CCD step whenever no
    # halving moved x in between.
    def _invalidate(self, x_changed: bool = True):
        if x_changed:
            self._cons_ok = False
            self._hits = None
        self._E = None

    def _ensure_constraints(self):
        if not getattr(self, "_cons_ok", False):
            self.find_constraints()

    def _energy(self) -> float:
        if getattr(self, "_E", None) is None:
            self.compute_energy()
        return self._E

    def set_state(self, x=None, v=None):
        self._invalidate()
        xa = _lib.f64(x) if x is not None else None
        va = _lib.f64(v) if v is not None else None
        _lib.call("fipc_set_state", self.ctx, _lib.dptr(xa) if xa is not None else None,
                  _lib.dptr(va) if va is not None else None)

    def vector(self, which: int) -> np.ndarray:
        out = np.empty((self.nv, 3))
        _lib.call("fipc_get_vector", self.ctx, which, _lib.dptr(out))
        return out

    def set_vector(self, which: int, a):
        if which in (0, 3):
            self._invalidate(x_changed=which == 0)
        a = _lib.f64(a).reshape(self.nv, 3)
        _lib.call("fipc_set_vector", self.ctx, which, _lib.dptr(a))

    # ------------------------------------------------- reference step functions
    def compute_xn_and_xTilde(self, t: float):
        """compute_xn_and_xTilde + move_nodes (NEWTON_IPC.py:497-511)."""
        rows, target = self.scene.dirichlet_rows(t)
        _lib.call("fipc_begin_step_rows", self.ctx, len(rows), _lib.iptr(rows), _lib.dptr(target))
        self._invalidate(x_changed=False)  # xTilde changed: the energy, not the constraints

    def find_constraints(self):
        counts = np.zeros(7, np.int32)
        _lib.call("fipc_find_constraints", self.ctx, _lib.iptr(counts))
        self.counts = [int(c) for c in counts]
        self._cons_ok = True
        self._E = None
        return self.counts

    def search_stats(self, on: bool = True):
        """Count the broad-phase candidates of the constraint searches from
        now on (off by default: the counting slows the queries)."""
        _lib.call("fipc_search_stats", self.ctx, int(on))

    def search_counts(self):
        """Broad-phase candidates (point-triangle, edge-edge) of the last
        constraint search: pairs whose boxes overlap within dHat (-1 unless
        search_stats was on for that search)."""
        cand = np.zeros(2, np.int32)
        _lib.call("fipc_search_counts", self.ctx, _lib.iptr(cand))
        return [int(c) for c in cand]

    def constraints(self, sort: bool = True):
        """Host copy of the seven row sets, in np.unique order when ``sort``."""
        out = {}
        for t, name in enumerate(TYPE_NAMES):
            n = self.counts[t]
            rows = np.zeros((n, TYPE_WIDTH[t]), np.int32)
            if n:
                _lib.call("fipc_get_constraints", self.ctx, t, _lib.iptr(rows))
            if sort and n:
                rows = np.unique(rows, axis=0)
            out[name] = rows
        return out

    def output_current_minimal_distance(self) -> float:
        d = C.c_double()
        _lib.call("fipc_min_distance", self.ctx, C.byref(d))
        return d.value

    def compute_hessian_and_gradient(self):
        _lib.call("fipc_assemble", self.ctx, self.precision)

    def solve_system(self, t: float = 0.0, tol: float | None = None):
        it = C.c_int()
        rel = C.c_double()
        tol = self.pcg_tol if tol is None else tol
        code = self.so.fipc_solve(self.ctx, tol, self.pcg_maxit, C.byref(it), C.byref(rel))
        if code not in (0, 5):
            raise _lib.FipcError("fipc_solve", code)
        self.last_pcg_iters = it.value
        self.last_solve_converged = code == 0
        return it.value, rel.value

    def refine_solve(self, tol: float | None = None, bound: float | None = None):
        """Continues the last solve of the current system to ``tol``
        (default ``pcg_tol``) from its solution; with ``bound``, stops as soon
        as max|p*| < bound is certified for the exact direction p*
        (fipc_solve_certify), which counts as converged."""
        it = C.c_int()
        rel = C.c_double()
        tol = self.pcg_tol if tol is None else tol
        maxit = max(self.pcg_maxit - self.last_pcg_iters, 1)
        if bound is None:
            code = self.so.fipc_solve_refine(self.ctx, tol, maxit, C.byref(it), C.byref(rel))
            name = "fipc_solve_refine"
        else:
            cert = C.c_int()
            code = self.so.fipc_solve_certify(self.ctx, tol, maxit, bound, C.byref(it), C.byref(rel), C.byref(cert))
            name = "fipc_solve_certify"
            self.certified_solves += cert.value
        if code not in (0, 5):
            raise _lib.FipcError(name, code)
        self.last_solve_converged = code == 0
        return it.value, rel.value

    def output_residual(self) -> float:
        r = C.c_double()
        _lib.call("fipc_residual", self.ctx, C.byref(r))
        return r.value

    def compute_energy(self, parts: bool = False):
        e = C.c_double()
        p = np.zeros(3)
        _lib.call("fipc_energy", self.ctx, C.byref(e), _lib.dptr(p))
        self._E = e.value
        return (e.value, p) if parts else e.value

    def compute_intersection_free_step_size(self, parts: bool = False):
        a = C.c_double()
        p = np.zeros(3)
        _lib.call("fipc_ccd_alpha", self.ctx, self.setup.eta, int(self.inversion), C.byref(a), _lib.dptr(p))
        return (a.value, p) if parts else a.value

    def check_collision(self) -> int:
        if getattr(self, "_hits", None) is None:
            h = C.c_int()
            _lib.call("fipc_check_collision", self.ctx, C.byref(h))
            self._hits = h.value
        return self._hits

    def save_xPrev(self):
        _lib.call("fipc_save_prev", self.ctx)

    def apply_sol(self, alpha: float):
        _lib.call("fipc_apply_sol", self.ctx, float(alpha))
        self._invalidate()

    def compute_v(self):
        _lib.call("fipc_end_step", self.ctx)

    # --------------------------------------------------------------- frame loop
    def step(self, warm_start=None, capture=None) -> StepStats:
        """One time step (NEWTON_IPC.py:956-999).  ``warm_start`` optionally
        proposes an increment before the first Newton iteration (a callable
        that returns a host array, or True after writing the device direction
        vector); it is kept only if it is intersection-free and lowers the
        incremental potential, otherwise x = x_n is used as in the reference.
        ``capture`` (fastipc_amd.warmstart.Capture) records training rows of
        the step."""
        st = StepStats()
        T = self.timers
        f = self.frame
        dt = self.setup.dt
        with T("step"):
            self.compute_xn_and_xTilde(f * dt)
            if capture is not None:
                self._ensure_constraints()
                capture.begin(self)
            if warm_start is not None:
                with T("warm_start"):
                    st.warm_accepted = self._try_warm_start(warm_start)
            while True:
                st.newton_iters += 1
                with T("build"):
                    with T("find"):
                        self._ensure_constraints()
                        st.counts.append(list(self.counts))
                    with T("assemble"):
                        self.compute_hessian_and_gradient()
                with T("solve"):
                    it, rel = self.solve_system(f * dt, self.pcg_tol_search)
                res = self.output_residual()
                if self.pcg_tol_search is not None and res < self.setup.newton_tol and self.last_solve_converged:
                    # the direction that decides convergence is solved to
                    # pcg_tol, or until its test is certified
                    with T("solve"):
                        it2, rel = self.refine_solve(bound=self.setup.newton_tol if self.certify else None)
                    it += it2
                    res = self.output_residual()
                st.pcg_iters.append(it)
                st.residuals.append(res)
                st.relres.append(rel)
                if not self.last_solve_converged:
                    # a direction short of the PCG tolerance underestimates
                    # max |p| (CG iterates grow from zero): it is applied,
                    # never trusted for the convergence test
                    st.unconverged_solves += 1
                elif res < self.setup.newton_tol:
                    break
                if st.newton_iters >= self.max_newton:
                    st.converged = False
                    break
                with T("line_search"):
                    alpha = self._line_search()
                    st.alphas.append(alpha)
                if alpha * res < 1e-3 * self.setup.newton_tol:
                    # The reference would iterate forever here: its energy is
                    # discontinuous where an edge pair's closest-feature type
                    # changes and np.unique had merged the point-edge row it
                    # emitted with one from the point-triangle search
                    # (NEWTON_IPC.py:235-357), so E > E0 for every step the
                    # halving tries.  The step ends at the last iterate.
                    st.stalled = True
                    st.converged = False
                    break
            if capture is not None:
                capture.end(self)
            self.compute_v()
        st.energy = float("nan")
        self.frame += 1
        T.end_round()
        return st

    def _line_search(self) -> float:
        T = self.timers
        with T("energy"):
            E0 = self._energy()
        self.save_xPrev()
        with T("ccd"):
            alpha = self.compute_intersection_free_step_size()
        self.apply_sol(alpha)
        n = 0
        with T("collision"):
            while self.check_collision() and n < self.max_halvings:
                alpha /= 2.0
                self.apply_sol(alpha)
                n += 1
        with T("find"):
            self.find_constraints()
        with T("energy"):
            E = self.compute_energy()
        n = 0
        while E > E0 and n < self.max_halvings:
            alpha *= 0.5
            self.apply_sol(alpha)
            with T("find"):
                self.find_constraints()
            with T("energy"):
                E = self.compute_energy()
            n += 1
        n = 0
        with T("collision"):
            while self.check_collision() and n < self.max_halvings:
                alpha /= 2.0
                self.apply_sol(alpha)
                n += 1
        return alpha

    def _try_warm_start(self, fn, max_halvings: int = 4) -> bool:
        """Line search on the proposed increment (the steps of :980-995):
        CCD / inversion bound, halve on collision, then halve while the
        incremental potential is not below its value at x_n; if no halving
        succeeds, x = x_n."""
        self._ensure_constraints()
        self.save_xPrev()
        # the prediction is queued first: E0 (at x_n; the prediction writes
        # only the direction) then waits for both at one synchronisation
        dx = fn(self)
        E0 = self._energy()
        if dx is None or dx is False:
            return False
        if dx is not True:
            self.set_vector(1, np.ascontiguousarray(dx, np.float64).reshape(-1, 3))
        alpha = min(1.0, self.compute_intersection_free_step_size())
        self.apply_sol(alpha)
        n = 0
        while self.check_collision():
            n += 1
            if n > self.max_halvings:
                self.apply_sol(0.0)
                return False
            alpha /= 2.0
            self.apply_sol(alpha)
        for _ in range(max_halvings + 1):
            self.find_constraints()
            if self.compute_energy() < E0:
                return True
            alpha *= 0.5
            self.apply_sol(alpha)
        self.apply_sol(0.0)
        return False
