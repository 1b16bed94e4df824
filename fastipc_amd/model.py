"""Set-up quantities of the Newton IPC step, computed once per scene.

Mirrors ``compute_restT_and_m`` (NEWTON_IPC.py:485-494), the Lame parameters
(:145-155), ``compute_adaptive_kappa`` (:116-124, times 10 at :943) and
``compute_mean_of_boundary_edges`` (:127-134, the hash cell size at :944).
"""
from __future__ import annotations

import math

import numpy as np


def barrier_H(d, dHat2, kappa):
    """Second derivative of the barrier (common/math/distance.py:38-44)."""
    if d >= dHat2:
        return 0.0
    t2 = d - dHat2
    return kappa * ((math.log(d / dHat2) * -2.0 - t2 * 4.0 / d) + 1.0 / (d * d) * (t2 * t2))


def rest_shape(x: np.ndarray, tets: np.ndarray):
    """D_m (columns x1-x0, x2-x0, x3-x0), its inverse and the rest volume."""
    x = np.asarray(x, np.float64)
    t = np.asarray(tets, np.int64)
    Dm = np.stack([x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]], x[t[:, 3]] - x[t[:, 0]]], axis=2)
    det = np.linalg.det(Dm)
    if np.any(det <= 0):
        raise ValueError("mesh inverted: a tet has non-positive rest volume")
    IB = np.linalg.inv(Dm)
    return Dm, IB, det / 6.0


def lumped_mass(nv, tets, vol, density):
    """Each tet gives vol * rho / 4 to each of its vertices (NEWTON_IPC.py:489-493)."""
    m = np.zeros(nv)
    share = vol * density / 4.0
    for a in range(4):
        np.add.at(m, tets[:, a], share)
    return m


def lame(E, nu):
    E = np.asarray(E, np.float64)
    nu = np.asarray(nu, np.float64)
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def adaptive_kappa(m, dHat2):
    """10 * 1e13 * mean mass / (4e-16 * b''(1e-16; dHat^2, 1))."""
    return 10.0 * 1.0e13 * float(np.sum(m)) / len(m) / (4.0e-16 * barrier_H(1.0e-16, dHat2, 1.0))


def mean_boundary_edge(x, edges):
    e = np.asarray(edges)
    return float(np.mean(np.linalg.norm(x[e[:, 0]] - x[e[:, 1]], axis=1)))


class SceneSetup:
    """Everything the solver derives from a scene before the first step."""

    def __init__(self, scene):
        self.scene = scene
        x = np.ascontiguousarray(scene.x, np.float64)
        tets = np.ascontiguousarray(scene.tets, np.int32)
        self.x0 = x.copy()
        self.tets = tets
        self.Dm, self.IB, self.vol = rest_shape(x, tets)
        self.mass = lumped_mass(len(x), tets, self.vol, scene.density)
        self.la, self.mu = lame(scene.youngs, scene.poisson)
        self.dHat = scene.dhat
        self.dHat2 = self.dHat * self.dHat
        self.kappa = adaptive_kappa(self.mass, self.dHat2)
        bp, be, bt = scene.boundary
        self.bp = np.ascontiguousarray(bp, np.int32)
        self.be = np.ascontiguousarray(be, np.int32)
        self.bt = np.ascontiguousarray(bt, np.int32)
        self.inv_dx = 1.0 / mean_boundary_edge(x, self.be)
        self.dirichlet = np.ascontiguousarray(scene.dirichlet_mask, np.uint8)
        self.dt = scene.dt
        self.gravity = scene.gravity
        self.newton_tol = scene.newton_tol
        self.eta = scene.eta
