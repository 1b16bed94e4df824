"""ctypes bindings of the C oracle (``libipc_oracle.so``) and, when it has been
built, of the reference's own native library (``_ref/libref.so``).

Test infrastructure only (see the package docstring)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.c_double
_P = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_LP = C.POINTER(C.c_int64)

KINDS = ("PP", "PE", "PT", "EE", "EEM", "PPM", "PEM")
WIDTH = (2, 3, 4, 4, 4, 4, 4)
NPTS = (2, 3, 4, 4, 4, 4, 4)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    path = os.path.join(HERE, "libipc_oracle.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    sig = {
        "orc_barrier_E": (_D, [_D, _D, _D]), "orc_barrier_g": (_D, [_D, _D, _D]), "orc_barrier_H": (_D, [_D, _D, _D]),
        "orc_PP_E": (_D, [_P] * 2), "orc_PE_E": (_D, [_P] * 3), "orc_PT_E": (_D, [_P] * 4), "orc_EE_E": (_D, [_P] * 4),
        "orc_EECN2_E": (_D, [_P] * 4), "orc_M_threshold": (_D, [_P] * 4),
        "orc_dist_gH": (None, [C.c_int, _P, _P, _P, _P]),
        "orc_PT_type": (C.c_int, [_P] * 4), "orc_EE_type": (C.c_int, [_P] * 4), "orc_PE_type": (C.c_int, [_P] * 3),
        "orc_PT_dist2": (_D, [_P] * 4 + [C.c_int]), "orc_EE_dist2": (_D, [_P] * 4 + [C.c_int]),
        "orc_make_pd": (None, [C.c_int, _P]), "orc_make_pd2": (None, [_P]),
        "orc_constraint_E": (_D, [C.c_int, _P, _P, _D, _D]),
        "orc_constraint_gH": (_D, [C.c_int, _P, _P, _D, _D, C.c_int, _P, _P]),
        "orc_svd3": (None, [_P] * 4), "orc_nh_energy": (_D, [_P, _D, _D]), "orc_nh_P": (None, [_P, _D, _D, _P]),
        "orc_nh_dPdF": (None, [_P, _D, _D, _P]),
        "orc_quad_root": (_D, [_D] * 4), "orc_cubic_root": (_D, [_D] * 5),
        "orc_tet_inversion_alpha": (_D, [_P] * 8 + [_D]),
        "orc_pt_ccd": (_D, [_P] * 8 + [_D, _D]), "orc_ee_ccd": (_D, [_P] * 8 + [_D, _D]),
        "orc_ccd_one": (_D, [C.c_int, C.c_int] + [_P] * 8 + [_D, _D]),
        "orc_seg_tri": (C.c_int, [_P] * 5),
        "orc_find_constraints": (None, [_P, _P, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, _D, _D,
                                        C.POINTER(_IP), _LP, _LP]),
        "orc_tet_restT": (None, [C.c_int, _P, _IP, _P, _P]),
        "orc_tet_gH": (None, [_P, _IP, _P, _D, _D, _D, _P, _P]),
        "orc_tet_E": (_D, [_P, _IP, _P, _D, _D, _D]),
        "orc_coo_size": (C.c_int64, [C.c_int, C.c_int, _LP]),
        "orc_assemble": (None, [C.c_int, _P, _P, _P, _P, C.c_int, _IP, _P, _P, _P, _D,
                                C.POINTER(_IP), _LP, _D, _D, _IP, _IP, _P, _P]),
        "orc_energy": (_D, [C.c_int, _P, _P, _P, _P, C.c_int, _IP, _P, _P, _P, _D,
                            C.POINTER(_IP), _LP, _D, _D, _P]),
        "orc_ccd_alpha": (_D, [_P, _P, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, _D, _D, _D, C.c_int, _P]),
        "orc_check_collision": (C.c_int, [_P, C.c_int, _IP, C.c_int, _IP, _D, _D]),
        "orc_num_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def dptr(a):
    return a.ctypes.data_as(_P)


def iptr(a):
    return a.ctypes.data_as(_IP)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------------------
# scalar helpers (numpy in, numpy out)
# ---------------------------------------------------------------------------
def barrier(d, dhat2, kappa):
    L = lib()
    return L.orc_barrier_E(d, dhat2, kappa), L.orc_barrier_g(d, dhat2, kappa), L.orc_barrier_H(d, dhat2, kappa)


def dist_gH(kind: int, pts):
    """kind 0 PP, 1 PE, 2 PT, 3 EE, 4 |(a1-a0)x(b1-b0)|^2 -> (E, g, H)."""
    pts = f64(pts).ravel()
    n = {0: 6, 1: 9}.get(kind, 12)
    E = np.zeros(1)
    g = np.zeros(n)
    H = np.zeros((n, n))
    lib().orc_dist_gH(kind, dptr(pts), dptr(E), dptr(g), dptr(H))
    return E[0], g, H


def dist_E(kind: int, pts):
    pts = [f64(p) for p in np.asarray(pts, np.float64).reshape(-1, 3)]
    L = lib()
    fn = [L.orc_PP_E, L.orc_PE_E, L.orc_PT_E, L.orc_EE_E, L.orc_EECN2_E][kind]
    return fn(*[dptr(p) for p in pts])


def pt_type(p, t0, t1, t2):
    return lib().orc_PT_type(*[dptr(f64(v)) for v in (p, t0, t1, t2)])


def ee_type(a0, a1, b0, b1):
    return lib().orc_EE_type(*[dptr(f64(v)) for v in (a0, a1, b0, b1)])


def make_pd(H):
    H = f64(H).copy()
    n = H.shape[0]
    if n == 2:
        lib().orc_make_pd2(dptr(H))
    else:
        lib().orc_make_pd(n, dptr(H))
    return H


def constraint_gH(kind: int, x, x0=None, dhat2=1e-6, kappa=1.0, project=True):
    x = f64(x).ravel()
    x0 = f64(x if x0 is None else x0).ravel()
    n = (6, 9, 12, 12, 12, 12, 12)[kind]
    g = np.zeros(n)
    H = np.zeros((n, n))
    E = lib().orc_constraint_gH(kind, dptr(x), dptr(x0), dhat2, kappa, int(project), dptr(g), dptr(H))
    return E, g, H


def constraint_E(kind: int, x, x0=None, dhat2=1e-6, kappa=1.0):
    x = f64(x).ravel()
    x0 = f64(x if x0 is None else x0).ravel()
    return lib().orc_constraint_E(kind, dptr(x), dptr(x0), dhat2, kappa)


def nh_dPdF(F, la, mu):
    F = f64(F)
    out = np.zeros((9, 9))
    lib().orc_nh_dPdF(dptr(F), la, mu, dptr(out))
    return out


def nh_P(F, la, mu):
    F = f64(F)
    out = np.zeros((3, 3))
    lib().orc_nh_P(dptr(F), la, mu, dptr(out))
    return out


def svd3(F):
    F = f64(F)
    U, s, V = np.zeros((3, 3)), np.zeros(3), np.zeros((3, 3))
    lib().orc_svd3(dptr(F), dptr(U), dptr(s), dptr(V))
    return U, s, V


def cubic_root(a, b, c, d, tol=1e-8):
    return lib().orc_cubic_root(a, b, c, d, tol)


def pt_ccd(p, t0, t1, t2, dp, dt0, dt1, dt2, eta, dist2):
    return lib().orc_pt_ccd(*[dptr(f64(v)) for v in (p, t0, t1, t2, dp, dt0, dt1, dt2)], eta, dist2)


def ee_ccd(a0, a1, b0, b1, da0, da1, db0, db1, eta, dist2):
    return lib().orc_ee_ccd(*[dptr(f64(v)) for v in (a0, a1, b0, b1, da0, da1, db0, db1)], eta, dist2)


def seg_tri(P, Q, A, B, Cc):
    return lib().orc_seg_tri(*[dptr(f64(v)) for v in (P, Q, A, B, Cc)])


def ccd_one(kind, fused, pts, disp, eta, dist2):
    """One rounding of the CCD alone (kind 0 point-triangle, 1 edge-edge):
    fused multiply-adds or strict IEEE; orc_pt_ccd / orc_ee_ccd return the
    smaller of the two."""
    return lib().orc_ccd_one(int(kind), int(fused), *[dptr(f64(v)) for v in list(pts) + list(disp)], eta, dist2)


def tet_inversion_alpha(xs, ps, eta=0.2):
    return lib().orc_tet_inversion_alpha(*[dptr(f64(v)) for v in list(xs) + list(ps)], eta)


# ---------------------------------------------------------------------------
# the reference's own compiled library (oracle/_ref/libref.so)
# ---------------------------------------------------------------------------
_ref = None


def ref_path():
    return os.path.join(HERE, "_ref", "libref.so")


def ref():
    """The reference's wrapper.cpp/CTCD.cpp compiled by ``build_ref.sh``, or None."""
    global _ref
    if _ref is None and os.path.exists(ref_path()):
        r = C.CDLL(ref_path())
        for n in (2, 3, 4, 6, 9, 12):
            f = getattr(r, f"project_pd_{n}")
            f.restype = None
            f.argtypes = [_D] * (n * n) + [_P] * (n * n)
        for n in (2, 3, 4, 6, 9):
            f = getattr(r, f"solve_{n}")
            f.restype = None
            f.argtypes = [_D] * (n * n + n) + [_P] * n
        r.svd_2.restype = None
        r.svd_2.argtypes = [_D] * 4 + [_P] * 12
        r.svd_3.restype = None
        r.svd_3.argtypes = [_D] * 9 + [_P] * 27
        r.point_triangle_ccd.restype = None
        r.point_triangle_ccd.argtypes = [_D] * 26 + [_P]
        r.edge_edge_ccd.restype = None
        r.edge_edge_ccd.argtypes = [_D] * 26 + [_P]
        r.get_smallest_positive_real_cubic_root.restype = None
        r.get_smallest_positive_real_cubic_root.argtypes = [_D] * 5 + [_P]
        _ref = r
    return _ref


def ref_project_pd(H):
    H = f64(H)
    n = H.shape[0]
    outs = [C.c_double() for _ in range(n * n)]
    getattr(ref(), f"project_pd_{n}")(*[float(v) for v in H.ravel()], *[C.byref(o) for o in outs])
    return np.array([o.value for o in outs]).reshape(n, n)


def ref_solve(A, b):
    A = f64(A)
    n = A.shape[0]
    outs = [C.c_double() for _ in range(n)]
    getattr(ref(), f"solve_{n}")(*[float(v) for v in A.ravel()], *[float(v) for v in np.ravel(b)],
                                  *[C.byref(o) for o in outs])
    return np.array([o.value for o in outs])


def ref_svd2(F):
    outs = [C.c_double() for _ in range(12)]
    ref().svd_2(*[float(v) for v in np.ravel(F)], *[C.byref(o) for o in outs])
    v = np.array([o.value for o in outs])
    return v[:4].reshape(2, 2), v[4:8].reshape(2, 2), v[8:].reshape(2, 2)


def ref_svd3(F):
    outs = [C.c_double() for _ in range(27)]
    ref().svd_3(*[float(v) for v in np.ravel(F)], *[C.byref(o) for o in outs])
    v = np.array([o.value for o in outs])
    return v[:9].reshape(3, 3), v[9:18].reshape(3, 3), v[18:].reshape(3, 3)


def ref_pt_ccd(p, t0, t1, t2, dp, dt0, dt1, dt2, eta, dist2):
    out = C.c_double()
    args = [float(v) for arr in (p, t0, t1, t2, dp, dt0, dt1, dt2) for v in np.asarray(arr).ravel()]
    ref().point_triangle_ccd(*args, eta, dist2, C.byref(out))
    return out.value


def ref_ee_ccd(a0, a1, b0, b1, da0, da1, db0, db1, eta, dist2):
    out = C.c_double()
    args = [float(v) for arr in (a0, a1, b0, b1, da0, da1, db0, db1) for v in np.asarray(arr).ravel()]
    ref().edge_edge_ccd(*args, eta, dist2, C.byref(out))
    return out.value


def ref_cubic_root(a, b, c, d, tol=1e-8):
    out = C.c_double()
    ref().get_smallest_positive_real_cubic_root(a, b, c, d, tol, C.byref(out))
    return out.value
