"""CPU-side tests of libfastipc.so: the library loads without a GPU, exports
every symbol that include/fastipc.h declares, and its host-compiled math (the
same fipc_math.h the kernels run) agrees with the oracle.  The scalar
drop-ins use the reference's by-value / double* calling convention."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fastipc_amd import _lib
from oracle import lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "fastipc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(\w+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    so = _lib.load()
    names = declared_symbols()
    assert len(names) > 30
    for n in names:
        assert hasattr(so, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_status_without_gpu():
    from tests.conftest import HAS_GPU
    if HAS_GPU:
        pytest.skip("a GPU is present")
    ctx = C.c_void_p()
    assert _lib.load().fipc_create(C.byref(ctx)) == 2  # FIPC_ERR_HIP


@pytest.mark.parametrize("n", [3, 6, 9, 12])
def test_project_pd_dropin_by_value(n):
    rng = np.random.default_rng(n)
    so = _lib.load()
    for _ in range(20):
        A = rng.normal(size=(n, n))
        A = A + A.T
        outs = [C.c_double() for _ in range(n * n)]
        getattr(so, f"project_pd_{n}")(*[float(v) for v in A.ravel()], *[C.byref(o) for o in outs])
        got = np.array([o.value for o in outs]).reshape(n, n)
        assert np.abs(got - O.make_pd(A)).max() <= 1e-12 * np.abs(A).max()


def test_array_project_pd_including_2x2():
    rng = np.random.default_rng(2)
    for n in (2, 3, 6, 9, 12):
        A = rng.normal(size=(n, n))
        A = np.ascontiguousarray(A + A.T)
        out = np.zeros_like(A)
        _lib.load().fipc_project_pd(n, _lib.dptr(A), _lib.dptr(out))
        assert np.abs(out - O.make_pd(A)).max() <= 1e-12 * np.abs(A).max()


def test_ccd_dropins_vs_oracle_and_reference():
    REF = np.load(os.path.join(ROOT, "tests", "golden", "reference.npz"))
    so = _lib.load()
    for name, fn in (("pt", so.point_triangle_ccd), ("ee", so.edge_edge_ccd)):
        P, D, d2, T = (REF[f"ccd_{name}_{s}"] for s in ("pts", "disp", "d2", "toc"))
        for i in range(0, len(P), 3):
            out = C.c_double()
            fn(*[float(v) for v in P[i].ravel()], *[float(v) for v in D[i].ravel()], 0.2, float(d2[i]), C.byref(out))
            orc = (O.pt_ccd if name == "pt" else O.ee_ccd)(*P[i], *D[i], 0.2, d2[i])
            assert abs(out.value - orc) <= 1e-12
            assert abs(out.value - T[i]) <= 1e-6 and out.value <= T[i] + 1e-9


def test_root_dropins():
    REF = np.load(os.path.join(ROOT, "tests", "golden", "reference.npz"))
    so = _lib.load()
    for c, r in zip(REF["cubic_in"], REF["cubic_out"]):
        out = C.c_double()
        so.get_smallest_positive_real_cubic_root(*[float(v) for v in c], 1e-8, C.byref(out))
        assert (out.value < 0) == (r < 0)
        if r >= 0:
            assert out.value == pytest.approx(r, rel=1e-6, abs=1e-9)
    out = C.c_double()
    so.get_smallest_positive_real_quad_root(1.0, -3.0, 2.0, 1e-8, C.byref(out))
    assert out.value == pytest.approx(1.0)


def test_svd3_dropin():
    rng = np.random.default_rng(5)
    so = _lib.load()
    for _ in range(50):
        F = rng.normal(size=(3, 3))
        o = [C.c_double() for _ in range(27)]
        so.svd_3(*[float(v) for v in F.ravel()], *[C.byref(v) for v in o])
        v = np.array([x.value for x in o])
        U, S, V = v[:9].reshape(3, 3), v[9:18].reshape(3, 3), v[18:].reshape(3, 3)
        assert np.abs(U @ S @ V.T - F).max() < 1e-12 * np.abs(F).max() * 10
        assert np.linalg.det(U) == pytest.approx(1.0) and np.linalg.det(V) == pytest.approx(1.0)
        s = np.abs(np.diag(S))
        assert np.all(np.diff(s[:2]) <= 1e-12)


@pytest.mark.parametrize("kind", range(7))
def test_host_constraint_math_vs_oracle(kind):
    AD = np.load(os.path.join(ROOT, "tests", "golden", "autodiff.npz"))
    X, X0 = AD[f"c{kind}_x"], AD[f"c{kind}_x0"]
    h, k = float(AD["c_dhat2"]), float(AD["c_kappa"])
    n = (6, 9, 12, 12, 12, 12, 12)[kind]
    for i in range(len(X)):
        for project in (0, 1):
            E = C.c_double()
            g = np.zeros(n)
            H = np.zeros((n, n))
            x = np.ascontiguousarray(X[i].ravel())
            x0 = np.ascontiguousarray(X0[i].ravel())
            _lib.call("fipc_host_constraint", kind, _lib.dptr(x), _lib.dptr(x0), h, k, project, C.byref(E),
                      _lib.dptr(g), _lib.dptr(H))
            Eo, go, Ho = O.constraint_gH(kind, X[i], X0[i], h, k, project=bool(project))
            assert E.value == pytest.approx(Eo, rel=1e-11)
            assert np.abs(g - go).max() <= 1e-10 * np.abs(go).max()
            assert np.abs(H - Ho).max() <= 1e-9 * np.abs(Ho).max()


def test_host_tet_math_vs_oracle():
    rng = np.random.default_rng(8)
    L = O.lib()
    for i in range(60):
        F = np.eye(3) + rng.normal(size=(3, 3)) * 0.3
        if np.linalg.det(F) < 0.2:
            continue
        Dm = np.eye(3) * 0.1 + rng.normal(size=(3, 3)) * 0.01
        IB = np.ascontiguousarray(np.linalg.inv(Dm))
        la, mu, dt = 3e5, 2e5, 0.01
        vol = np.linalg.det(Dm) / 6
        psi = C.c_double()
        P = np.zeros(9)
        H = np.zeros(144)
        _lib.call("fipc_host_tet", _lib.dptr(np.ascontiguousarray(F)), _lib.dptr(IB), dt * dt * vol, la, mu,
                  C.byref(psi), _lib.dptr(P), _lib.dptr(H))
        x = np.zeros((4, 3))
        x[1:] = (F @ Dm).T
        g12 = np.zeros(12)
        H12 = np.zeros(144)
        L.orc_tet_gH(O.dptr(x), O.iptr(np.arange(4, dtype=np.int32)), O.dptr(np.ascontiguousarray(Dm)), la, mu, dt,
                     O.dptr(g12), O.dptr(H12))
        assert np.abs(H - H12).max() <= 1e-9 * np.abs(H12).max()
        assert np.abs(P.reshape(3, 3) - O.nh_P(F, la, mu)).max() <= 1e-10 * np.abs(P).max()


def _dense():
    return np.load(os.path.join(ROOT, "tests", "golden", "reference_dense.npz"))


@pytest.mark.parametrize("n", [2, 3, 4, 6, 9])
def test_solve_dropins_vs_reference(n):
    """solve_N against the reference's own compiled wrapper.cpp (goldens):
    SPD inputs to rounding; indefinite inputs (D clamped at 1e-8, so the
    solved matrix has a condition number near 1e8 |D|) to that conditioning.
    The reference keeps LDL^T's pivot permutation in the solved matrix, so
    for SPD inputs x differs from F^-1 rhs whenever the pivoting reorders."""
    d = _dense()
    so = _lib.load()
    A, B, X = d[f"solve{n}_A"], d[f"solve{n}_b"], d[f"solve{n}_x"]
    differs = 0
    for k, (a, b, x) in enumerate(zip(A, B, X)):
        outs = [C.c_double() for _ in range(n)]
        getattr(so, f"solve_{n}")(*[float(v) for v in a.ravel()], *[float(v) for v in b], *[C.byref(o) for o in outs])
        got = np.array([o.value for o in outs])
        rel = np.abs(got - x).max() / np.abs(x).max()
        if k < 30:  # SPD half
            assert rel <= 1e-12, (k, rel)
            differs += np.abs(np.linalg.solve(a, b) - x).max() > 1e-6 * np.abs(x).max()
        else:
            assert rel <= 1e-2, (k, rel)
    if n >= 3:
        assert differs > 0  # the permutation quirk is exercised


def test_svd2_dropin_vs_reference():
    d = _dense()
    so = _lib.load()
    for F, U, S, V in zip(d["svd2_F"], d["svd2_U"], d["svd2_S"], d["svd2_V"]):
        outs = [C.c_double() for _ in range(12)]
        so.svd_2(*[float(v) for v in F.ravel()], *[C.byref(o) for o in outs])
        v = np.array([o.value for o in outs])
        u, s, w = v[:4].reshape(2, 2), v[4:8].reshape(2, 2), v[8:].reshape(2, 2)
        assert np.abs(u - U).max() <= 1e-14 and np.abs(s - S).max() <= 1e-13 and np.abs(w - V).max() <= 1e-14
        # and what the mathematics fixes: A = U S V^T with rotations, sorted
        assert np.abs(u @ s @ w.T - F).max() <= 1e-13 * max(1.0, np.abs(F).max())
        assert np.linalg.det(u) == pytest.approx(1.0) and np.linalg.det(w) == pytest.approx(1.0)
        assert s[0, 0] >= s[1, 1] and s[0, 0] >= abs(s[1, 1]) - 1e-12


@pytest.mark.parametrize("n", [2, 4])
def test_project_pd_2_4_dropins_vs_reference(n):
    d = _dense()
    so = _lib.load()
    for a, o in zip(d[f"pd{n}_in"], d[f"pd{n}_out"]):
        outs = [C.c_double() for _ in range(n * n)]
        getattr(so, f"project_pd_{n}")(*[float(v) for v in a.ravel()], *[C.byref(x) for x in outs])
        got = np.array([x.value for x in outs]).reshape(n, n)
        assert np.abs(got - o).max() <= 1e-12 * np.abs(a).max()
        # PSD and equal to the eigen-clamp of the leading negative run (eigh is unique for it)
        assert np.linalg.eigvalsh(got).min() >= -1e-12 * np.abs(a).max()
