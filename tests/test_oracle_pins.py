"""Pins the C oracle before it is trusted as the parity checker.

* squared distances, constraint energies and neo-Hookean: torch float64
  autodiff fixtures (tests/golden/autodiff.npz, made by make_goldens.py);
* makePD, cubic root, CCD: outputs of the reference's own native library
  (tests/golden/reference.npz, compiled from the reference's wrapper.cpp and
  EVCTCD by oracle/build_ref.sh) and closed-form checks (numpy eigh / roots);
* closest-feature classification: brute-force minimum distance;
* constraint search: brute-force enumeration of all pairs;
* tet gradient / Hessian: autodiff of the tet energy.
"""
import os

import numpy as np
import pytest

from oracle import lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
AD = np.load(os.path.join(HERE, "golden", "autodiff.npz"))
REF = np.load(os.path.join(HERE, "golden", "reference.npz"))


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def test_barrier_closed_form_and_fd():
    h, k = 1e-3, 2.0
    for d in (1e-7, 1e-5, 3e-4, 9.99e-4):
        E, g, H = O.barrier(d, h, k)
        assert E == pytest.approx(-k * (d - h) ** 2 * np.log(d / h), rel=1e-14)
        # finite differences (the tests/TEST.py:8-19 pattern, asserted)
        s = d * 1e-6
        Ep, gp, _ = O.barrier(d + s, h, k)
        Em, gm, _ = O.barrier(d - s, h, k)
        assert (Ep - Em) / (2 * s) == pytest.approx(g, rel=1e-6)
        assert (gp - gm) / (2 * s) == pytest.approx(H, rel=1e-6)
    assert O.barrier(h, h, k) == (0.0, 0.0, 0.0)
    assert O.barrier(2 * h, h, k) == (0.0, 0.0, 0.0)


@pytest.mark.parametrize("kind", range(5))
def test_distance_derivatives_vs_autodiff(kind):
    X, E, G, H = (AD[f"dist{kind}_{s}"] for s in "xEgH")
    for i in range(len(X)):
        e, g, h = O.dist_gH(kind, X[i])
        assert e == pytest.approx(E[i], rel=1e-12)
        assert rel(g, G[i]) < 1e-11
        assert rel(h, H[i]) < 1e-10
        assert O.dist_E(kind, X[i]) == pytest.approx(E[i], rel=1e-12)


@pytest.mark.parametrize("kind", range(7))
def test_constraint_energy_gradient_hessian_vs_autodiff(kind):
    X, X0, E, G, H = (AD[f"c{kind}_{s}"] for s in ("x", "x0", "E", "g", "H"))
    h, k = float(AD["c_dhat2"]), float(AD["c_kappa"])
    assert len(X) > 10
    for i in range(len(X)):
        e, g, hh = O.constraint_gH(kind, X[i], X0[i], h, k, project=False)
        assert e == pytest.approx(E[i], rel=1e-10)
        assert O.constraint_E(kind, X[i], X0[i], h, k) == pytest.approx(E[i], rel=1e-10)
        assert rel(g, G[i]) < 1e-9
        assert rel(hh, H[i]) < 1e-8


def _colvec_perm():
    # row-major vec index r*3+c  ->  column-major index c*3+r
    return np.array([(i % 3) * 3 + i // 3 for i in range(9)])


def test_neohookean_vs_autodiff():
    F, P, H = AD["nh_F"], AD["nh_P"], AD["nh_H_rowvec"]
    la, mu = float(AD["nh_la"]), float(AD["nh_mu"])
    perm = _colvec_perm()
    for i in range(len(F)):
        assert rel(O.nh_P(F[i], la, mu), P[i]) < 1e-10
        U, s, V = O.svd3(F[i])
        psi = O.lib().orc_nh_energy(O.dptr(np.ascontiguousarray(s)), la, mu)
        assert psi == pytest.approx(AD["nh_psi"][i], rel=1e-8, abs=1e-6)
        # pure stretch: the energy Hessian is PD, so the projected dPdF equals it
        Hc = np.empty((9, 9))
        Hr = H[i]
        for a in range(9):
            for b in range(9):
                Hc[perm[a], perm[b]] = Hr[a, b]
        assert rel(O.nh_dPdF(F[i], la, mu), Hc) < 1e-8


def test_neohookean_projection_properties():
    rng = np.random.default_rng(3)
    for _ in range(50):
        F = np.eye(3) + rng.normal(size=(3, 3)) * 0.6
        if np.linalg.det(F) < 0.1:
            continue
        la, mu = 4e5, 1e5
        D = O.nh_dPdF(F, la, mu)
        assert np.abs(D - D.T).max() <= 1e-9 * np.abs(D).max()
        assert np.linalg.eigvalsh(D).min() >= -1e-7 * np.abs(D).max()
        # rotation invariance: dPdF(RF) = (I x R) dPdF(F) (I x R)^T in column-major vec
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        R = q * np.sign(np.linalg.det(q))
        K = np.kron(np.eye(3), R)
        assert rel(O.nh_dPdF(R @ F, la, mu), K @ D @ K.T) < 1e-8


@pytest.mark.parametrize("n", [3, 6, 9, 12])
def test_make_pd_vs_reference_and_eigh(n):
    A, B = REF[f"pd{n}_in"], REF[f"pd{n}_out"]
    for a, b in zip(A, B):
        o = O.make_pd(a)
        assert np.abs(o - b).max() <= 1e-12 * np.abs(a).max()
        w, V = np.linalg.eigh(a)
        assert np.abs(o - (V * np.maximum(w, 0)) @ V.T).max() <= 1e-12 * np.abs(a).max()


def test_make_pd2_closed_form():
    # math_tools.py:39-64: unchanged when PSD, zero when NSD, and on the
    # positive eigenvector the value L1md^2/L1 (not the exact projection)
    rng = np.random.default_rng(4)
    for _ in range(200):
        a, b, d = rng.normal(size=3)
        M = np.array([[a, b], [b, d]])
        w, V = np.linalg.eigh(M)
        out = O.make_pd(M)
        if w.min() >= 0:
            assert np.allclose(out, M, rtol=0, atol=1e-15)
        elif w.max() <= 0:
            assert np.all(out == 0)
        else:
            L1 = w.max()
            assert out[0, 0] == pytest.approx((L1 - d) ** 2 / L1, rel=1e-9)
            assert out[1, 1] == pytest.approx(b * b / L1, rel=1e-9)


def test_cubic_root_vs_reference_and_numpy():
    C, R = REF["cubic_in"], REF["cubic_out"]
    for c, r in zip(C, R):
        o = O.cubic_root(*c)
        assert (o < 0) == (r < 0)
        if r >= 0:
            assert o == pytest.approx(r, rel=1e-6, abs=1e-9)
            # smallest positive real root of a t^3 + b t^2 + c t + d
            if abs(c[0]) > 1e-8:
                roots = np.roots(c)
                pos = [z.real for z in roots if abs(z.imag) < 1e-6 and z.real > 0]
                if pos:
                    assert o == pytest.approx(min(pos), rel=1e-5, abs=1e-9)


def _det_cubic(xs, ps):
    """Coefficients of det[x2 - x1 + t (p2 - p1), ...] by sampling the cubic
    at four t and solving the Vandermonde system (independent of the kernel's
    cross-product expansion)."""
    ts = np.array([0.0, 0.5, 1.0, 2.0])
    def det(t):
        y = [xs[i] + t * ps[i] for i in range(4)]
        return np.linalg.det(np.stack([y[1] - y[0], y[2] - y[0], y[3] - y[0]]))
    v = np.array([det(t) for t in ts])
    a, b, c, d = np.linalg.solve(np.vander(ts, 4), v)
    return a, b, c, d


def test_inversion_certificate_is_sound():
    """Tets that the inversion kernel skips (2 (|a| T^3 + |b| T^2 + |c| T) < d,
    T = 1 / (1 - eta)) have an oracle bound of at least 1, so skipping them
    cannot change the minimum, which starts at 1."""
    rng = np.random.default_rng(11)
    eta, T = 0.2, 1.25
    cleared = kept = 0
    for trial in range(3000):
        xs = [np.zeros(3), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])]
        xs = [x + rng.normal(size=3) * 0.1 for x in xs]
        scale = 10.0 ** rng.uniform(-3, 0.3)
        ps = [rng.normal(size=3) * scale for _ in range(4)]
        a, b, c, d = _det_cubic(xs, ps)
        clear = d > 0 and 2 * (abs(a) * T ** 3 + abs(b) * T ** 2 + abs(c) * T) < d
        al = O.tet_inversion_alpha(xs, ps, eta)
        if clear:
            cleared += 1
            assert al >= 1.0, (trial, a, b, c, d, al)
        else:
            kept += 1
    # both branches are exercised, and small steps are cleared
    assert cleared > 300 and kept > 300, (cleared, kept)


def test_inversion_bound_value():
    """The bound's value: 0.8 times the smallest positive real root of the
    cubic t -> det[D_s + t dD_s] (coefficients fitted by _det_cubic from four
    determinants, roots by numpy.roots), or 1 when there is none
    (math_tools.py:160-181 with wrapper.cpp:220-252).  A swapped, negated or
    dropped coefficient moves the root, so such a mistake in the oracle's
    expansion fails here.  Tets whose roots are numerically ambiguous for the
    1e-8 imaginary-part test (near-double roots, tiny imaginary parts) and
    those on the reference's |a| <= 1e-8 quadratic branch are skipped."""
    rng = np.random.default_rng(12)
    eta = 0.2
    checked = hits = skipped = 0
    for trial in range(3000):
        xs = [np.zeros(3), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])]
        xs = [x + rng.normal(size=3) * 0.1 for x in xs]
        scale = 10.0 ** rng.uniform(-1, 0.5)
        ps = [rng.normal(size=3) * scale for _ in range(4)]
        a, b, c, d = _det_cubic(xs, ps)
        if abs(a) <= 1e-6:
            skipped += 1
            continue
        roots = np.roots([a, b, c, d])
        mag = max(1.0, np.abs(roots).max())
        im = np.abs(roots.imag)
        sep = min(abs(roots[i] - roots[j]) for i in range(3) for j in range(i + 1, 3))
        if np.any((im > 1e-12 * mag) & (im < 1e-5 * mag)) or sep < 1e-4 * mag:
            skipped += 1
            continue
        pos = [z.real for z in roots if abs(z.imag) <= 1e-12 * mag and z.real > 1e-9]
        al = O.tet_inversion_alpha(xs, ps, eta)
        if pos:
            want = (1.0 - eta) * min(pos)
            assert al == pytest.approx(want, rel=1e-9), (trial, (a, b, c, d), al, want)
            hits += 1
        else:
            assert al == 1.0, (trial, (a, b, c, d), al)
        checked += 1
    assert checked > 2500 and hits > 500, (checked, hits, skipped)


@pytest.mark.parametrize("name,fn", [("pt", "pt_ccd"), ("ee", "ee_ccd")])
def test_ccd_vs_reference(name, fn):
    P, D, d2, T = (REF[f"ccd_{name}_{s}"] for s in ("pts", "disp", "d2", "toc"))
    f = getattr(O, fn)
    out = np.array([f(*P[i], *D[i], 0.2, d2[i]) for i in range(len(P))])
    assert (T < 1).sum() > 100
    assert np.abs(out - T).max() <= 1e-6
    # never later than the reference beyond root-finding precision
    assert np.all(out <= T + 1e-9)


def test_ccd_invariants():
    rng = np.random.default_rng(9)
    for _ in range(300):
        pts = rng.uniform(0, 1, size=(4, 3))
        disp = rng.normal(size=(4, 3)) * 0.8
        d2 = O.lib().orc_PT_dist2(*[O.dptr(np.ascontiguousarray(v)) for v in pts], O.pt_type(*pts))
        toc = O.pt_ccd(*pts, *disp, 0.2, d2)
        maxd = np.linalg.norm(disp[0]) + np.linalg.norm(disp[1:], axis=1).max()
        lb = 0.8 * np.sqrt(d2) / maxd
        assert toc == 1.0 or lb - 1e-12 <= toc <= 1.0
        # along the trajectory before toc the point stays off the triangle's
        # plane interior (distance to the triangle never reaches zero)
        for t in np.linspace(0, toc, 20)[:-1]:
            y = pts + t * disp
            d = O.lib().orc_PT_dist2(*[O.dptr(np.ascontiguousarray(v)) for v in y], O.pt_type(*y))
            assert d > 0.0


def _seg_seg_d2(a0, a1, b0, b1, n=301):
    s = np.linspace(0, 1, n)
    A = a0[None] + s[:, None] * (a1 - a0)
    B = b0[None] + s[:, None] * (b1 - b0)
    return ((A[:, None] - B[None]) ** 2).sum(-1).min()


def _pt_tri_d2(p, t0, t1, t2, n=201):
    u, v = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    m = u + v <= 1
    Q = t0 + u[m, None] * (t1 - t0) + v[m, None] * (t2 - t0)
    return ((Q - p) ** 2).sum(-1).min()


def test_closest_feature_types_vs_bruteforce():
    rng = np.random.default_rng(12)
    L = O.lib()
    for _ in range(60):
        p, t0, t1, t2 = rng.uniform(-1, 1, size=(4, 3))
        c = O.pt_type(p, t0, t1, t2)
        d = L.orc_PT_dist2(*[O.dptr(np.ascontiguousarray(v)) for v in (p, t0, t1, t2)], c)
        bf = _pt_tri_d2(p, t0, t1, t2)
        assert d <= bf + 1e-12 and d == pytest.approx(bf, rel=2e-2, abs=1e-4)
        a0, a1, b0, b1 = rng.uniform(-1, 1, size=(4, 3))
        c = O.ee_type(a0, a1, b0, b1)
        d = L.orc_EE_dist2(*[O.dptr(np.ascontiguousarray(v)) for v in (a0, a1, b0, b1)], c)
        bf = _seg_seg_d2(a0, a1, b0, b1)
        assert d <= bf + 1e-12 and d == pytest.approx(bf, rel=2e-2, abs=1e-4)
    # special cases: interior, edge and vertex regions of the triangle
    t0, t1, t2 = np.array([0., 0, 0]), np.array([1., 0, 0]), np.array([0., 1, 0])
    assert O.pt_type(np.array([0.2, 0.2, 0.1]), t0, t1, t2) == 6
    assert O.pt_type(np.array([0.5, -0.3, 0.1]), t0, t1, t2) == 3
    assert O.pt_type(np.array([0.6, 0.6, 0.1]), t0, t1, t2) == 4
    assert O.pt_type(np.array([-0.3, 0.5, 0.1]), t0, t1, t2) == 5
    assert O.pt_type(np.array([-0.3, -0.3, 0.1]), t0, t1, t2) == 0
    assert O.pt_type(np.array([1.3, -0.1, 0.1]), t0, t1, t2) == 1
    assert O.pt_type(np.array([-0.1, 1.3, 0.1]), t0, t1, t2) == 2
    # parallel edges overlapping: reduced to a point-edge case (distance.py:1730)
    a0, a1 = np.array([0., 0, 0]), np.array([1., 0, 0])
    b0, b1 = np.array([0.2, 0.1, 0]), np.array([1.4, 0.1, 0])
    assert O.ee_type(a0, a1, b0, b1) in (2, 5, 6, 7)


def test_find_constraints_vs_bruteforce():
    from fastipc_amd import scene as S
    from oracle.newton import OracleIPC
    sc = S.two_cubes(4)
    o = OracleIPC(sc)
    rng = np.random.default_rng(0)
    # bring the cubes and the ground within dHat of each other
    x = o.x.copy()
    free = ~o.dmask
    x[free, 1] -= x[free, 1].min() - 0.4 * o.dHat
    x[free] += rng.normal(size=x[free].shape) * 1e-4
    o.x = np.ascontiguousarray(x)
    o.find_constraints()
    got = o.constraints()
    assert sum(len(v) for v in got.values()) > 10
    want = _brute(o)
    for k in got:
        assert np.array_equal(got[k], want[k]), k


def _brute(o):
    """All-pairs enumeration with the reference's attempt_PT / attempt_EE
    (NEWTON_IPC.py:181-312): hash-independent definition of the sets."""
    L = O.lib()
    x, x0, dH = o.x, o.x0, o.dHat
    dH2 = dH * dH
    rows = {k: [] for k in ("PP", "PE", "PT", "EE", "EEM", "PPM", "PEM")}
    P = lambda i: O.dptr(np.ascontiguousarray(x[i]))
    for p in o.bp:
        for t0, t1, t2 in o.bt:
            if p in (t0, t1, t2):
                continue
            T = x[[t0, t1, t2]]
            if not (np.all(x[p] < T.max(0) + dH) and np.all(T.min(0) - dH < x[p])):
                continue
            c = O.pt_type(x[p], x[t0], x[t1], x[t2])
            if not L.orc_PT_dist2(P(p), P(t0), P(t1), P(t2), c) < dH2:
                continue
            r = {0: ("PP", (p, t0)), 1: ("PP", (p, t1)), 2: ("PP", (p, t2)), 3: ("PE", (p, t0, t1)),
                 4: ("PE", (p, t1, t2)), 5: ("PE", (p, t2, t0)), 6: ("PT", (p, t0, t1, t2))}[c]
            rows[r[0]].append(r[1])
    be = o.be
    for i in range(len(be)):
        a0, a1 = be[i]
        for j in range(i + 1, len(be)):
            b0, b1 = be[j]
            if len({a0, a1, b0, b1}) < 4:
                continue
            A, B = x[[a0, a1]], x[[b0, b1]]
            if not (np.all(A.min(0) < B.max(0) + dH) and np.all(B.min(0) - dH < A.max(0))):
                continue
            c = O.ee_type(x[a0], x[a1], x[b0], x[b1])
            if not L.orc_EE_dist2(P(a0), P(a1), P(b0), P(b1), c) < dH2:
                continue
            moll = O.dist_E(4, x[[a0, a1, b0, b1]]) < 1e-3 * np.sum((x0[a0] - x0[a1]) ** 2) * np.sum((x0[b0] - x0[b1]) ** 2)
            tab = {0: (("PPM", (a0, a1, b0, b1)), ("PP", (a0, b0))), 1: (("PPM", (a0, a1, b1, b0)), ("PP", (a0, b1))),
                   3: (("PPM", (a1, a0, b0, b1)), ("PP", (a1, b0))), 4: (("PPM", (a1, a0, b1, b0)), ("PP", (a1, b1))),
                   2: (("PEM", (a0, a1, b0, b1)), ("PE", (a0, b0, b1))), 5: (("PEM", (a1, a0, b0, b1)), ("PE", (a1, b0, b1))),
                   6: (("PEM", (b0, b1, a0, a1)), ("PE", (b0, a0, a1))), 7: (("PEM", (b1, b0, a0, a1)), ("PE", (b1, a0, a1))),
                   8: (("EEM", (a0, a1, b0, b1)), ("EE", (a0, a1, b0, b1)))}[c]
            k, r = tab[0] if moll else tab[1]
            rows[k].append(r)
    w = {"PP": 2, "PE": 3}
    return {k: (np.unique(np.array(v, np.int32), axis=0) if v else np.zeros((0, w.get(k, 4)), np.int32))
            for k, v in rows.items()}


def test_tet_gradient_hessian_vs_autodiff():
    torch = pytest.importorskip("torch")
    torch.set_default_dtype(torch.float64)
    rng = np.random.default_rng(21)
    L = O.lib()
    for _ in range(10):
        Dm = np.eye(3) * 0.1 + rng.normal(size=(3, 3)) * 0.01
        # pure stretch: the tet Hessian is positive definite and unprojected
        q1, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        q2, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        q1 *= np.sign(np.linalg.det(q1))
        q2 *= np.sign(np.linalg.det(q2))
        F = q1 @ np.diag(rng.uniform(1.02, 1.15, size=3)) @ q2.T
        x = np.zeros((4, 3))
        x[1:] = (F @ Dm).T
        x += rng.normal(size=3)
        la, mu, dt = 3e5, 1e5, 0.01
        vol = np.linalg.det(Dm) / 6
        IB = torch.tensor(np.linalg.inv(Dm))

        def energy(z):
            z = z.reshape(4, 3)
            Ds = torch.stack([z[1] - z[0], z[2] - z[0], z[3] - z[0]], dim=1)
            Fz = Ds @ IB
            J = torch.det(Fz)
            lj = torch.log(J)
            return dt * dt * vol * (mu / 2 * ((Fz * Fz).sum() - 3) - mu * lj + la / 2 * lj * lj)

        t = torch.tensor(x.ravel())
        G = torch.autograd.functional.jacobian(energy, t).numpy()
        H = torch.autograd.functional.hessian(energy, t).numpy()
        g12 = np.zeros(12)
        H12 = np.zeros(144)
        L.orc_tet_gH(O.dptr(np.ascontiguousarray(x)), O.iptr(np.arange(4, dtype=np.int32)), O.dptr(np.ascontiguousarray(Dm)),
                     la, mu, dt, O.dptr(g12), O.dptr(H12))
        assert rel(g12, G) < 1e-9
        assert rel(H12.reshape(12, 12), H) < 1e-7


def test_reference_line_search_stall():
    """Why the reference's Newton loop can stall: at the stalled iterate of
    a cube-grid step (tests/golden/stall_local.npz, made by
    make_stall_fixture.py) the edge pair a = (1, 3), b = (4, 5) is closest
    at b's end point 5 (EE type 7), so the edge-edge search emits the
    point-edge row (5, 1, 3) (NEWTON_IPC.py:235-312) -- the same tuple the
    point-triangle search emits for point 5 and triangle (1, 3, 0), and
    np.unique (:348-356) keeps one.  A step of 1e-3 along the Newton
    direction moves the closest points inside both edges (type 8): the
    edge-edge row (1, 3, 4, 5) appears while the point-edge row stays (the
    triangle still yields it), so the barrier energy jumps up by a whole
    barrier term at an unchanged distance.  E > E0 then holds for every
    step the line search tries and alpha is halved without end."""
    F = np.load(os.path.join(HERE, "golden", "stall_local.npz"))
    x, p, x0 = F["x"], F["p"], F["x0"]
    tris, edges = F["tris"], F["edges"]
    dhat, kappa, inv_dx = float(F["dhat"]), float(F["kappa"]), float(F["inv_dx"])
    bp = np.arange(len(x), dtype=np.int32)
    L = O.lib()
    import ctypes as C

    def search(xx):
        cap = 64
        bufs = [np.zeros((cap, w), np.int32) for w in O.WIDTH]
        arr = (C.POINTER(C.c_int) * 7)(*[O.iptr(b) for b in bufs])
        caps = np.full(7, cap, np.int64)
        cnt = np.zeros(7, np.int64)
        xx = np.ascontiguousarray(xx)
        L.orc_find_constraints(O.dptr(xx), O.dptr(np.ascontiguousarray(x0)), len(bp), O.iptr(bp), len(edges),
                               O.iptr(np.ascontiguousarray(edges)), len(tris), O.iptr(np.ascontiguousarray(tris)),
                               dhat, inv_dx, arr, caps.ctypes.data_as(C.POINTER(C.c_int64)),
                               cnt.ctypes.data_as(C.POINTER(C.c_int64)))
        rows = {O.KINDS[t]: set(map(tuple, np.unique(bufs[t][:cnt[t]], axis=0))) for t in range(7)}
        E = sum(O.constraint_E(t, xx[list(r)], x0[list(r)], dhat * dhat, kappa)
                for t in range(7) for r in rows[O.KINDS[t]])
        return rows, E

    r0, E0 = search(x)
    r1, E1 = search(x + 1e-3 * p)
    pts = [np.ascontiguousarray(v) for v in x]
    assert O.ee_type(pts[1], pts[3], pts[4], pts[5]) == 7
    q = [np.ascontiguousarray(v) for v in x + 1e-3 * p]
    assert O.ee_type(q[1], q[3], q[4], q[5]) == 8
    assert (5, 1, 3) in r0["PE"] and (5, 1, 3) in r1["PE"]
    assert (1, 3, 4, 5) not in r0["EE"] and (1, 3, 4, 5) in r1["EE"]
    jump = O.constraint_E(3, x[[1, 3, 4, 5]], x0[[1, 3, 4, 5]], dhat * dhat, kappa)
    assert jump > 0
    # the energy of these rows rises by about one barrier term, far more
    # than the first-order change along the direction could explain
    assert E1 - E0 == pytest.approx(jump, rel=1e-2)


def test_ccd_scene_pairs_vs_reference():
    """The CCD step bounds of the pairs real scene states send to the narrow
    phase, against the reference's own compiled CCD (golden
    reference_ccd_scenes.npz): never later than the reference, except where
    the reference's hit is false (brute force: no contact).  Lattice-
    aligned edges falling rigidly (the stacked cubes before contact) make
    coefficients that vanish in exact arithmetic; the strict-IEEE rounding
    alone is then later than the reference on some pairs (what the
    fused-multiply-add twin, and so the minimum of the two, corrects), and
    the pairs outside 1e-6 are all of that kind and all earlier."""
    F = np.load(os.path.join(HERE, "golden", "reference_ccd_scenes.npz"))
    kind, P, D, d2, toc, state = F["kind"], F["pts"], F["disp"], F["d2"], F["toc"], F["state"]
    orc = np.array([(O.pt_ccd if k == 0 else O.ee_ccd)(*P[i], *D[i], 0.2, d2[i]) for i, k in enumerate(kind)])
    strict = np.array([O.ccd_one(k, 0, P[i], D[i], 0.2, d2[i]) for i, k in enumerate(kind)])
    assert len(toc) > 4000 and (toc < 1).sum() > 300
    later = np.flatnonzero(orc > toc + 1e-9)
    L = O.lib()
    for i in later:
        # later than the reference only where the reference reports a
        # contact that does not happen: the exact segment distance along
        # the whole step never comes near eta * d (a false hit of its root
        # finding, which our bound is not bound to repeat)
        assert kind[i] == 1 and orc[i] == 1.0
        eta_d = 0.2 * np.sqrt(d2[i])
        dmin = min(L.orc_EE_dist2(*[O.dptr(np.ascontiguousarray(v)) for v in P[i] + t * D[i]],
                                  O.ee_type(*(P[i] + t * D[i])))
                   for t in np.linspace(0.0, 1.0, 2001))
        assert np.sqrt(dmin) > 2 * eta_d, (i, np.sqrt(dmin), eta_d)
    assert len(later) <= 2
    outside = (np.abs(orc - toc) > 1e-6) & (orc <= toc)
    print(f"{len(toc)} pairs: {int(outside.sum())} earlier than the reference by more than 1e-6, "
          f"{len(later)} later (false hits of the reference); strict rounding alone later than the "
          f"reference on {int(np.sum(strict > toc + 1e-9))}")
    assert outside.sum() <= 50
    assert set(state[outside]) <= {"cubes:8:30", "cubes:8:300"}
    assert np.sum(strict > toc + 1e-9) > 0
