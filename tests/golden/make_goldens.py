"""Writes the golden fixtures that pin the oracle.

    python tests/golden/make_goldens.py

* autodiff.npz -- torch float64 autodiff (CPU) of the squared distances
  PP/PE/PT/EE and |(a1-a0)x(b1-b0)|^2 (common/math/distance.py), of the seven
  barrier-composed constraint energies b(d(x)) [* m(x)] (common/math/ipc.py)
  and of the neo-Hookean energy Psi(F) (common/physics/neo_hookean.py:5-16):
  inputs, values, gradients, Hessians.  The formulas here are the plain
  definitions; derivatives come from torch, independent of the oracle's jets.
* reference_dense.npz (or alone with --dense) -- solve_N, svd_2, svd_3 and
  project_pd_{2,4} of the same native library, on seeded SPD / indefinite /
  diagonal / rank-deficient inputs;
* reference_ccd_scenes.npz (--scene-ccd) -- the reference's CCD step bounds
  on the candidate pairs of stacked-cube and slab states (scene_ccd_set);
* reference.npz -- outputs of the reference's own native library
  (common/math/wrapper/wrapper.cpp + EVCTCD, built by oracle/build_ref.sh into
  oracle/_ref/libref.so): project_pd_{3,6,9,12}, the cubic root and the
  point-triangle / edge-edge CCD on seeded random inputs.
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
torch.set_default_dtype(torch.float64)


def d_pp(x):
    a, b = x[0:3], x[3:6]
    return ((a - b) ** 2).sum()


def d_pe(x):
    p, e0, e1 = x[0:3], x[3:6], x[6:9]
    return (torch.linalg.cross(e0 - p, e1 - p) ** 2).sum() / ((e1 - e0) ** 2).sum()


def d_pt(x):
    p, t0, t1, t2 = x[0:3], x[3:6], x[6:9], x[9:12]
    n = torch.linalg.cross(t1 - t0, t2 - t0)
    return ((p - t0) @ n) ** 2 / (n @ n)


def d_ee(x):
    a0, a1, b0, b1 = x[0:3], x[3:6], x[6:9], x[9:12]
    n = torch.linalg.cross(a1 - a0, b1 - b0)
    return ((b0 - a0) @ n) ** 2 / (n @ n)


def d_cn(x):
    a0, a1, b0, b1 = x[0:3], x[3:6], x[6:9], x[9:12]
    n = torch.linalg.cross(a1 - a0, b1 - b0)
    return n @ n


DIST = [(d_pp, 6), (d_pe, 9), (d_pt, 12), (d_ee, 12), (d_cn, 12)]


def barrier(d, h, k):
    return -k * (d - h) ** 2 * torch.log(d / h)


def constraint_energy(kind, x, x0, h, k):
    if kind == 0:
        d = d_pp(x)
    elif kind == 1:
        d = d_pe(x)
    elif kind == 2:
        d = d_pt(x)
    elif kind in (3, 4):
        d = d_ee(x)
    elif kind == 5:
        d = d_pp(torch.cat([x[0:3], x[6:9]]))
    else:
        d = d_pe(torch.cat([x[0:3], x[6:12]]))
    b = barrier(d, h, k)
    if kind >= 4:
        a0, a1, b0, b1 = x0[0:3], x0[3:6], x0[6:9], x0[9:12]
        eps = 1e-3 * ((a0 - a1) ** 2).sum() * ((b0 - b1) ** 2).sum()
        c = d_cn(x)
        r = c / eps
        b = b * ((-r + 2.0) * r)
    return b


def psi(F, la, mu):
    J = torch.det(F)
    lj = torch.log(J)
    return mu / 2 * ((F * F).sum() - 3) - mu * lj + la / 2 * lj * lj


def autodiff_set(rng):
    out = {}
    for i, (fn, n) in enumerate(DIST):
        X = rng.uniform(-1, 1, size=(40, n))
        E, G, H = [], [], []
        for x in X:
            t = torch.tensor(x)
            E.append(fn(t).item())
            G.append(torch.autograd.functional.jacobian(fn, t).numpy())
            H.append(torch.autograd.functional.hessian(fn, t).numpy())
        out[f"dist{i}_x"] = X
        out[f"dist{i}_E"] = np.array(E)
        out[f"dist{i}_g"] = np.array(G)
        out[f"dist{i}_H"] = np.array(H)
    h, k = 1e-3, 1.0
    for kind in range(7):
        npts = (2, 3, 4, 4, 4, 4, 4)[kind]
        xs, x0s, E, G, H = [], [], [], [], []
        while len(xs) < 24:
            x = rng.uniform(-1, 1, size=(npts, 3)) * 0.05
            x0 = x.copy()
            if kind >= 4:
                d = rng.normal(size=3)
                x[1] = x[0] + d * 0.05
                x[3] = x[2] + (d + rng.normal(size=3) * 0.02) * 0.05
                x0 = x.copy()
                x0[3] = x0[2] + (d + rng.normal(size=3) * 0.3) * 0.05
            t, t0 = torch.tensor(x.ravel()), torch.tensor(x0.ravel())
            f = lambda z: constraint_energy(kind, z, t0, h, k)
            e = f(t).item()
            if kind >= 4:
                eps = 1e-3 * np.sum((x0[0] - x0[1]) ** 2) * np.sum((x0[2] - x0[3]) ** 2)
                c = np.sum(np.cross(x[1] - x[0], x[3] - x[2]) ** 2)
                if not c < eps:
                    continue
            dd = [d_pp, d_pe, d_pt, d_ee, d_ee, None, None][kind]
            if not (e > 0) or not np.isfinite(e):
                continue
            xs.append(x)
            x0s.append(x0)
            E.append(e)
            G.append(torch.autograd.functional.jacobian(f, t).numpy())
            H.append(torch.autograd.functional.hessian(f, t).numpy())
        out[f"c{kind}_x"] = np.array(xs)
        out[f"c{kind}_x0"] = np.array(x0s)
        out[f"c{kind}_E"] = np.array(E)
        out[f"c{kind}_g"] = np.array(G)
        out[f"c{kind}_H"] = np.array(H)
    out["c_dhat2"] = np.array(h)
    out["c_kappa"] = np.array(k)
    # neo-Hookean under pure stretch (singular values in [1.02, 1.15]): every
    # 2x2 twist block and the singular-value Hessian are positive, so the
    # energy Hessian is positive definite and equals the projected dP/dF
    def rot(r):
        q, _ = np.linalg.qr(r.normal(size=(3, 3)))
        return q * np.sign(np.linalg.det(q))
    Fs = np.array([rot(rng) @ np.diag(rng.uniform(1.02, 1.15, size=3)) @ rot(rng).T for _ in range(40)])
    la, mu = 2.0e5, 1.0e5
    P, HH, PS = [], [], []
    for F in Fs:
        t = torch.tensor(F)
        f = lambda z: psi(z.reshape(3, 3), la, mu)
        PS.append(f(t.ravel()).item())
        P.append(torch.autograd.functional.jacobian(f, t.ravel()).numpy().reshape(3, 3))
        HH.append(torch.autograd.functional.hessian(f, t.ravel()).numpy())
    out["nh_F"] = Fs
    out["nh_psi"] = np.array(PS)
    out["nh_P"] = np.array(P)
    out["nh_H_rowvec"] = np.array(HH)  # Hessian w.r.t. row-major vec(F)
    out["nh_la"] = np.array(la)
    out["nh_mu"] = np.array(mu)
    return out


def reference_set(rng):
    from oracle import lib as O
    if O.ref() is None:
        print("oracle/_ref/libref.so not built; run oracle/build_ref.sh", file=sys.stderr)
        return None
    out = {}
    for n in (3, 6, 9, 12):
        A = rng.normal(size=(50, n, n))
        A = A + np.transpose(A, (0, 2, 1))
        out[f"pd{n}_in"] = A
        out[f"pd{n}_out"] = np.array([O.ref_project_pd(a) for a in A])
    C = rng.normal(size=(2000, 4))
    C[::10, 0] *= 1e-9
    out["cubic_in"] = C
    out["cubic_out"] = np.array([O.ref_cubic_root(*c) for c in C])
    for kind, name in ((0, "pt"), (1, "ee")):
        pts = rng.uniform(0, 1, size=(3000, 4, 3))
        disp = rng.normal(size=(3000, 4, 3)) * rng.uniform(0.1, 1.5, size=(3000, 1, 1))
        d2 = np.zeros(3000)
        toc = np.zeros(3000)
        L = O.lib()
        for i in range(3000):
            P = [O.dptr(np.ascontiguousarray(v)) for v in pts[i]]
            if kind == 0:
                d2[i] = L.orc_PT_dist2(*P, O.pt_type(*pts[i]))
                toc[i] = O.ref_pt_ccd(*pts[i], *disp[i], 0.2, d2[i])
            else:
                d2[i] = L.orc_EE_dist2(*P, O.ee_type(*pts[i]))
                toc[i] = O.ref_ee_ccd(*pts[i], *disp[i], 0.2, d2[i])
        out[f"ccd_{name}_pts"] = pts
        out[f"ccd_{name}_disp"] = disp
        out[f"ccd_{name}_d2"] = d2
        out[f"ccd_{name}_toc"] = toc
    return out


def dense_set(rng):
    """solve_N, svd_2, project_pd_{2,4} and svd_3 of the reference's native library."""
    from oracle import lib as O
    if O.ref() is None:
        return None
    out = {}
    for n in (2, 3, 4, 6, 9):
        A = rng.normal(size=(60, n, n))
        A = A + np.transpose(A, (0, 2, 1))
        # half of them SPD, half indefinite (the clamp of D at 1e-8 matters)
        A[:30] = np.einsum("kij,klj->kil", A[:30], A[:30]) + 0.1 * np.eye(n)
        b = rng.normal(size=(60, n))
        out[f"solve{n}_A"] = A
        out[f"solve{n}_b"] = b
        out[f"solve{n}_x"] = np.array([O.ref_solve(a, bb) for a, bb in zip(A, b)])
    F = rng.normal(size=(200, 2, 2))
    F[::7, 0, 1] = 0.0
    F[::7, 1, 0] = 0.0
    F[::11] = np.abs(F[::11]) * np.array([[1, 0], [0, 1]])
    out["svd2_F"] = F
    r = [O.ref_svd2(f) for f in F]
    out["svd2_U"] = np.array([u for u, _, _ in r])
    out["svd2_S"] = np.array([s for _, s, _ in r])
    out["svd2_V"] = np.array([v for _, _, v in r])
    for n in (2, 4):
        A = rng.normal(size=(60, n, n))
        A = A + np.transpose(A, (0, 2, 1))
        out[f"pd{n}_in"] = A
        out[f"pd{n}_out"] = np.array([O.ref_project_pd(a) for a in A])
    # svd_3: general, diagonal (signs), rank-deficient and repeated-value inputs
    F = list(rng.normal(size=(200, 3, 3)))
    F += [np.diag(rng.normal(size=3)) for _ in range(20)]
    F += [rng.normal(size=(3, 2)) @ rng.normal(size=(2, 3)) for _ in range(20)]
    for _ in range(20):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        F.append(Q @ np.diag([2.0, 2.0, 1.0]) @ R)
    F = np.array(F)
    out["svd3_F"] = F
    r = [O.ref_svd3(f) for f in F]
    out["svd3_U"] = np.array([u for u, _, _ in r])
    out["svd3_S"] = np.array([s for _, s, _ in r])
    out["svd3_V"] = np.array([v for _, _, v in r])
    return out


def _ccd_candidates(o, p, eta=0.2):
    """All (point, triangle) and (edge i < j) pairs of an oracle state that
    reach the CCD narrow phase for direction p, by brute force: index
    filter, the swept boxes overlapping within dHat (ipc.py:339-345,
    :357-363) and the lower bound (1 - eta) d / maxDisp below 1
    (wrapper.cpp:254-328).  Returns (kind, points, displacements, d^2)."""
    from oracle import lib as O
    L = O.lib()
    x, dh = o.x, o.dHat
    bp, be, bt = o.bp, o.be, o.bt
    out = []
    lo_p, hi_p = np.minimum(x[bp], x[bp] + p[bp]), np.maximum(x[bp], x[bp] + p[bp])
    xt, pt = x[bt], p[bt]
    lo_t, hi_t = np.minimum(xt, xt + pt).min(1), np.maximum(xt, xt + pt).max(1)
    for i in range(len(bp)):
        ok = np.all(lo_p[i] < hi_t + dh, axis=1) & np.all(lo_t - dh < hi_p[i], axis=1) & ~np.any(bt == bp[i], axis=1)
        for j in np.flatnonzero(ok):
            ids = [bp[i], *bt[j]]
            P, D = x[ids], p[ids]
            md = np.sqrt(D[0] @ D[0]) + np.sqrt(max(D[1] @ D[1], D[2] @ D[2], D[3] @ D[3]))
            if md == 0:
                continue
            d2 = L.orc_PT_dist2(*[O.dptr(np.ascontiguousarray(v)) for v in P], O.pt_type(*P))
            if (1 - eta) * np.sqrt(d2) / md < 1:
                out.append((0, P, D, d2))
    xe, pe = x[be], p[be]
    lo_e, hi_e = np.minimum(xe, xe + pe).min(1), np.maximum(xe, xe + pe).max(1)
    for i in range(len(be)):
        a0, a1 = be[i]
        ok = np.all(lo_e[i] < hi_e + dh, axis=1) & np.all(lo_e - dh < hi_e[i], axis=1)
        ok[:i + 1] = False
        ok &= ~(np.isin(be[:, 0], [a0, a1]) | np.isin(be[:, 1], [a0, a1]))
        for j in np.flatnonzero(ok):
            ids = [a0, a1, *be[j]]
            P, D = x[ids], p[ids]
            md = np.sqrt(max(D[0] @ D[0], D[1] @ D[1])) + np.sqrt(max(D[2] @ D[2], D[3] @ D[3]))
            if md == 0:
                continue
            d2 = L.orc_EE_dist2(*[O.dptr(np.ascontiguousarray(v)) for v in P], O.ee_type(*P))
            if (1 - eta) * np.sqrt(d2) / md < 1:
                out.append((1, P, D, d2))
    return out


def scene_ccd_set(rng):
    """Golden step bounds of the reference's compiled CCD for the pairs real
    scene states send to the narrow phase: the stacked cubes (two 10^3
    cubes, steps 8 and 20: before contact the cubes fall rigidly on the
    lattice, so many edge pairs are exactly parallel) and the 10x16x10 slab
    between plates (step 15), each with the oracle's Newton direction scaled
    1, 30 and 300.  Kept: every pair whose strict-IEEE and fused-multiply-add
    evaluations differ by more than 1e-9, or where either differs from the
    reference by more than that, and a seeded sample of 600 others per state
    and scale."""
    from fastipc_amd import scene as S
    from oracle import lib as O
    from oracle.newton import OracleIPC
    L = O.lib()
    import ctypes as C
    rows = {k: [] for k in ("kind", "pts", "disp", "d2", "toc", "state")}
    for label, sc, steps in (("cubes", S.two_cubes(10), (8, 20)), ("slab", S.compressed_slab(10, 16, 10), (15,))):
        o = OracleIPC(sc)
        for k in range(max(steps) + 1):
            if k in steps:
                o.compute_xn_and_xTilde(o.frame * o.dt)
                o.find_constraints()
                o.assemble()
                o.solve_system(o.frame * o.dt)
                p0 = o.p.copy()
                for si, scale in enumerate((1.0, 30.0, 300.0)):
                    cand = _ccd_candidates(o, p0 * scale)
                    ref = np.array([(O.ref_pt_ccd if kk == 0 else O.ref_ee_ccd)(*P, *D, 0.2, d2) for kk, P, D, d2 in cand])
                    one = np.array([[O.ccd_one(kk, f, P, D, 0.2, d2) for f in (0, 1)] for kk, P, D, d2 in cand])
                    keep = np.flatnonzero((np.abs(ref[:, None] - one) > 1e-9).any(axis=1) | (np.abs(one[:, 0] - one[:, 1]) > 1e-9))
                    rest = np.setdiff1d(np.arange(len(cand)), keep)
                    keep = np.concatenate([keep, rng.choice(rest, min(600, len(rest)), replace=False)])
                    for i in np.sort(keep):
                        kk, P, D, d2 = cand[i]
                        rows["kind"].append(kk)
                        rows["pts"].append(P)
                        rows["disp"].append(D)
                        rows["d2"].append(d2)
                        rows["toc"].append(ref[i])
                        rows["state"].append(f"{label}:{k}:{scale:g}")
                    print(f"{label} step {k} scale {scale:g}: {len(cand)} pairs, kept {len(keep)}", flush=True)
            o.step()
    return {"kind": np.array(rows["kind"], np.int32), "pts": np.array(rows["pts"]), "disp": np.array(rows["disp"]),
            "d2": np.array(rows["d2"]), "toc": np.array(rows["toc"]), "state": np.array(rows["state"])}


if __name__ == "__main__" and "--scene-ccd" in sys.argv:
    np.savez_compressed(os.path.join(HERE, "reference_ccd_scenes.npz"), **scene_ccd_set(np.random.default_rng(31)))
    print("wrote reference_ccd_scenes.npz")
    sys.exit(0)

if __name__ == "__main__" and "--dense" in sys.argv:
    np.savez_compressed(os.path.join(HERE, "reference_dense.npz"), **dense_set(np.random.default_rng(777)))
    print("wrote reference_dense.npz")
    sys.exit(0)

if __name__ == "__main__":
    rng = np.random.default_rng(20261202)
    np.savez_compressed(os.path.join(HERE, "autodiff.npz"), **autodiff_set(rng))
    ref = reference_set(np.random.default_rng(424242))
    if ref is not None:
        np.savez_compressed(os.path.join(HERE, "reference.npz"), **ref)
    dense = dense_set(np.random.default_rng(777))
    if dense is not None:
        np.savez_compressed(os.path.join(HERE, "reference_dense.npz"), **dense)
    print("wrote goldens")
