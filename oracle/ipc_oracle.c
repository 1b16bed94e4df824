/*
 * ipc_oracle.c -- CPU restatement (FP64, strict IEEE, OpenMP) of the FastIPC
 * Newton-IPC hot path.  TEST INFRASTRUCTURE ONLY: this library is the parity
 * checker and the CPU baseline; the product library never links or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates.
 * Derivatives of the squared distances are obtained with second-order forward
 * mode differentiation (jets), not with the reference's generated closed
 * forms, so the two derivations are independent.
 *
 * Pinning (tests/test_oracle_*.py):
 *   barrier_*          closed form + finite differences (tests/TEST.py:8-19 pattern)
 *   *_E / dist g,H     torch float64 autodiff fixtures in tests/golden
 *   make_pd (n>2)      compiled reference project_pd_{3,6,9,12} + numpy eigh
 *   make_pd2           eigen-decomposition closed form (numpy)
 *   neo-Hookean        Psi/P: invariants; dPdF: autodiff Hessian where PD,
 *                      symmetry, PSD-ness, rotation invariance
 *   cubic root         compiled reference + numpy.roots
 *   CCD                compiled reference point_triangle_ccd / edge_edge_ccd
 *   broad phase        brute-force enumeration on small scenes
 */
#include <math.h>
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* small vector helpers                                                      */
/* ------------------------------------------------------------------------ */
static inline void vsub(const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline void vadd(const double* a, const double* b, double* o) { o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2]; }
static inline double vdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void vcross(const double* a, const double* b, double* o)
{
    double t0 = a[1] * b[2] - a[2] * b[1];
    double t1 = a[2] * b[0] - a[0] * b[2];
    double t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0; o[1] = t1; o[2] = t2;
}
static inline double vnorm2(const double* a) { return vdot(a, a); }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* barrier  (common/math/distance.py:21-44)                                  */
/* ------------------------------------------------------------------------ */
ORC_API double orc_barrier_E(double d, double dHat2, double kappa)
{
    double E = 0.0;
    if (d < dHat2) E = -kappa * (d - dHat2) * (d - dHat2) * log(d / dHat2);
    return E;
}
ORC_API double orc_barrier_g(double d, double dHat2, double kappa)
{
    double g = 0.0, t2 = d - dHat2;
    if (d < dHat2) g = kappa * (t2 * log(d / dHat2) * -2.0 - (t2 * t2) / d);
    return g;
}
ORC_API double orc_barrier_H(double d, double dHat2, double kappa)
{
    double H = 0.0, t2 = d - dHat2;
    if (d < dHat2) H = kappa * ((log(d / dHat2) * -2.0 - t2 * 4.0 / d) + 1.0 / (d * d) * (t2 * t2));
    return H;
}

/* ------------------------------------------------------------------------ */
/* second-order jets over up to 12 variables                                 */
/* ------------------------------------------------------------------------ */
#define JN 12
typedef struct { double v; double g[JN]; double h[JN][JN]; } jet;

static void j_var(jet* o, double v, int k) { memset(o, 0, sizeof(jet)); o->v = v; o->g[k] = 1.0; }
static void j_add(const jet* a, const jet* b, jet* o)
{
    jet r; r.v = a->v + b->v;
    for (int i = 0; i < JN; i++) { r.g[i] = a->g[i] + b->g[i]; for (int j = 0; j < JN; j++) r.h[i][j] = a->h[i][j] + b->h[i][j]; }
    *o = r;
}
static void j_sub(const jet* a, const jet* b, jet* o)
{
    jet r; r.v = a->v - b->v;
    for (int i = 0; i < JN; i++) { r.g[i] = a->g[i] - b->g[i]; for (int j = 0; j < JN; j++) r.h[i][j] = a->h[i][j] - b->h[i][j]; }
    *o = r;
}
static void j_mul(const jet* a, const jet* b, jet* o)
{
    jet r; r.v = a->v * b->v;
    for (int i = 0; i < JN; i++) {
        r.g[i] = a->v * b->g[i] + b->v * a->g[i];
        for (int j = 0; j < JN; j++)
            r.h[i][j] = a->v * b->h[i][j] + b->v * a->h[i][j] + a->g[i] * b->g[j] + b->g[i] * a->g[j];
    }
    *o = r;
}
static void j_div(const jet* a, const jet* b, jet* o)
{
    jet r; r.v = a->v / b->v;
    for (int i = 0; i < JN; i++) r.g[i] = (a->g[i] - r.v * b->g[i]) / b->v;
    for (int i = 0; i < JN; i++)
        for (int j = 0; j < JN; j++)
            r.h[i][j] = (a->h[i][j] - r.v * b->h[i][j] - r.g[i] * b->g[j] - b->g[i] * r.g[j]) / b->v;
    *o = r;
}
typedef struct { jet c[3]; } jvec;
static void jv_load(jvec* o, const double* p, int k0)
{
    for (int d = 0; d < 3; d++) j_var(&o->c[d], p[d], k0 + d);
}
static void jv_sub(const jvec* a, const jvec* b, jvec* o) { for (int d = 0; d < 3; d++) j_sub(&a->c[d], &b->c[d], &o->c[d]); }
static void jv_dot(const jvec* a, const jvec* b, jet* o)
{
    jet t0, t1, t2;
    j_mul(&a->c[0], &b->c[0], &t0); j_mul(&a->c[1], &b->c[1], &t1); j_mul(&a->c[2], &b->c[2], &t2);
    j_add(&t0, &t1, &t0); j_add(&t0, &t2, o);
}
static void jv_cross(const jvec* a, const jvec* b, jvec* o)
{
    jet p, q; jvec r;
    j_mul(&a->c[1], &b->c[2], &p); j_mul(&a->c[2], &b->c[1], &q); j_sub(&p, &q, &r.c[0]);
    j_mul(&a->c[2], &b->c[0], &p); j_mul(&a->c[0], &b->c[2], &q); j_sub(&p, &q, &r.c[1]);
    j_mul(&a->c[0], &b->c[1], &p); j_mul(&a->c[1], &b->c[0], &q); j_sub(&p, &q, &r.c[2]);
    *o = r;
}

/* ------------------------------------------------------------------------ */
/* squared distances (common/math/distance.py:954-979, :684-686, :191-196)   */
/* ------------------------------------------------------------------------ */
ORC_API double orc_PP_E(const double* a, const double* b) { double d[3]; vsub(a, b, d); return vnorm2(d); }
ORC_API double orc_PE_E(const double* p, const double* e0, const double* e1)
{
    double a[3], b[3], c[3], e[3];
    vsub(e0, p, a); vsub(e1, p, b); vcross(a, b, c); vsub(e1, e0, e);
    return vnorm2(c) / vnorm2(e);
}
ORC_API double orc_PT_E(const double* p, const double* t0, const double* t1, const double* t2)
{
    double a[3], b[3], n[3], d[3];
    vsub(t1, t0, a); vsub(t2, t0, b); vcross(a, b, n); vsub(p, t0, d);
    double aTb = vdot(d, n);
    return aTb * aTb / vnorm2(n);
}
ORC_API double orc_EE_E(const double* a0, const double* a1, const double* b0, const double* b1)
{
    double u[3], v[3], n[3], w[3];
    vsub(a1, a0, u); vsub(b1, b0, v); vcross(u, v, n); vsub(b0, a0, w);
    double aTb = vdot(w, n);
    return aTb * aTb / vnorm2(n);
}
/* |(a1-a0) x (b1-b0)|^2  (distance.py:1345-1346) */
ORC_API double orc_EECN2_E(const double* a0, const double* a1, const double* b0, const double* b1)
{
    double u[3], v[3], n[3];
    vsub(a1, a0, u); vsub(b1, b0, v); vcross(u, v, n);
    return vnorm2(n);
}
/* M_threshold (distance.py:1670-1672) */
ORC_API double orc_M_threshold(const double* a0, const double* a1, const double* b0, const double* b1)
{
    double u[3], v[3];
    vsub(a0, a1, u); vsub(b0, b1, v);
    return 1.0e-3 * vnorm2(u) * vnorm2(v);
}

/* jet versions: kind 0=PP(2 pts) 1=PE(3) 2=PT(4) 3=EE(4) 4=EECN2(4) */
static void dist_jet(int kind, const double* pts, jet* out)
{
    jvec P[4];
    int np = kind == 0 ? 2 : kind == 1 ? 3 : 4;
    for (int i = 0; i < np; i++) jv_load(&P[i], pts + 3 * i, 3 * i);
    jvec a, b, n;
    jet t, s;
    switch (kind) {
    case 0: jv_sub(&P[0], &P[1], &a); jv_dot(&a, &a, out); break;
    case 1: jv_sub(&P[1], &P[0], &a); jv_sub(&P[2], &P[0], &b); jv_cross(&a, &b, &n); jv_dot(&n, &n, &t);
            jv_sub(&P[2], &P[1], &a); jv_dot(&a, &a, &s); j_div(&t, &s, out); break;
    case 2: jv_sub(&P[2], &P[1], &a); jv_sub(&P[3], &P[1], &b); jv_cross(&a, &b, &n); jv_sub(&P[0], &P[1], &a);
            jv_dot(&a, &n, &t); j_mul(&t, &t, &t); jv_dot(&n, &n, &s); j_div(&t, &s, out); break;
    case 3: jv_sub(&P[1], &P[0], &a); jv_sub(&P[3], &P[2], &b); jv_cross(&a, &b, &n); jv_sub(&P[2], &P[0], &a);
            jv_dot(&a, &n, &t); j_mul(&t, &t, &t); jv_dot(&n, &n, &s); j_div(&t, &s, out); break;
    default: jv_sub(&P[1], &P[0], &a); jv_sub(&P[3], &P[2], &b); jv_cross(&a, &b, &n); jv_dot(&n, &n, out); break;
    }
}
/* value, gradient (3k) and Hessian (3k x 3k, row-major) of a squared distance */
ORC_API void orc_dist_gH(int kind, const double* pts, double* E, double* g, double* H)
{
    jet r;
    dist_jet(kind, pts, &r);
    int n = kind == 0 ? 6 : kind == 1 ? 9 : 12;
    *E = r.v;
    for (int i = 0; i < n; i++) { g[i] = r.g[i]; for (int j = 0; j < n; j++) H[i * n + j] = r.h[i][j]; }
}

/* ------------------------------------------------------------------------ */
/* closest-feature classification (distance.py:1675-1767)                    */
/* ------------------------------------------------------------------------ */
static void pt_param(const double* e, const double* nvec, const double* d, double* out)
{
    /* param = (B B^T)^{-1} B d with B rows [e, e x n]  (distance.py:1678-1679) */
    double r1[3];
    vcross(e, nvec, r1);
    double m00 = vdot(e, e), m01 = vdot(e, r1), m10 = vdot(r1, e), m11 = vdot(r1, r1);
    double b0 = vdot(e, d), b1 = vdot(r1, d);
    double inv = 1.0 / (m00 * m11 - m01 * m10);
    double i00 = m11 * inv, i01 = -m01 * inv, i10 = -m10 * inv, i11 = m00 * inv;
    out[0] = i00 * b0 + i01 * b1;
    out[1] = i10 * b0 + i11 * b1;
}
ORC_API int orc_PT_type(const double* p, const double* t0, const double* t1, const double* t2)
{
    double e0[3], e1[3], e2[3], f[3], nvec[3], d[3], p0[2], p1[2], p2[2];
    vsub(t1, t0, e0); vsub(t2, t0, f); vcross(e0, f, nvec);
    vsub(p, t0, d); pt_param(e0, nvec, d, p0);
    if (0.0 < p0[0] && p0[0] < 1.0 && 0.0 <= p0[1]) return 3;
    vsub(t2, t1, e1); vsub(p, t1, d); pt_param(e1, nvec, d, p1);
    if (0.0 < p1[0] && p1[0] < 1.0 && 0.0 <= p1[1]) return 4;
    vsub(t0, t2, e2); vsub(p, t2, d); pt_param(e2, nvec, d, p2);
    if (0.0 < p2[0] && p2[0] < 1.0 && 0.0 < p2[1]) return 5;
    if (p0[0] <= 0.0 && p2[0] >= 1.0) return 0;
    if (p1[0] <= 0.0 && p0[0] >= 1.0) return 1;
    if (p2[0] <= 0.0 && p1[0] >= 1.0) return 2;
    return 6;
}
ORC_API int orc_EE_type(const double* a0, const double* a1, const double* b0, const double* b1)
{
    double u[3], v[3], w[3], c[3];
    vsub(a1, a0, u); vsub(b1, b0, v); vsub(a0, b0, w);
    double a = vdot(u, u), b = vdot(u, v), cc = vdot(v, v), d = vdot(u, w), e = vdot(v, w);
    double D = a * cc - b * b, tD = D, sN = b * e - cc * d, tN = 0.0;
    int kase = 8;
    if (sN <= 0.0) { tN = e; tD = cc; kase = 2; }
    else if (sN >= D) { tN = e + b; tD = cc; kase = 5; }
    else {
        tN = a * e - b * d;
        vcross(u, v, c);
        if (0.0 < tN && tN < tD && (vdot(c, w) == 0.0 || vnorm2(c) < 1.0e-20 * a * cc)) {
            if (sN < D / 2) { tN = e; tD = cc; kase = 2; }
            else { tN = e + b; tD = cc; kase = 5; }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) kase = 0; else if (-d >= a) kase = 3; else kase = 6;
    } else if (tN >= tD) {
        if (-d + b <= 0.0) kase = 1; else if (-d + b >= a) kase = 4; else kase = 7;
    }
    return kase;
}
ORC_API int orc_PE_type(const double* p, const double* e0, const double* e1)
{
    double e[3], d[3];
    vsub(e1, e0, e); vsub(p, e0, d);
    double ratio = vdot(e, d) / vnorm2(e);
    if (ratio < 0) return 0;
    if (ratio > 1) return 1;
    return 2;
}
/* PT_dist2 / EE_dist2 (distance.py:1770-1811) */
ORC_API double orc_PT_dist2(const double* p, const double* t0, const double* t1, const double* t2, int c)
{
    switch (c) {
    case 0: return orc_PP_E(p, t0);
    case 1: return orc_PP_E(p, t1);
    case 2: return orc_PP_E(p, t2);
    case 3: return orc_PE_E(p, t0, t1);
    case 4: return orc_PE_E(p, t1, t2);
    case 5: return orc_PE_E(p, t2, t0);
    case 6: return orc_PT_E(p, t0, t1, t2);
    }
    return 0.0;
}
ORC_API double orc_EE_dist2(const double* a0, const double* a1, const double* b0, const double* b1, int c)
{
    switch (c) {
    case 0: return orc_PP_E(a0, b0);
    case 1: return orc_PP_E(a0, b1);
    case 2: return orc_PE_E(a0, b0, b1);
    case 3: return orc_PP_E(a1, b0);
    case 4: return orc_PP_E(a1, b1);
    case 5: return orc_PE_E(a1, b0, b1);
    case 6: return orc_PE_E(b0, a0, a1);
    case 7: return orc_PE_E(b1, a0, a1);
    case 8: return orc_EE_E(a0, a1, b0, b1);
    }
    return 0.0;
}

/* ------------------------------------------------------------------------ */
/* SPD projection                                                           */
/* ------------------------------------------------------------------------ */
/* cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major).
 * On return A holds the eigenvalues on its diagonal, V the eigenvectors
 * (columns).  Used by make_pd and the oracle SVD. */
static void jacobi_eig(int n, double* A, double* V)
{
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; sweep++) {
        int rotated = 0;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                double app = A[p * n + p], aqq = A[q * n + q];
                /* negligible off-diagonal entry: drop it */
                if (fabs(apq) <= 1e-18 * (fabs(app) + fabs(aqq)) || apq == 0.0) { A[p * n + q] = A[q * n + p] = 0.0; continue; }
                rotated = 1;
                double theta = (aqq - app) / (2.0 * apq);
                double t = fabs(theta) > 1e150 ? 0.5 / theta
                         : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
        if (!rotated) break;
    }
}
/* makePD (common/math/wrapper/UTILS.h:7-25): if the smallest eigenvalue is
 * negative, clamp every negative eigenvalue to zero and rebuild V D V^T. */
ORC_API void orc_make_pd(int n, double* M)
{
    double A[144], V[144], lam[12];
    memcpy(A, M, sizeof(double) * n * n);
    jacobi_eig(n, A, V);
    double lmin = 1e300;
    for (int i = 0; i < n; i++) { lam[i] = A[i * n + i]; lmin = dmin(lmin, lam[i]); }
    if (lmin >= 0) return;
    for (int i = 0; i < n; i++) if (lam[i] < 0) lam[i] = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += V[i * n + k] * lam[k] * V[j * n + k];
            M[i * n + j] = s;
        }
}
/* 2x2 closed form (common/math/math_tools.py:39-64) */
ORC_API void orc_make_pd2(double* M)
{
    double a = M[0], b = (M[1] + M[2]) / 2.0, d = M[3];
    double b2 = b * b, D = a * d - b2, T = (a + d) / 2.0;
    double sq = sqrt(T * T - D), L2 = T - sq;
    if (L2 < 0.0) {
        double L1 = T + sq;
        if (L1 <= 0.0) { M[0] = M[1] = M[2] = M[3] = 0.0; }
        else if (b2 == 0.0) { M[0] = L1; M[1] = M[2] = M[3] = 0.0; }
        else {
            double L1md = L1 - d, r = L1md / L1;
            M[0] = r * L1md; M[1] = b * r; M[2] = b * r; M[3] = b2 / L1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* per-constraint energy / gradient / Hessian (common/math/ipc.py:30-329)    */
/* kind: 0 PP, 1 PE, 2 PT, 3 EE, 4 EEM, 5 PPM, 6 PEM                        */
/* x: the constraint's points (2,3 or 4 x 3); x0: rest positions (mollified) */
/* ------------------------------------------------------------------------ */
static double mollifier(const double* x, double eps, double* g, double* H)
{
    /* M_E/M_g/M_H (distance.py:1641-1667) with EEM_E/g/H (:1624-1638) */
    jet c;
    dist_jet(4, x, &c);
    memset(g, 0, sizeof(double) * 12);
    memset(H, 0, sizeof(double) * 144);
    if (!(c.v < eps)) return 1.0;
    double r = c.v / eps;
    double m = (-r + 2.0) * r;
    double m1 = 2.0 / eps * (-c.v / eps + 1.0);
    double m2 = -2.0 / (eps * eps);
    for (int i = 0; i < 12; i++) {
        g[i] = m1 * c.g[i];
        for (int j = 0; j < 12; j++) H[i * 12 + j] = m1 * c.h[i][j] + m2 * c.g[i] * c.g[j];
    }
    return m;
}
ORC_API double orc_constraint_E(int kind, const double* x, const double* x0, double dHat2, double kappa)
{
    double d;
    switch (kind) {
    case 0: d = orc_PP_E(x, x + 3); break;
    case 1: d = orc_PE_E(x, x + 3, x + 6); break;
    case 2: d = orc_PT_E(x, x + 3, x + 6, x + 9); break;
    case 3: case 4: d = orc_EE_E(x, x + 3, x + 6, x + 9); break;
    case 5: d = orc_PP_E(x, x + 6); break;
    default: d = orc_PE_E(x, x + 6, x + 9); break;
    }
    double b = orc_barrier_E(d, dHat2, kappa);
    if (kind >= 4) {
        double eps = orc_M_threshold(x0, x0 + 3, x0 + 6, x0 + 9);
        double g[12], H[144];
        b *= mollifier(x, eps, g, H);
    }
    return b;
}
ORC_API int orc_constraint_dof(int kind) { return kind == 0 ? 6 : kind == 1 ? 9 : 12; }
/* g and (optionally projected) H; returns the energy */
ORC_API double orc_constraint_gH(int kind, const double* x, const double* x0, double dHat2, double kappa,
                                 int project, double* g, double* H)
{
    double pts[12];
    int dkind, map[12], nd;
    switch (kind) {
    case 0: dkind = 0; nd = 6; memcpy(pts, x, 6 * 8); break;
    case 1: dkind = 1; nd = 9; memcpy(pts, x, 9 * 8); break;
    case 2: dkind = 2; nd = 12; memcpy(pts, x, 12 * 8); break;
    case 3: case 4: dkind = 3; nd = 12; memcpy(pts, x, 12 * 8); break;
    case 5: dkind = 0; nd = 6; memcpy(pts, x, 3 * 8); memcpy(pts + 3, x + 6, 3 * 8); break;
    default: dkind = 1; nd = 9; memcpy(pts, x, 3 * 8); memcpy(pts + 3, x + 6, 6 * 8); break;
    }
    /* placement of the distance's dofs in 12-space (fill_vec, ipc.py:254,300) */
    for (int i = 0; i < nd; i++) map[i] = i;
    if (kind == 5) { int m[6] = {0, 1, 2, 6, 7, 8}; memcpy(map, m, sizeof m); }
    if (kind == 6) { int m[9] = {0, 1, 2, 6, 7, 8, 9, 10, 11}; memcpy(map, m, sizeof m); }
    jet dj;
    dist_jet(dkind, pts, &dj);
    double d = dj.v;
    double b = orc_barrier_E(d, dHat2, kappa), bg = orc_barrier_g(d, dHat2, kappa), bh = orc_barrier_H(d, dHat2, kappa);
    int n = kind >= 4 ? 12 : nd;
    double lg[12] = {0}, lH[144];
    memset(lH, 0, sizeof lH);
    for (int i = 0; i < nd; i++) {
        lg[map[i]] = bg * dj.g[i];
        for (int j = 0; j < nd; j++) lH[map[i] * n + map[j]] = bh * dj.g[i] * dj.g[j] + bg * dj.h[i][j];
    }
    double E = b;
    if (kind >= 4) {
        double eps = orc_M_threshold(x0, x0 + 3, x0 + 6, x0 + 9), mg[12], mH[144];
        double m = mollifier(x, eps, mg, mH);
        E = b * m;
        for (int i = 0; i < 12; i++) g[i] = lg[i] * m + b * mg[i];
        for (int i = 0; i < 12; i++)
            for (int j = 0; j < 12; j++)
                H[i * 12 + j] = lH[i * 12 + j] * m + lg[i] * mg[j] + mg[i] * lg[j] + b * mH[i * 12 + j];
    } else {
        for (int i = 0; i < n; i++) g[i] = lg[i];
        memcpy(H, lH, sizeof(double) * n * n);
    }
    if (project) orc_make_pd(n, H);
    return E;
}

/* ------------------------------------------------------------------------ */
/* neo-Hookean (common/physics/neo_hookean.py)                               */
/* ------------------------------------------------------------------------ */
static double det3(const double* F)
{
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
}
/* SVD via eigenvectors of F^T F, rotation-variant: det U = det V = 1 and the
 * last singular value carries the sign of det F. */
ORC_API void orc_svd3(const double* F, double* U, double* s, double* V)
{
    double A[9], W[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            A[i * 3 + j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j] + F[2 * 3 + i] * F[2 * 3 + j];
    jacobi_eig(3, A, W);
    double lam[3] = {A[0], A[4], A[8]};
    int idx[3] = {0, 1, 2};
    for (int i = 0; i < 3; i++) for (int j = i + 1; j < 3; j++) if (lam[idx[j]] > lam[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) V[i * 3 + k] = W[i * 3 + idx[k]];
    if (det3(V) < 0) for (int i = 0; i < 3; i++) V[i * 3 + 2] = -V[i * 3 + 2];
    double FV[9];
    for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) FV[i * 3 + k] = F[i * 3 + 0] * V[0 * 3 + k] + F[i * 3 + 1] * V[1 * 3 + k] + F[i * 3 + 2] * V[2 * 3 + k];
    double u0[3] = {FV[0], FV[3], FV[6]}, u1[3] = {FV[1], FV[4], FV[7]}, u2[3];
    s[0] = sqrt(vnorm2(u0));
    for (int d = 0; d < 3; d++) u0[d] /= s[0];
    double pr = vdot(u0, u1);
    for (int d = 0; d < 3; d++) u1[d] -= pr * u0[d];
    s[1] = sqrt(vnorm2(u1));
    if (s[1] > 0) for (int d = 0; d < 3; d++) u1[d] /= s[1];
    vcross(u0, u1, u2);
    s[1] = vdot(u1, (double[]){FV[1], FV[4], FV[7]});
    s[2] = vdot(u2, (double[]){FV[2], FV[5], FV[8]});
    for (int d = 0; d < 3; d++) { U[d * 3 + 0] = u0[d]; U[d * 3 + 1] = u1[d]; U[d * 3 + 2] = u2[d]; }
}
/* Psi(sigma) (neo_hookean.py:5-16) */
ORC_API double orc_nh_energy(const double* sig, double la, double mu)
{
    double s2 = sig[0] * sig[0] + sig[1] * sig[1] + sig[2] * sig[2];
    double lp = log(sig[0] * sig[1] * sig[2]);
    return mu / 2.0 * (s2 - 3) - (mu - la / 2.0 * lp) * lp;
}
/* P(F) (neo_hookean.py:61-67) */
ORC_API void orc_nh_P(const double* F, double la, double mu, double* P)
{
    double J = det3(F), logJ = log(J), scale = la * logJ - mu;
    double C[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                   F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                   F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
    for (int i = 0; i < 9; i++) P[i] = mu * F[i] + scale * (C[i] / J);
}
/* projected dP/dF, 9x9 in column-major vec ordering (neo_hookean.py:70-131) */
ORC_API void orc_nh_dPdF(const double* F, double la, double mu, double* dPdF)
{
    double U[9], s[3], V[9];
    orc_svd3(F, U, s, V);
    double prod = s[0] * s[1] * s[2], lp = log(prod), middle = mu - la * lp;
    double g[3], H[9];
    for (int i = 0; i < 3; i++) g[i] = mu * (s[i] - 1.0 / s[i]) + la * (1.0 / s[i]) * lp;
    for (int i = 0; i < 3; i++) {
        double inv2 = 1.0 / s[i] / s[i];
        H[i * 4] = mu * (1.0 + inv2) - la * inv2 * (lp - 1.0);
    }
    H[1] = H[3] = la / s[0] / s[1];
    H[5] = H[7] = la / s[1] / s[2];
    H[2] = H[6] = la / s[0] / s[2];
    orc_make_pd(3, H);
    double B[3][4];
    int pr[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (int k = 0; k < 3; k++) {
        int i = pr[k][0], j = pr[k][1];
        double left = (mu + middle / s[i] / s[j]) / 2.0;
        double right = (g[i] + g[j]) / (2 * dmax(s[i] + s[j], 0.000001));
        B[k][0] = left + right; B[k][1] = left - right; B[k][2] = left - right; B[k][3] = left + right;
        orc_make_pd2(B[k]);
    }
    /* 9x9 M in the (U col, V col) pair basis index a*3+b */
    double M[81];
    memset(M, 0, sizeof M);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) M[(i * 4) * 9 + j * 4] = H[i * 3 + j];
    M[1 * 9 + 1] = B[0][0]; M[1 * 9 + 3] = B[0][1]; M[3 * 9 + 1] = B[0][2]; M[3 * 9 + 3] = B[0][3];
    M[5 * 9 + 5] = B[1][0]; M[5 * 9 + 7] = B[1][1]; M[7 * 9 + 5] = B[1][2]; M[7 * 9 + 7] = B[1][3];
    M[2 * 9 + 2] = B[2][3]; M[2 * 9 + 6] = B[2][2]; M[6 * 9 + 2] = B[2][1]; M[6 * 9 + 6] = B[2][0];
    /* dPdF = sum_kl M[k][l] vec(u_a v_b^T) vec(u_c v_d^T)^T, vec column-major */
    double w[9][9];
    for (int k = 0; k < 9; k++) {
        int a = k / 3, b = k % 3;
        for (int j = 0; j < 3; j++) for (int i = 0; i < 3; i++) w[k][j * 3 + i] = U[i * 3 + a] * V[j * 3 + b];
    }
    for (int r = 0; r < 9; r++)
        for (int c = 0; c < 9; c++) {
            double acc = 0;
            for (int k = 0; k < 9; k++) for (int l = 0; l < 9; l++) if (M[k * 9 + l] != 0.0) acc += M[k * 9 + l] * w[k][r] * w[l][c];
            dPdF[r * 9 + c] = acc;
        }
}

/* ------------------------------------------------------------------------ */
/* roots  (common/math/wrapper/wrapper.cpp:200-252)                          */
/* ------------------------------------------------------------------------ */
ORC_API double orc_quad_root(double a, double b, double c, double tol)
{
    double t;
    if (fabs(a) <= tol) t = -c / b;
    else {
        double desc = b * b - 4 * a * c;
        if (desc > 0) {
            t = (-b - sqrt(desc)) / (2 * a);
            if (t < 0) t = (-b + sqrt(desc)) / (2 * a);
        } else t = -1;
    }
    return t;
}
ORC_API double orc_cubic_root(double a, double b, double c, double d, double tol)
{
    double t = -1;
    if (fabs(a) <= tol) return orc_quad_root(b, c, d, tol);
    double complex d0 = b * b - 3 * a * c;
    double complex d1 = 2 * b * b * b - 9 * a * b * c + 27 * a * a * d;
    double complex C = cpow((d1 + csqrt(d1 * d1 - 4.0 * d0 * d0 * d0)) / 2.0, 1.0 / 3.0);
    if (cabs(C) == 0.0) C = cpow((d1 - csqrt(d1 * d1 - 4.0 * d0 * d0 * d0)) / 2.0, 1.0 / 3.0);
    double complex u2 = (-1.0 + sqrt(3.0) * I) / 2.0, u3 = (-1.0 - sqrt(3.0) * I) / 2.0;
    double complex r1 = (b + C + d0 / C) / (-3.0 * a);
    double complex r2 = (b + u2 * C + d0 / (u2 * C)) / (-3.0 * a);
    double complex r3 = (b + u3 * C + d0 / (u3 * C)) / (-3.0 * a);
    if (fabs(cimag(r1)) < tol && creal(r1) > 0) t = creal(r1);
    if (fabs(cimag(r2)) < tol && creal(r2) > 0 && (creal(r2) < t || t < 0)) t = creal(r2);
    if (fabs(cimag(r3)) < tol && creal(r3) > 0 && (creal(r3) < t || t < 0)) t = creal(r3);
    return t;
}
/* inversion-free step bound for one tet (math_tools.py:160-181) */
ORC_API double orc_tet_inversion_alpha(const double* x1, const double* x2, const double* x3, const double* x4,
                                       const double* p1, const double* p2, const double* p3, const double* p4, double eta)
{
    double v01[3], v02[3], v03[3], p01[3], p02[3], p03[3], a1[3], a2[3], s[3], c0[3];
    vsub(x2, x1, v01); vsub(x3, x1, v02); vsub(x4, x1, v03);
    vsub(p2, p1, p01); vsub(p3, p1, p02); vsub(p4, p1, p03);
    double pp[3]; vcross(p01, p02, pp);
    vcross(v01, p02, a1); vcross(p01, v02, a2); vadd(a1, a2, s);
    vcross(v01, v02, c0);
    double a = vdot(p03, pp);
    double b = vdot(v03, pp) + vdot(p03, s);
    double c = vdot(p03, c0) + vdot(v03, s);
    double d = vdot(v03, c0);
    double r = orc_cubic_root(a, b, c, d, 1e-8);
    return r < 0.0 ? 1.0 : r * (1.0 - eta);
}

/* ------------------------------------------------------------------------ */
/* CCD: interval-based exact CCD restated from EVCTCD/CTCD.cpp:10-472       */
/* (ccd_body.h), evaluated twice: multiply-adds rounded twice (strict) and  */
/* once (fma).  On exactly degenerate pairs (parallel, lattice-aligned      */
/* edges moving rigidly) the coefficients vanish in exact arithmetic and    */
/* the interval decisions follow the rounding of the cancelling products;   */
/* the reference's build (README.md:15, -mfma) fuses them.  The step bound  */
/* is the smaller of the two.                                               */
/* ------------------------------------------------------------------------ */
typedef struct { double l, u; } ivl;
typedef struct { int n; ivl v[16]; } ivlset;
#define MAD(a, b, c) ((a) * (b) + (c))
#define CCDF(f) f##_strict
#include "ccd_body.h"
#undef MAD
#undef CCDF
#define MAD(a, b, c) fma((a), (b), (c))
#define CCDF(f) f##_fused
#include "ccd_body.h"
#undef MAD
#undef CCDF
/* point_triangle_ccd (wrapper.cpp:254-290) */
ORC_API double orc_pt_ccd(const double* p, const double* t0, const double* t1, const double* t2,
                          const double* dp, const double* dt0, const double* dt1, const double* dt2, double eta, double dist2)
{
    return dmin(pt_ccd_strict(p, t0, t1, t2, dp, dt0, dt1, dt2, eta, dist2),
                pt_ccd_fused(p, t0, t1, t2, dp, dt0, dt1, dt2, eta, dist2));
}
/* one rounding alone: fused != 0 the fma evaluation, else the strict one */
ORC_API double orc_ccd_one(int type, int fused, const double* a, const double* b, const double* c, const double* d,
                           const double* da, const double* db, const double* dc, const double* dd, double eta, double dist2)
{
    if (type == 0) return fused ? pt_ccd_fused(a, b, c, d, da, db, dc, dd, eta, dist2)
                                : pt_ccd_strict(a, b, c, d, da, db, dc, dd, eta, dist2);
    return fused ? ee_ccd_fused(a, b, c, d, da, db, dc, dd, eta, dist2) : ee_ccd_strict(a, b, c, d, da, db, dc, dd, eta, dist2);
}
/* edge_edge_ccd (wrapper.cpp:292-328) */
ORC_API double orc_ee_ccd(const double* a0, const double* a1, const double* b0, const double* b1,
                          const double* da0, const double* da1, const double* db0, const double* db1, double eta, double dist2)
{
    return dmin(ee_ccd_strict(a0, a1, b0, b1, da0, da1, db0, db1, eta, dist2),
                ee_ccd_fused(a0, a1, b0, b1, da0, da1, db0, db1, eta, dist2));
}
/* segment_intersect_triangle (common/math/ipc.py:511-526) */
ORC_API int orc_seg_tri(const double* P, const double* Q, const double* A, const double* B, const double* C)
{
    double R[3], E1[3], E2[3], N[3], AO[3], DAO[3];
    vsub(Q, P, R);
    double RLen = sqrt(vnorm2(R));
    for (int d = 0; d < 3; d++) R[d] /= RLen;
    vsub(B, A, E1); vsub(C, A, E2); vcross(E1, E2, N);
    double det = -vdot(R, N), invdet = 1.0 / det;
    vsub(P, A, AO); vcross(AO, R, DAO);
    double u = vdot(E2, DAO) * invdet, v = -vdot(E1, DAO) * invdet, t = vdot(AO, N) * invdet;
    return det >= 1e-12 && t >= 0.0 && u >= 0.0 && v >= 0.0 && (u + v) <= 1.0 && t <= RLen;
}

/* ------------------------------------------------------------------------ */
/* uniform-grid broad phase (NEWTON_IPC.py:97-113, 213-232, 313-334)         */
/* items are inserted into every cell their box overlaps; a query visits     */
/* every cell of its own box and tests a pair only in the first common cell  */
/* so that each pair is seen once.                                           */
/* ------------------------------------------------------------------------ */
typedef struct { int64_t key; int id; } cell_item;
typedef struct { cell_item* it; int64_t n; } grid_t;
#define CELL_BIAS (1 << 20)
static inline int64_t cell_key(int i, int j, int k)
{
    return (((int64_t)(i + CELL_BIAS)) << 42) | (((int64_t)(j + CELL_BIAS)) << 21) | (int64_t)(k + CELL_BIAS);
}
static int cmp_item(const void* a, const void* b)
{
    const cell_item* x = a; const cell_item* y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id - y->id;
}
static void box_cells(const double* lo, const double* hi, double inv, int* clo, int* chi)
{
    for (int d = 0; d < 3; d++) { clo[d] = (int)floor(lo[d] * inv); chi[d] = (int)floor(hi[d] * inv); }
}
static void grid_build(grid_t* g, int n, const double* lo, const double* hi, double inv)
{
    int64_t* cnt = malloc(sizeof(int64_t) * (n + 1));
    cnt[0] = 0;
    for (int i = 0; i < n; i++) {
        int a[3], b[3];
        box_cells(lo + 3 * i, hi + 3 * i, inv, a, b);
        cnt[i + 1] = cnt[i] + (int64_t)(b[0] - a[0] + 1) * (b[1] - a[1] + 1) * (b[2] - a[2] + 1);
    }
    g->n = cnt[n];
    g->it = malloc(sizeof(cell_item) * (g->n + 1));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        int a[3], b[3];
        box_cells(lo + 3 * i, hi + 3 * i, inv, a, b);
        int64_t o = cnt[i];
        for (int x = a[0]; x <= b[0]; x++) for (int y = a[1]; y <= b[1]; y++) for (int z = a[2]; z <= b[2]; z++) {
            g->it[o].key = cell_key(x, y, z); g->it[o].id = i; o++;
        }
    }
    qsort(g->it, g->n, sizeof(cell_item), cmp_item);
    free(cnt);
}
static int64_t grid_find(const grid_t* g, int64_t key)
{
    int64_t lo = 0, hi = g->n;
    while (lo < hi) { int64_t m = (lo + hi) / 2; if (g->it[m].key < key) lo = m + 1; else hi = m; }
    return lo;
}
/* callback-style enumeration of overlapping (query, item) pairs */
typedef void (*pair_fn)(void* ctx, int q, int item, int tid);
static void grid_query(const grid_t* g, int nq, const double* qlo, const double* qhi,
                       const double* ilo, const double* ihi, double inv, pair_fn fn, void* ctx)
{
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 64)
        for (int q = 0; q < nq; q++) {
            int a[3], b[3];
            box_cells(qlo + 3 * q, qhi + 3 * q, inv, a, b);
            for (int x = a[0]; x <= b[0]; x++) for (int y = a[1]; y <= b[1]; y++) for (int z = a[2]; z <= b[2]; z++) {
                int64_t key = cell_key(x, y, z);
                for (int64_t s = grid_find(g, key); s < g->n && g->it[s].key == key; s++) {
                    int id = g->it[s].id, c[3], e[3];
                    box_cells(ilo + 3 * id, ihi + 3 * id, inv, c, e);
                    /* first common cell */
                    if (x != (a[0] > c[0] ? a[0] : c[0]) || y != (a[1] > c[1] ? a[1] : c[1]) || z != (a[2] > c[2] ? a[2] : c[2])) continue;
                    fn(ctx, q, id, tid);
                }
            }
        }
    }
}

/* growable per-thread row buffers */
typedef struct { int* d; int64_t n, cap; } ibuf;
static void ib_push(ibuf* b, const int* v, int k)
{
    if (b->n + k > b->cap) { b->cap = (b->cap + k) * 2 + 64; b->d = realloc(b->d, sizeof(int) * b->cap); }
    memcpy(b->d + b->n, v, sizeof(int) * k);
    b->n += k;
}
#define MAXT 256
#define NTYPES 7
static const int type_width[NTYPES] = {2, 3, 4, 4, 4, 4, 4}; /* PP PE PT EE EEM PPM PEM */
typedef struct {
    const double *x, *x0;
    const int *bp, *be, *bt;
    double dHat, dHat2;
    ibuf rows[MAXT][NTYPES];
} fc_ctx;

/* attempt_PT (NEWTON_IPC.py:181-212) */
static void fc_pt(void* vc, int tri, int pi, int tid)
{
    fc_ctx* c = vc;
    int p = c->bp[pi], t0 = c->bt[3 * tri], t1 = c->bt[3 * tri + 1], t2 = c->bt[3 * tri + 2];
    if (p == t0 || p == t1 || p == t2) return;
    const double *X = c->x, *P = X + 3 * p, *T0 = X + 3 * t0, *T1 = X + 3 * t1, *T2 = X + 3 * t2;
    for (int d = 0; d < 3; d++) {
        double mx = dmax(dmax(T0[d], T1[d]), T2[d]), mn = dmin(dmin(T0[d], T1[d]), T2[d]);
        if (!(P[d] < mx + c->dHat && mn - c->dHat < P[d])) return;
    }
    int kase = orc_PT_type(P, T0, T1, T2);
    double d2 = orc_PT_dist2(P, T0, T1, T2, kase);
    if (!(d2 < c->dHat2)) return;
    int r[4];
    switch (kase) {
    case 0: r[0] = p; r[1] = t0; ib_push(&c->rows[tid][0], r, 2); break;
    case 1: r[0] = p; r[1] = t1; ib_push(&c->rows[tid][0], r, 2); break;
    case 2: r[0] = p; r[1] = t2; ib_push(&c->rows[tid][0], r, 2); break;
    case 3: r[0] = p; r[1] = t0; r[2] = t1; ib_push(&c->rows[tid][1], r, 3); break;
    case 4: r[0] = p; r[1] = t1; r[2] = t2; ib_push(&c->rows[tid][1], r, 3); break;
    case 5: r[0] = p; r[1] = t2; r[2] = t0; ib_push(&c->rows[tid][1], r, 3); break;
    default: r[0] = p; r[1] = t0; r[2] = t1; r[3] = t2; ib_push(&c->rows[tid][2], r, 4); break;
    }
}
/* attempt_EE (NEWTON_IPC.py:235-312); query edge i, item edge j, keep i < j */
static void fc_ee(void* vc, int i, int j, int tid)
{
    fc_ctx* c = vc;
    if (!(i < j)) return;
    int a0 = c->be[2 * i], a1 = c->be[2 * i + 1], b0 = c->be[2 * j], b1 = c->be[2 * j + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
    const double *X = c->x, *A0 = X + 3 * a0, *A1 = X + 3 * a1, *B0 = X + 3 * b0, *B1 = X + 3 * b1;
    for (int d = 0; d < 3; d++) {
        double mxa = dmax(A0[d], A1[d]), mna = dmin(A0[d], A1[d]), mxb = dmax(B0[d], B1[d]), mnb = dmin(B0[d], B1[d]);
        if (!(mna < mxb + c->dHat && mnb - c->dHat < mxa)) return;
    }
    double cn2 = orc_EECN2_E(A0, A1, B0, B1);
    const double* Y = c->x0;
    double eps = orc_M_threshold(Y + 3 * a0, Y + 3 * a1, Y + 3 * b0, Y + 3 * b1);
    int kase = orc_EE_type(A0, A1, B0, B1);
    double d2 = orc_EE_dist2(A0, A1, B0, B1, kase);
    if (!(d2 < c->dHat2)) return;
    int moll = cn2 < eps;
    int r[4];
    static const int pp_near[5][2] = {{0, 0}, {0, 1}, {0, 0}, {1, 0}, {1, 1}};
    (void)pp_near;
    int ids[4] = {a0, a1, b0, b1};
    switch (kase) {
    case 0: case 1: case 3: case 4: {
        int ia = (kase == 3 || kase == 4) ? 1 : 0, ib = (kase == 1 || kase == 4) ? 1 : 0;
        if (moll) { r[0] = ids[ia]; r[1] = ids[1 - ia]; r[2] = ids[2 + ib]; r[3] = ids[3 - ib]; ib_push(&c->rows[tid][5], r, 4); }
        else { r[0] = ids[ia]; r[1] = ids[2 + ib]; ib_push(&c->rows[tid][0], r, 2); }
        break;
    }
    case 2: case 5: {
        int ia = kase == 5 ? 1 : 0;
        if (moll) { r[0] = ids[ia]; r[1] = ids[1 - ia]; r[2] = b0; r[3] = b1; ib_push(&c->rows[tid][6], r, 4); }
        else { r[0] = ids[ia]; r[1] = b0; r[2] = b1; ib_push(&c->rows[tid][1], r, 3); }
        break;
    }
    case 6: case 7: {
        int ib = kase == 7 ? 1 : 0;
        if (moll) { r[0] = ids[2 + ib]; r[1] = ids[3 - ib]; r[2] = a0; r[3] = a1; ib_push(&c->rows[tid][6], r, 4); }
        else { r[0] = ids[2 + ib]; r[1] = a0; r[2] = a1; ib_push(&c->rows[tid][1], r, 3); }
        break;
    }
    default:
        r[0] = a0; r[1] = a1; r[2] = b0; r[3] = b1;
        ib_push(&c->rows[tid][moll ? 4 : 3], r, 4);
        break;
    }
}
static int nthreads(void)
{
#ifdef _OPENMP
    int n = omp_get_max_threads();
    return n > MAXT ? MAXT : n;
#else
    return 1;
#endif
}
/* find_constraints (NEWTON_IPC.py:337-357) without the final np.unique, which
 * the Python driver applies.  rows_out[t] receives up to cap[t] rows of width
 * type_width[t]; counts[t] the number found (may exceed cap: then truncated). */
ORC_API void orc_find_constraints(const double* x, const double* x0, int nbp, const int* bp, int nbe, const int* be,
                                  int nbt, const int* bt, double dHat, double inv_dx,
                                  int** rows_out, const int64_t* cap, int64_t* counts)
{
#ifdef _OPENMP
    if (omp_get_max_threads() > MAXT) omp_set_num_threads(MAXT);
#endif
    fc_ctx* c = calloc(1, sizeof(fc_ctx));
    c->x = x; c->x0 = x0; c->bp = bp; c->be = be; c->bt = bt; c->dHat = dHat; c->dHat2 = dHat * dHat;
    int nmax = nbp > nbe ? nbp : nbe;
    nmax = nmax > nbt ? nmax : nbt;
    double* lo = malloc(sizeof(double) * 3 * nmax);
    double* hi = malloc(sizeof(double) * 3 * nmax);
    double* qlo = malloc(sizeof(double) * 3 * nmax);
    double* qhi = malloc(sizeof(double) * 3 * nmax);
    /* PT: points inserted with +-dHat, triangles query their box */
    for (int i = 0; i < nbp; i++) for (int d = 0; d < 3; d++) { lo[3 * i + d] = x[3 * bp[i] + d] - dHat; hi[3 * i + d] = x[3 * bp[i] + d] + dHat; }
    for (int i = 0; i < nbt; i++) for (int d = 0; d < 3; d++) {
        double a = x[3 * bt[3 * i] + d], b = x[3 * bt[3 * i + 1] + d], e = x[3 * bt[3 * i + 2] + d];
        qlo[3 * i + d] = dmin(dmin(a, b), e); qhi[3 * i + d] = dmax(dmax(a, b), e);
    }
    grid_t g;
    grid_build(&g, nbp, lo, hi, inv_dx);
    grid_query(&g, nbt, qlo, qhi, lo, hi, inv_dx, fc_pt, c);
    free(g.it);
    /* EE: edges inserted with their box, queried with box +- dHat */
    for (int i = 0; i < nbe; i++) for (int d = 0; d < 3; d++) {
        double a = x[3 * be[2 * i] + d], b = x[3 * be[2 * i + 1] + d];
        lo[3 * i + d] = dmin(a, b); hi[3 * i + d] = dmax(a, b);
        qlo[3 * i + d] = lo[3 * i + d] - dHat; qhi[3 * i + d] = hi[3 * i + d] + dHat;
    }
    grid_build(&g, nbe, lo, hi, inv_dx);
    grid_query(&g, nbe, qlo, qhi, lo, hi, inv_dx, fc_ee, c);
    free(g.it);
    for (int t = 0; t < NTYPES; t++) {
        int64_t n = 0;
        for (int th = 0; th < MAXT; th++) {
            ibuf* b = &c->rows[th][t];
            int64_t k = b->n / type_width[t];
            for (int64_t r = 0; r < k; r++, n++)
                if (n < cap[t]) memcpy(rows_out[t] + n * type_width[t], b->d + r * type_width[t], sizeof(int) * type_width[t]);
            free(b->d);
        }
        counts[t] = n;
    }
    free(lo); free(hi); free(qlo); free(qhi); free(c);
}

/* ------------------------------------------------------------------------ */
/* assembly (NEWTON_IPC.py:545-747) into COO triplets + rhs                  */
/* ------------------------------------------------------------------------ */
/* elastic: restT (3x3 row-major, rest D_m) per tet */
ORC_API void orc_tet_restT(int nt, const double* x, const int* tets, double* restT, double* vol)
{
    for (int e = 0; e < nt; e++) {
        const double* a = x + 3 * tets[4 * e];
        double* T = restT + 9 * e;
        for (int c = 1; c <= 3; c++) for (int r = 0; r < 3; r++) T[r * 3 + c - 1] = x[3 * tets[4 * e + c] + r] - a[r];
        vol[e] = det3(T) / 6.0;
    }
}
static void inv3(const double* A, double* B)
{
    double d = det3(A);
    B[0] = (A[4] * A[8] - A[5] * A[7]) / d; B[1] = (A[2] * A[7] - A[1] * A[8]) / d; B[2] = (A[1] * A[5] - A[2] * A[4]) / d;
    B[3] = (A[5] * A[6] - A[3] * A[8]) / d; B[4] = (A[0] * A[8] - A[2] * A[6]) / d; B[5] = (A[2] * A[3] - A[0] * A[5]) / d;
    B[6] = (A[3] * A[7] - A[4] * A[6]) / d; B[7] = (A[1] * A[6] - A[0] * A[7]) / d; B[8] = (A[0] * A[4] - A[1] * A[3]) / d;
}
static void tet_F(const double* x, const int* t, const double* IB, double* F)
{
    double Ds[9];
    for (int c = 1; c <= 3; c++) for (int r = 0; r < 3; r++) Ds[r * 3 + c - 1] = x[3 * t[c] + r] - x[3 * t[0] + r];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) F[i * 3 + j] = Ds[i * 3] * IB[j] + Ds[i * 3 + 1] * IB[3 + j] + Ds[i * 3 + 2] * IB[6 + j];
}
/* per-tet 12x12 Hessian (dofs vertex-major) and 12 gradient entries, following
 * the intermediate contraction of NEWTON_IPC.py:626-690 */
ORC_API void orc_tet_gH(const double* x, const int* t, const double* restT, double la, double mu, double dt,
                        double* g12, double* H144)
{
    double IB[9], F[9], dPdF[81], P[9];
    inv3(restT, IB);
    double vol0 = det3(restT) / 6.0;
    tet_F(x, t, IB, F);
    orc_nh_dPdF(F, la, mu, dPdF);
    orc_nh_P(F, la, mu, P);
    double sc = dt * dt * vol0;
    for (int i = 0; i < 81; i++) dPdF[i] *= sc;
    for (int i = 0; i < 9; i++) P[i] *= sc;
    double im[12][9];
    for (int c = 0; c < 9; c++) {
        for (int a = 0; a < 3; a++) for (int r = 0; r < 3; r++)
            im[3 + 3 * a + r][c] = IB[a * 3 + 0] * dPdF[r * 9 + c] + IB[a * 3 + 1] * dPdF[(3 + r) * 9 + c] + IB[a * 3 + 2] * dPdF[(6 + r) * 9 + c];
        for (int r = 0; r < 3; r++) im[r][c] = -im[3 + r][c] - im[6 + r][c] - im[9 + r][c];
    }
    for (int R = 0; R < 12; R++) {
        for (int a = 0; a < 3; a++) for (int r = 0; r < 3; r++)
            H144[R * 12 + 3 + 3 * a + r] = IB[a * 3 + 0] * im[R][r] + IB[a * 3 + 1] * im[R][3 + r] + IB[a * 3 + 2] * im[R][6 + r];
        for (int r = 0; r < 3; r++) H144[R * 12 + r] = -H144[R * 12 + 3 + r] - H144[R * 12 + 6 + r] - H144[R * 12 + 9 + r];
    }
    for (int a = 0; a < 3; a++) for (int r = 0; r < 3; r++)
        g12[3 + 3 * a + r] = IB[a * 3 + 0] * P[r * 3 + 0] + IB[a * 3 + 1] * P[r * 3 + 1] + IB[a * 3 + 2] * P[r * 3 + 2];
    for (int r = 0; r < 3; r++) g12[r] = -g12[3 + r] - g12[6 + r] - g12[9 + r];
}
/* elastic energy of one tet (NEWTON_IPC.py:521-525) */
ORC_API double orc_tet_E(const double* x, const int* t, const double* restT, double la, double mu, double dt)
{
    double IB[9], F[9], U[9], s[3], V[9];
    inv3(restT, IB);
    tet_F(x, t, IB, F);
    orc_svd3(F, U, s, V);
    return orc_nh_energy(s, la, mu) * dt * dt * (det3(restT) / 6.0);
}
/* full system: inertia, elasticity, seven contact types.  rows[t] with n[t]
 * rows each.  COO output sized by the caller (orc_coo_size). */
ORC_API int64_t orc_coo_size(int nv, int nt, const int64_t* n)
{
    int64_t s = (int64_t)nv * 3 + (int64_t)nt * 144;
    for (int t = 0; t < NTYPES; t++) { int k = type_width[t] == 2 ? 2 : type_width[t] == 3 ? 3 : 4; s += n[t] * 9 * k * k; }
    return s;
}
static void add_rhs(double* rhs, int i, double v)
{
#pragma omp atomic
    rhs[i] -= v;
}
ORC_API void orc_assemble(int nv, const double* x, const double* xt, const double* m, const double* x0,
                          int nt, const int* tets, const double* restT, const double* la, const double* mu, double dt,
                          const int* const* rows, const int64_t* n, double dHat2, double kappa,
                          int* ci, int* cj, double* cv, double* rhs)
{
    memset(rhs, 0, sizeof(double) * 3 * nv);
    int64_t base = 0;
    for (int i = 0; i < nv; i++) for (int d = 0; d < 3; d++) {
        int64_t c = base + i * 3 + d;
        ci[c] = i * 3 + d; cj[c] = i * 3 + d; cv[c] = m[i];
        rhs[i * 3 + d] -= m[i] * (x[3 * i + d] - xt[3 * i + d]);
    }
    base += (int64_t)nv * 3;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nt; e++) {
        double g[12], H[144];
        const int* t = tets + 4 * e;
        orc_tet_gH(x, t, restT + 9 * e, la[e], mu[e], dt, g, H);
        int64_t c = base + (int64_t)e * 144;
        for (int R = 0; R < 12; R++) for (int C = 0; C < 12; C++, c++) {
            ci[c] = t[R / 3] * 3 + R % 3; cj[c] = t[C / 3] * 3 + C % 3; cv[c] = H[R * 12 + C];
        }
        for (int R = 0; R < 12; R++) add_rhs(rhs, t[R / 3] * 3 + R % 3, g[R]);
    }
    base += (int64_t)nt * 144;
    for (int ty = 0; ty < NTYPES; ty++) {
        int w = type_width[ty], k = ty == 0 ? 2 : ty == 1 ? 3 : 4, nd = 3 * k;
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t r = 0; r < n[ty]; r++) {
            const int* id = rows[ty] + r * w;
            double xx[12], yy[12], g[12], H[144];
            for (int a = 0; a < k; a++) for (int d = 0; d < 3; d++) { xx[3 * a + d] = x[3 * id[a] + d]; yy[3 * a + d] = x0[3 * id[a] + d]; }
            orc_constraint_gH(ty, xx, yy, dHat2, kappa, 1, g, H);
            int64_t c = base + r * nd * nd;
            for (int R = 0; R < nd; R++) for (int C = 0; C < nd; C++, c++) {
                ci[c] = id[R / 3] * 3 + R % 3; cj[c] = id[C / 3] * 3 + C % 3; cv[c] = H[R * nd + C];
            }
            for (int R = 0; R < nd; R++) add_rhs(rhs, id[R / 3] * 3 + R % 3, g[R]);
        }
        base += n[ty] * nd * nd;
    }
}
/* total incremental potential (NEWTON_IPC.py:514-542); contributions are
 * summed per category in index order (deterministic) */
ORC_API double orc_energy(int nv, const double* x, const double* xt, const double* m, const double* x0,
                          int nt, const int* tets, const double* restT, const double* la, const double* mu, double dt,
                          const int* const* rows, const int64_t* n, double dHat2, double kappa, double* parts)
{
    double Ei = 0, Ee = 0, Eb = 0;
    for (int i = 0; i < nv; i++) {
        double d[3];
        vsub(x + 3 * i, xt + 3 * i, d);
        Ei += 0.5 * m[i] * vnorm2(d);
    }
    double* te = malloc(sizeof(double) * (nt > 0 ? nt : 1));
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nt; e++) te[e] = orc_tet_E(x, tets + 4 * e, restT + 9 * e, la[e], mu[e], dt);
    for (int e = 0; e < nt; e++) Ee += te[e];
    free(te);
    for (int ty = 0; ty < NTYPES; ty++) {
        int w = type_width[ty], k = ty == 0 ? 2 : ty == 1 ? 3 : 4;
        for (int64_t r = 0; r < n[ty]; r++) {
            const int* id = rows[ty] + r * w;
            double xx[12], yy[12];
            for (int a = 0; a < k; a++) for (int d = 0; d < 3; d++) { xx[3 * a + d] = x[3 * id[a] + d]; yy[3 * a + d] = x0[3 * id[a] + d]; }
            Eb += orc_constraint_E(ty, xx, yy, dHat2, kappa);
        }
    }
    if (parts) { parts[0] = Ei; parts[1] = Ee; parts[2] = Eb; }
    return Ei + Ee + Eb;
}

/* ------------------------------------------------------------------------ */
/* line-search filters (NEWTON_IPC.py:377-468) and collision check (:855-886)*/
/* ------------------------------------------------------------------------ */
typedef struct {
    const double *x, *p;
    const int *bp, *be, *bt;
    double dHat, eta;
    double amin[MAXT];
    int hit[MAXT];
} ccd_ctx;
static void ccd_pt(void* vc, int tri, int pi, int tid)
{
    ccd_ctx* c = vc;
    int p = c->bp[pi], t0 = c->bt[3 * tri], t1 = c->bt[3 * tri + 1], t2 = c->bt[3 * tri + 2];
    if (p == t0 || p == t1 || p == t2) return;
    const double *X = c->x, *D = c->p;
    for (int d = 0; d < 3; d++) {
        double pv[2] = {X[3 * p + d], X[3 * p + d] + D[3 * p + d]};
        double mxp = dmax(pv[0], pv[1]), mnp = dmin(pv[0], pv[1]);
        double mxt = -1e300, mnt = 1e300;
        int tv[3] = {t0, t1, t2};
        for (int k = 0; k < 3; k++) {
            double a = X[3 * tv[k] + d], b = a + D[3 * tv[k] + d];
            mxt = dmax(mxt, dmax(a, b)); mnt = dmin(mnt, dmin(a, b));
        }
        if (!(mnp < mxt + c->dHat && mnt - c->dHat < mxp)) return;
    }
    const double *P = X + 3 * p, *T0 = X + 3 * t0, *T1 = X + 3 * t1, *T2 = X + 3 * t2;
    double d2 = orc_PT_dist2(P, T0, T1, T2, orc_PT_type(P, T0, T1, T2));
    double a = orc_pt_ccd(P, T0, T1, T2, D + 3 * p, D + 3 * t0, D + 3 * t1, D + 3 * t2, c->eta, d2);
    if (a < c->amin[tid]) c->amin[tid] = a;
}
static void ccd_ee(void* vc, int i, int j, int tid)
{
    ccd_ctx* c = vc;
    if (!(i < j)) return;
    int a0 = c->be[2 * i], a1 = c->be[2 * i + 1], b0 = c->be[2 * j], b1 = c->be[2 * j + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
    const double *X = c->x, *D = c->p;
    for (int d = 0; d < 3; d++) {
        double va[4] = {X[3 * a0 + d], X[3 * a0 + d] + D[3 * a0 + d], X[3 * a1 + d], X[3 * a1 + d] + D[3 * a1 + d]};
        double vb[4] = {X[3 * b0 + d], X[3 * b0 + d] + D[3 * b0 + d], X[3 * b1 + d], X[3 * b1 + d] + D[3 * b1 + d]};
        double mxa = dmax(dmax(va[0], va[1]), dmax(va[2], va[3])), mna = dmin(dmin(va[0], va[1]), dmin(va[2], va[3]));
        double mxb = dmax(dmax(vb[0], vb[1]), dmax(vb[2], vb[3])), mnb = dmin(dmin(vb[0], vb[1]), dmin(vb[2], vb[3]));
        if (!(mna < mxb + c->dHat && mnb - c->dHat < mxa)) return;
    }
    const double *A0 = X + 3 * a0, *A1 = X + 3 * a1, *B0 = X + 3 * b0, *B1 = X + 3 * b1;
    double d2 = orc_EE_dist2(A0, A1, B0, B1, orc_EE_type(A0, A1, B0, B1));
    double a = orc_ee_ccd(A0, A1, B0, B1, D + 3 * a0, D + 3 * a1, D + 3 * b0, D + 3 * b1, c->eta, d2);
    if (a < c->amin[tid]) c->amin[tid] = a;
}
/* compute_intersection_free_step_size (NEWTON_IPC.py:456-468) */
ORC_API double orc_ccd_alpha(const double* x, const double* p, int nbp, const int* bp, int nbe, const int* be,
                             int nbt, const int* bt, int nt, const int* tets, double dHat, double inv_dx, double eta,
                             int inversion, double* parts)
{
    ccd_ctx* c = calloc(1, sizeof(ccd_ctx));
    c->x = x; c->p = p; c->bp = bp; c->be = be; c->bt = bt; c->dHat = dHat; c->eta = eta;
    for (int t = 0; t < MAXT; t++) c->amin[t] = 1.0;
    int nmax = nbp > nbe ? nbp : nbe;
    nmax = nmax > nbt ? nmax : nbt;
    double* lo = malloc(sizeof(double) * 3 * nmax);
    double* hi = malloc(sizeof(double) * 3 * nmax);
    double* qlo = malloc(sizeof(double) * 3 * nmax);
    double* qhi = malloc(sizeof(double) * 3 * nmax);
    for (int i = 0; i < nbp; i++) for (int d = 0; d < 3; d++) {
        double a = x[3 * bp[i] + d], b = a + p[3 * bp[i] + d];
        lo[3 * i + d] = dmin(a, b) - dHat; hi[3 * i + d] = dmax(a, b) + dHat;
    }
    for (int i = 0; i < nbt; i++) for (int d = 0; d < 3; d++) {
        double mn = 1e300, mx = -1e300;
        for (int k = 0; k < 3; k++) { double a = x[3 * bt[3 * i + k] + d], b = a + p[3 * bt[3 * i + k] + d]; mn = dmin(mn, dmin(a, b)); mx = dmax(mx, dmax(a, b)); }
        qlo[3 * i + d] = mn; qhi[3 * i + d] = mx;
    }
    grid_t g;
    grid_build(&g, nbp, lo, hi, inv_dx);
    grid_query(&g, nbt, qlo, qhi, lo, hi, inv_dx, ccd_pt, c);
    free(g.it);
    double apt = 1.0;
    for (int t = 0; t < MAXT; t++) { apt = dmin(apt, c->amin[t]); c->amin[t] = 1.0; }
    for (int i = 0; i < nbe; i++) for (int d = 0; d < 3; d++) {
        double mn = 1e300, mx = -1e300;
        for (int k = 0; k < 2; k++) { double a = x[3 * be[2 * i + k] + d], b = a + p[3 * be[2 * i + k] + d]; mn = dmin(mn, dmin(a, b)); mx = dmax(mx, dmax(a, b)); }
        lo[3 * i + d] = mn; hi[3 * i + d] = mx; qlo[3 * i + d] = mn - dHat; qhi[3 * i + d] = mx + dHat;
    }
    grid_build(&g, nbe, lo, hi, inv_dx);
    grid_query(&g, nbe, qlo, qhi, lo, hi, inv_dx, ccd_ee, c);
    free(g.it);
    double aee = 1.0;
    for (int t = 0; t < MAXT; t++) aee = dmin(aee, c->amin[t]);
    double ainv = 1.0;
    if (inversion) {
#pragma omp parallel for reduction(min:ainv)
        for (int e = 0; e < nt; e++) {
            const int* t = tets + 4 * e;
            double a = orc_tet_inversion_alpha(x + 3 * t[0], x + 3 * t[1], x + 3 * t[2], x + 3 * t[3],
                                               p + 3 * t[0], p + 3 * t[1], p + 3 * t[2], p + 3 * t[3], 0.2);
            if (a < ainv) ainv = a;
        }
    }
    free(lo); free(hi); free(qlo); free(qhi); free(c);
    if (parts) { parts[0] = apt; parts[1] = aee; parts[2] = ainv; }
    return dmin(dmin(apt, aee), ainv);
}
typedef struct { const double* x; const int *be, *bt; int hit; } col_ctx;
static void col_pair(void* vc, int tri, int ei, int tid)
{
    (void)tid;
    col_ctx* c = vc;
    int a0 = c->be[2 * ei], a1 = c->be[2 * ei + 1], t0 = c->bt[3 * tri], t1 = c->bt[3 * tri + 1], t2 = c->bt[3 * tri + 2];
    if (a0 == t0 || a0 == t1 || a0 == t2 || a1 == t0 || a1 == t1 || a1 == t2) return;
    const double* X = c->x;
    if (orc_seg_tri(X + 3 * a0, X + 3 * a1, X + 3 * t0, X + 3 * t1, X + 3 * t2)) c->hit = 1;
}
/* check_collision_3d (NEWTON_IPC.py:855-880) */
ORC_API int orc_check_collision(const double* x, int nbe, const int* be, int nbt, const int* bt, double dHat, double inv_dx)
{
    col_ctx c = {x, be, bt, 0};
    int nmax = nbe > nbt ? nbe : nbt;
    double* lo = malloc(sizeof(double) * 3 * nmax);
    double* hi = malloc(sizeof(double) * 3 * nmax);
    double* qlo = malloc(sizeof(double) * 3 * nmax);
    double* qhi = malloc(sizeof(double) * 3 * nmax);
    for (int i = 0; i < nbe; i++) for (int d = 0; d < 3; d++) {
        double a = x[3 * be[2 * i] + d], b = x[3 * be[2 * i + 1] + d];
        lo[3 * i + d] = dmin(a, b) - dHat; hi[3 * i + d] = dmax(a, b) + dHat;
    }
    for (int i = 0; i < nbt; i++) for (int d = 0; d < 3; d++) {
        double a = x[3 * bt[3 * i] + d], b = x[3 * bt[3 * i + 1] + d], e = x[3 * bt[3 * i + 2] + d];
        qlo[3 * i + d] = dmin(dmin(a, b), e); qhi[3 * i + d] = dmax(dmax(a, b), e);
    }
    grid_t g;
    grid_build(&g, nbe, lo, hi, inv_dx);
    grid_query(&g, nbt, qlo, qhi, lo, hi, inv_dx, col_pair, &c);
    free(g.it);
    free(lo); free(hi); free(qlo); free(qhi);
    return c.hit;
}
ORC_API int orc_num_threads(void) { return nthreads(); }
