// fipc_math.h -- FP64 per-element and per-constraint math of the IPC Newton
// step, written once for the gfx950 kernels (__device__) and the host-side
// scalar drop-ins (__host__).
//
// Reference semantics (file:line) are cited per function.  Gradients and
// Hessians of the squared distances are derived in reduced coordinates
// (difference vectors) and expanded to vertex coordinates with constant
// +-1 maps; the reference uses generated closed forms of the same functions.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define FI __host__ __device__ __forceinline__

namespace fipc {

// ---------------------------------------------------------------------------
// 3-vectors and 3x3 matrices (row-major)
// ---------------------------------------------------------------------------
struct v3 { double x, y, z; };
FI v3 mk(double a, double b, double c) { return v3{a, b, c}; }
FI v3 ld3(const double* p) { return v3{p[0], p[1], p[2]}; }
FI v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
FI v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
FI v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
FI v3 operator*(double s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
FI double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FI v3 cross(v3 a, v3 b) { return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
FI double n2(v3 a) { return dot(a, a); }
FI double at(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
FI v3 vmin(v3 a, v3 b) { return v3{fmin(a.x, b.x), fmin(a.y, b.y), fmin(a.z, b.z)}; }
FI v3 vmax(v3 a, v3 b) { return v3{fmax(a.x, b.x), fmax(a.y, b.y), fmax(a.z, b.z)}; }

struct m3 { double a[9]; };
FI m3 mzero() { m3 r; for (int i = 0; i < 9; i++) r.a[i] = 0.0; return r; }
FI m3 madd(const m3& A, const m3& B) { m3 r; for (int i = 0; i < 9; i++) r.a[i] = A.a[i] + B.a[i]; return r; }
FI m3 mneg(const m3& A) { m3 r; for (int i = 0; i < 9; i++) r.a[i] = -A.a[i]; return r; }
FI m3 mtrans(const m3& A) { m3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.a[i * 3 + j] = A.a[j * 3 + i]; return r; }
// s*I
FI m3 mdiag(double s) { m3 r = mzero(); r.a[0] = r.a[4] = r.a[8] = s; return r; }
// a b^T
FI m3 outer(v3 a, v3 b)
{
    m3 r;
    double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.a[i * 3 + j] = A[i] * B[j];
    return r;
}
// [c]_x (cross-product matrix: [c]_x v = c x v)
FI m3 skew(v3 c)
{
    m3 r;
    r.a[0] = 0; r.a[1] = -c.z; r.a[2] = c.y;
    r.a[3] = c.z; r.a[4] = 0; r.a[5] = -c.x;
    r.a[6] = -c.y; r.a[7] = c.x; r.a[8] = 0;
    return r;
}
FI double det3(const double* F)
{
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
}

// ---------------------------------------------------------------------------
// barrier (common/math/distance.py:21-44)
// ---------------------------------------------------------------------------
FI double barrier_E(double d, double h, double k)
{
    return d < h ? -k * (d - h) * (d - h) * log(d / h) : 0.0;
}
FI double barrier_g(double d, double h, double k)
{
    double t = d - h;
    return d < h ? k * (t * log(d / h) * -2.0 - (t * t) / d) : 0.0;
}
FI double barrier_H(double d, double h, double k)
{
    double t = d - h;
    return d < h ? k * ((log(d / h) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t)) : 0.0;
}

// ---------------------------------------------------------------------------
// squared distances (distance.py:955-956, :684-685, :976-979, :192-195, :1345)
// ---------------------------------------------------------------------------
FI double PP_E(v3 a, v3 b) { return n2(a - b); }
FI double PE_E(v3 p, v3 e0, v3 e1) { return n2(cross(e0 - p, e1 - p)) / n2(e1 - e0); }
FI double PT_E(v3 p, v3 t0, v3 t1, v3 t2)
{
    v3 n = cross(t1 - t0, t2 - t0);
    double s = dot(p - t0, n);
    return s * s / n2(n);
}
FI double EE_E(v3 a0, v3 a1, v3 b0, v3 b1)
{
    v3 n = cross(a1 - a0, b1 - b0);
    double s = dot(b0 - a0, n);
    return s * s / n2(n);
}
FI double EECN2(v3 a0, v3 a1, v3 b0, v3 b1) { return n2(cross(a1 - a0, b1 - b0)); }
FI double M_threshold(v3 a0, v3 a1, v3 b0, v3 b1) { return 1.0e-3 * n2(a0 - a1) * n2(b0 - b1); }

// ---------------------------------------------------------------------------
// closest-feature classification (distance.py:1675-1767)
// ---------------------------------------------------------------------------
FI void pt_param(v3 e, v3 nv, v3 d, double& s0, double& s1)
{
    v3 r1 = cross(e, nv);
    double m00 = dot(e, e), m01 = dot(e, r1), m10 = dot(r1, e), m11 = dot(r1, r1);
    double b0 = dot(e, d), b1 = dot(r1, d);
    double inv = 1.0 / (m00 * m11 - m01 * m10);
    s0 = (m11 * inv) * b0 + (-m01 * inv) * b1;
    s1 = (-m10 * inv) * b0 + (m00 * inv) * b1;
}
FI int PT_type(v3 p, v3 t0, v3 t1, v3 t2)
{
    v3 e0 = t1 - t0, nv = cross(e0, t2 - t0);
    double a0, b0, a1, b1, a2, b2;
    pt_param(e0, nv, p - t0, a0, b0);
    if (0.0 < a0 && a0 < 1.0 && 0.0 <= b0) return 3;
    pt_param(t2 - t1, nv, p - t1, a1, b1);
    if (0.0 < a1 && a1 < 1.0 && 0.0 <= b1) return 4;
    pt_param(t0 - t2, nv, p - t2, a2, b2);
    if (0.0 < a2 && a2 < 1.0 && 0.0 < b2) return 5;
    if (a0 <= 0.0 && a2 >= 1.0) return 0;
    if (a1 <= 0.0 && a0 >= 1.0) return 1;
    if (a2 <= 0.0 && a1 >= 1.0) return 2;
    return 6;
}
FI int EE_type(v3 a0, v3 a1, v3 b0, v3 b1)
{
    v3 u = a1 - a0, v = b1 - b0, w = a0 - b0;
    double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b, tD = D, sN = b * e - c * d, tN = 0.0;
    int k = 8;
    if (sN <= 0.0) { tN = e; tD = c; k = 2; }
    else if (sN >= D) { tN = e + b; tD = c; k = 5; }
    else {
        tN = a * e - b * d;
        v3 uv = cross(u, v);
        if (0.0 < tN && tN < tD && (dot(uv, w) == 0.0 || n2(uv) < 1.0e-20 * a * c)) {
            if (sN < D / 2) { tN = e; tD = c; k = 2; }
            else { tN = e + b; tD = c; k = 5; }
        }
    }
    if (tN <= 0.0) k = (-d <= 0.0) ? 0 : ((-d >= a) ? 3 : 6);
    else if (tN >= tD) k = (-d + b <= 0.0) ? 1 : ((-d + b >= a) ? 4 : 7);
    return k;
}
// PE_type (distance.py:1756-1767): where p projects on the line of (e0, e1)
FI int PE_type(v3 p, v3 e0, v3 e1)
{
    v3 e = e1 - e0;
    double ratio = dot(e, p - e0) / n2(e);
    return ratio < 0 ? 0 : (ratio > 1 ? 1 : 2);
}
FI double PT_dist2(v3 p, v3 t0, v3 t1, v3 t2, int c)
{
    switch (c) {
    case 0: return PP_E(p, t0);
    case 1: return PP_E(p, t1);
    case 2: return PP_E(p, t2);
    case 3: return PE_E(p, t0, t1);
    case 4: return PE_E(p, t1, t2);
    case 5: return PE_E(p, t2, t0);
    default: return PT_E(p, t0, t1, t2);
    }
}
FI double EE_dist2(v3 a0, v3 a1, v3 b0, v3 b1, int c)
{
    switch (c) {
    case 0: return PP_E(a0, b0);
    case 1: return PP_E(a0, b1);
    case 2: return PE_E(a0, b0, b1);
    case 3: return PP_E(a1, b0);
    case 4: return PP_E(a1, b1);
    case 5: return PE_E(a1, b0, b1);
    case 6: return PE_E(b0, a0, a1);
    case 7: return PE_E(b1, a0, a1);
    default: return EE_E(a0, a1, b0, b1);
    }
}

// ---------------------------------------------------------------------------
// reduced-coordinate derivatives
//
// A distance is written in terms of R difference vectors y_r (R = 2 or 3) and
// expanded to K vertices by x-coefficients C[k][r] in {-1, 0, 1}:
//     grad_x[k] = sum_r C[k][r] grad_y[r]
//     H_x[k][l] = sum_rs C[k][r] C[l][s] H_y[r][s]
// ---------------------------------------------------------------------------
template <int R> struct RGrad { v3 g[R]; m3 H[R][R]; double v; };

// N(u, w) = |u x w|^2 : value, gradient, Hessian in (u, w)
FI void cross2_derivs(v3 u, v3 w, RGrad<2>& o)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    v3 c = cross(u, w);
    o.v = n2(c);
    o.g[0] = 2.0 * cross(w, c);
    o.g[1] = 2.0 * cross(c, u);
    double ww = n2(w), uu = n2(u), uw = dot(u, w);
    m3 Huu = madd(mdiag(2.0 * ww), mneg(outer(2.0 * w, w)));
    m3 Hww = madd(mdiag(2.0 * uu), mneg(outer(2.0 * u, u)));
    // N_uw = 2 (u w^T - (u.w) I) - 2 [c]_x
    m3 Huw = madd(madd(outer(2.0 * u, w), mdiag(-2.0 * uw)), mneg(skew(2.0 * c)));
    o.H[0][0] = Huu; o.H[1][1] = Hww; o.H[0][1] = Huw; o.H[1][0] = mtrans(Huw);
}
// d = (r . (e1 x e2))^2 / |e1 x e2|^2 in (r, e1, e2)
FI void tri_derivs(v3 r, v3 e1, v3 e2, RGrad<3>& o)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    v3 n = cross(e1, e2);
    double s = dot(r, n);
    RGrad<2> q;
    cross2_derivs(e1, e2, q);
    double Q = q.v, d = s * s / Q;
    o.v = d;
    // gradient of s
    v3 gs[3] = {n, cross(e2, r), cross(r, e1)};
    // Hessian of s (trilinear): s_r,e1 = -[e2]x, s_r,e2 = [e1]x, s_e1,e2 = -[r]x
    m3 Hs[3][3];
    Hs[0][0] = Hs[1][1] = Hs[2][2] = mzero();
    Hs[0][1] = mneg(skew(e2)); Hs[1][0] = mtrans(Hs[0][1]);
    Hs[0][2] = skew(e1);       Hs[2][0] = mtrans(Hs[0][2]);
    Hs[1][2] = mneg(skew(r));  Hs[2][1] = mtrans(Hs[1][2]);
    v3 gQ[3] = {mk(0, 0, 0), q.g[0], q.g[1]};
    double invQ = 1.0 / Q;
    for (int i = 0; i < 3; i++) o.g[i] = invQ * (2.0 * s * gs[i] - d * gQ[i]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            m3 h = madd(outer(2.0 * gs[i], gs[j]), mzero());
            for (int t = 0; t < 9; t++) h.a[t] += 2.0 * s * Hs[i][j].a[t];
            m3 a = outer(o.g[i], gQ[j]), b = outer(gQ[i], o.g[j]);
            for (int t = 0; t < 9; t++) h.a[t] -= a.a[t] + b.a[t];
            if (i > 0 && j > 0) for (int t = 0; t < 9; t++) h.a[t] -= d * q.H[i - 1][j - 1].a[t];
            for (int t = 0; t < 9; t++) h.a[t] *= invQ;
            o.H[i][j] = h;
        }
}
// d = |u x w|^2 / |w - u|^2 in (u = e0 - p, w = e1 - p)
FI void pe_derivs(v3 u, v3 w, RGrad<2>& o)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    RGrad<2> Nn;
    cross2_derivs(u, w, Nn);
    v3 e = w - u;
    double L = n2(e), d = Nn.v / L, invL = 1.0 / L;
    o.v = d;
    v3 gL[2] = {-2.0 * e, 2.0 * e};
    for (int i = 0; i < 2; i++) o.g[i] = invL * (Nn.g[i] - d * gL[i]);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            m3 a = outer(o.g[i], gL[j]), b = outer(gL[i], o.g[j]);
            double hl = (i == j) ? 2.0 : -2.0;  // Hessian of L: [[2I, -2I], [-2I, 2I]]
            m3 h;
            for (int t = 0; t < 9; t++) h.a[t] = Nn.H[i][j].a[t] - a.a[t] - b.a[t];
            h.a[0] -= d * hl; h.a[4] -= d * hl; h.a[8] -= d * hl;
            for (int t = 0; t < 9; t++) h.a[t] *= invL;
            o.H[i][j] = h;
        }
}

// expand reduced derivatives to K vertices; g has 3K entries, H is K*3 x K*3
// row-major.  C is K x R coefficients.
template <int K, int R>
FI void expand(const RGrad<R>& y, const int (&C)[K][R], double* g, double* H)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
#pragma unroll
    for (int k = 0; k < K; k++) {
        v3 s = mk(0, 0, 0);
#pragma unroll
        for (int r = 0; r < R; r++) if (C[k][r]) s = s + (double)C[k][r] * y.g[r];
        g[3 * k] = s.x; g[3 * k + 1] = s.y; g[3 * k + 2] = s.z;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int l = 0; l < K; l++) {
            double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int s = 0; s < R; s++) {
                    int c = C[k][r] * C[l][s];
                    if (c) for (int t = 0; t < 9; t++) b[t] += c * y.H[r][s].a[t];
                }
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) H[(3 * k + i) * (3 * K) + 3 * l + j] = b[i * 3 + j];
        }
}

// squared distance value, gradient and Hessian in vertex coordinates
FI double PP_gH(v3 a, v3 b, double* g, double* H)
{
    v3 d = a - b;
    g[0] = 2 * d.x; g[1] = 2 * d.y; g[2] = 2 * d.z; g[3] = -2 * d.x; g[4] = -2 * d.y; g[5] = -2 * d.z;
    for (int i = 0; i < 36; i++) H[i] = 0.0;
    for (int i = 0; i < 3; i++) { H[i * 6 + i] = 2.0; H[(i + 3) * 6 + i + 3] = 2.0; H[i * 6 + i + 3] = -2.0; H[(i + 3) * 6 + i] = -2.0; }
    return n2(d);
}
FI double PE_gH(v3 p, v3 e0, v3 e1, double* g, double* H)
{
    RGrad<2> y;
    pe_derivs(e0 - p, e1 - p, y);
    const int C[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
    expand<3, 2>(y, C, g, H);
    return y.v;
}
FI double PT_gH(v3 p, v3 t0, v3 t1, v3 t2, double* g, double* H)
{
    RGrad<3> y;
    tri_derivs(p - t0, t1 - t0, t2 - t0, y);
    const int C[4][3] = {{1, 0, 0}, {-1, -1, -1}, {0, 1, 0}, {0, 0, 1}};
    expand<4, 3>(y, C, g, H);
    return y.v;
}
FI double EE_gH(v3 a0, v3 a1, v3 b0, v3 b1, double* g, double* H)
{
    RGrad<3> y;
    tri_derivs(b0 - a0, a1 - a0, b1 - b0, y);
    const int C[4][3] = {{-1, -1, 0}, {0, 1, 0}, {1, 0, -1}, {0, 0, 1}};
    expand<4, 3>(y, C, g, H);
    return y.v;
}
FI double EECN2_gH(v3 a0, v3 a1, v3 b0, v3 b1, double* g, double* H)
{
    RGrad<2> y;
    cross2_derivs(a1 - a0, b1 - b0, y);
    const int C[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    expand<4, 2>(y, C, g, H);
    return y.v;
}

// ---------------------------------------------------------------------------
// per-constraint energy, gradient and Hessian (common/math/ipc.py:30-329)
// kind: 0 PP, 1 PE, 2 PT, 3 EE, 4 EEM, 5 PPM, 6 PEM
// x holds the constraint's points (K = 2, 3, 4), x0 their rest positions.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int kind_points(int kind) { return kind == 0 ? 2 : (kind == 1 ? 3 : 4); }
__host__ __device__ constexpr int kind_dofs(int kind) { return 3 * kind_points(kind); }

FI double constraint_dist(int kind, const v3* x)
{
    switch (kind) {
    case 0: return PP_E(x[0], x[1]);
    case 1: return PE_E(x[0], x[1], x[2]);
    case 2: return PT_E(x[0], x[1], x[2], x[3]);
    case 3: case 4: return EE_E(x[0], x[1], x[2], x[3]);
    case 5: return PP_E(x[0], x[2]);
    default: return PE_E(x[0], x[2], x[3]);
    }
}
FI double mollifier_E(const v3* x, const v3* x0)
{
    double eps = M_threshold(x0[0], x0[1], x0[2], x0[3]);
    double c = EECN2(x[0], x[1], x[2], x[3]);
    if (!(c < eps)) return 1.0;
    double r = c / eps;
    return (-r + 2.0) * r;
}
FI double constraint_E(int kind, const v3* x, const v3* x0, double dHat2, double kappa)
{
    double b = barrier_E(constraint_dist(kind, x), dHat2, kappa);
    return kind >= 4 ? b * mollifier_E(x, x0) : b;
}

// g (ND) and the unprojected H (ND x ND, row-major) of one constraint; returns E
template <int KIND>
FI double constraint_gH(const v3* x, const v3* x0, double dHat2, double kappa, double* g, double* H)
{
    constexpr int ND = kind_dofs(KIND);
    double dg[12], dH[144];
    double d;
    if (KIND == 0) d = PP_gH(x[0], x[1], dg, dH);
    else if (KIND == 1) d = PE_gH(x[0], x[1], x[2], dg, dH);
    else if (KIND == 2) d = PT_gH(x[0], x[1], x[2], x[3], dg, dH);
    else if (KIND == 3 || KIND == 4) d = EE_gH(x[0], x[1], x[2], x[3], dg, dH);
    else if (KIND == 5) d = PP_gH(x[0], x[2], dg, dH);
    else d = PE_gH(x[0], x[2], x[3], dg, dH);
    double b = barrier_E(d, dHat2, kappa), b1 = barrier_g(d, dHat2, kappa), b2 = barrier_H(d, dHat2, kappa);
    if (KIND <= 3) {
#pragma unroll
        for (int i = 0; i < ND; i++) g[i] = b1 * dg[i];
#pragma unroll
        for (int i = 0; i < ND; i++)
#pragma unroll
            for (int j = 0; j < ND; j++) H[i * ND + j] = b2 * dg[i] * dg[j] + b1 * dH[i * ND + j];
        return b;
    } else {
        // distance dofs placed in 12-space (fill_vec / fill_mat, ipc.py:254, :300)
        constexpr int nd = KIND == 5 ? 6 : (KIND == 6 ? 9 : 12);
        int map[12];
#pragma unroll
        for (int i = 0; i < 12; i++) map[i] = i;
        if (KIND == 5 || KIND == 6) {
#pragma unroll
            for (int i = 3; i < 12; i++) map[i] = i + 3;
        }
        double lg[12];
#pragma unroll
        for (int i = 0; i < 12; i++) lg[i] = 0.0;
#pragma unroll
        for (int i = 0; i < nd; i++) lg[map[i]] = b1 * dg[i];
        // mollifier (distance.py:1641-1667)
        double eps = M_threshold(x0[0], x0[1], x0[2], x0[3]);
        double cg[12], cH[144];
        double c = EECN2_gH(x[0], x[1], x[2], x[3], cg, cH);
        double m = 1.0, m1 = 0.0, m2 = 0.0;
        bool act = c < eps;
        if (act) {
            double r = c / eps;
            m = (-r + 2.0) * r;
            m1 = 2.0 * (1.0 / eps) * (-(1.0 / eps) * c + 1.0);
            m2 = -2.0 / (eps * eps);
        }
#pragma unroll
        for (int i = 0; i < 12; i++) g[i] = lg[i] * m + b * (m1 * cg[i]);
#pragma unroll
        for (int i = 0; i < 12; i++)
#pragma unroll
            for (int j = 0; j < 12; j++) H[i * 12 + j] = b * (m1 * cH[i * 12 + j] + m2 * cg[i] * cg[j]) + lg[i] * (m1 * cg[j]) + (m1 * cg[i]) * lg[j];
#pragma unroll
        for (int i = 0; i < nd; i++)
#pragma unroll
            for (int j = 0; j < nd; j++) H[map[i] * 12 + map[j]] += m * (b2 * dg[i] * dg[j] + b1 * dH[i * nd + j]);
        return b * m;
    }
}

// ---------------------------------------------------------------------------
// symmetric eigen-decomposition by cyclic Jacobi on packed upper-triangular
// storage.  The caller owns R rows of the eigenvector matrix, starting at
// row0; rotations act on rows independently, so R = N gives the full V and
// a group of lanes can each hold a slice.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int pk(int N, int i, int j)
{
    return i <= j ? i * N - i * (i - 1) / 2 + (j - i) : j * N - j * (j - 1) / 2 + (i - j);
}
template <int N, int R>
FI void jacobi_eig(double* A, double (&V)[R][N], int row0, int max_sweeps)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int c = 0; c < N; c++) V[r][c] = (row0 + r == c) ? 1.0 : 0.0;
    double fro = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = i; j < N; j++) fro += (i == j ? 1.0 : 2.0) * A[pk(N, i, j)] * A[pk(N, i, j)];
    const double tiny = 1e-17 * sqrt(fro);
    for (int sweep = 0; sweep < max_sweeps; sweep++) {
        bool rotated = false;
#pragma unroll
        for (int p = 0; p < N - 1; p++) {
#pragma unroll
            for (int q = p + 1; q < N; q++) {
                double apq = A[pk(N, p, q)];
                if (fabs(apq) > tiny) {
                    rotated = true;
                    double app = A[pk(N, p, p)], aqq = A[pk(N, q, q)];
                    double theta = (aqq - app) / (2.0 * apq);
                    double t = fabs(theta) > 1e150 ? 0.5 / theta
                             : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                    A[pk(N, p, p)] = app - t * apq;
                    A[pk(N, q, q)] = aqq + t * apq;
                    A[pk(N, p, q)] = 0.0;
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        if (k == p || k == q) continue;
                        double akp = A[pk(N, k, p)], akq = A[pk(N, k, q)];
                        A[pk(N, k, p)] = c * akp - s * akq;
                        A[pk(N, k, q)] = s * akp + c * akq;
                    }
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        double vp = V[r][p], vq = V[r][q];
                        V[r][p] = c * vp - s * vq;
                        V[r][q] = s * vp + c * vq;
                    }
                }
            }
        }
        if (!rotated) break;
    }
}
// makePD (common/math/wrapper/UTILS.h:7-25) on a full N x N row-major matrix,
// single thread: if the smallest eigenvalue is negative, clamp negative
// eigenvalues to zero and rebuild V D V^T.
template <int N>
FI void make_pd_full(double* M)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    double A[N * (N + 1) / 2];
    for (int i = 0; i < N; i++) for (int j = i; j < N; j++) A[pk(N, i, j)] = 0.5 * (M[i * N + j] + M[j * N + i]);
    double V[N][N];
    jacobi_eig<N, N>(A, V, 0, 30);
    double lam[N];
    bool neg = false;
    for (int i = 0; i < N; i++) { lam[i] = A[pk(N, i, i)]; if (lam[i] < 0) { neg = true; lam[i] = 0.0; } }
    if (!neg) return;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            double s = 0.0;
            for (int k = 0; k < N; k++) s += V[i][k] * lam[k] * V[j][k];
            M[i * N + j] = s;
        }
}
// 2x2 closed form (common/math/math_tools.py:39-64); M row-major
FI void make_pd2(double* M)
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

// ---------------------------------------------------------------------------
// neo-Hookean (common/physics/neo_hookean.py)
// ---------------------------------------------------------------------------
// SVD of F (row-major) through the eigenvectors of F^T F: det U = det V = 1,
// singular values in descending order, the last one signed by det F.
FI void svd3(const double* F, double* U, double* s, double* V)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    double A[6];
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) A[pk(3, i, j)] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    double W[3][3];
    jacobi_eig<3, 3>(A, W, 0, 30);
    double lam[3] = {A[pk(3, 0, 0)], A[pk(3, 1, 1)], A[pk(3, 2, 2)]};
    // sort eigenpairs by descending eigenvalue (column swaps with static indices)
    auto cswap = [&](int a, int b) {
        if (lam[b] > lam[a]) {
            double t = lam[a]; lam[a] = lam[b]; lam[b] = t;
            for (int i = 0; i < 3; i++) { t = W[i][a]; W[i][a] = W[i][b]; W[i][b] = t; }
        }
    };
    cswap(0, 1);
    cswap(1, 2);
    cswap(0, 1);
    for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) V[i * 3 + k] = W[i][k];
    if (det3(V) < 0) for (int i = 0; i < 3; i++) V[i * 3 + 2] = -V[i * 3 + 2];
    v3 f0 = mk(F[0] * V[0] + F[1] * V[3] + F[2] * V[6], F[3] * V[0] + F[4] * V[3] + F[5] * V[6], F[6] * V[0] + F[7] * V[3] + F[8] * V[6]);
    v3 f1 = mk(F[0] * V[1] + F[1] * V[4] + F[2] * V[7], F[3] * V[1] + F[4] * V[4] + F[5] * V[7], F[6] * V[1] + F[7] * V[4] + F[8] * V[7]);
    v3 f2 = mk(F[0] * V[2] + F[1] * V[5] + F[2] * V[8], F[3] * V[2] + F[4] * V[5] + F[5] * V[8], F[6] * V[2] + F[7] * V[5] + F[8] * V[8]);
    s[0] = sqrt(n2(f0));
    v3 u0 = (1.0 / s[0]) * f0;
    v3 u1 = f1 - dot(u0, f1) * u0;
    double l1 = sqrt(n2(u1));
    u1 = l1 > 0 ? (1.0 / l1) * u1 : u1;
    v3 u2 = cross(u0, u1);
    s[1] = dot(u1, f1);
    s[2] = dot(u2, f2);
    U[0] = u0.x; U[3] = u0.y; U[6] = u0.z;
    U[1] = u1.x; U[4] = u1.y; U[7] = u1.z;
    U[2] = u2.x; U[5] = u2.y; U[8] = u2.z;
}
// Psi from invariants: mu/2 (|F|^2 - 3) - mu ln J + la/2 (ln J)^2, equal to
// the singular-value form of neo_hookean.py:5-16
FI double nh_energy_F(const double* F, double la, double mu)
{
    double f2 = 0;
    for (int i = 0; i < 9; i++) f2 += F[i] * F[i];
    double lj = log(det3(F));
    return mu / 2.0 * (f2 - 3) - (mu - la / 2.0 * lj) * lj;
}
// P(F) (neo_hookean.py:61-67)
FI void nh_P(const double* F, double la, double mu, double* P)
{
    double J = det3(F), scale = la * log(J) - mu;
    double C[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                   F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                   F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
    for (int i = 0; i < 9; i++) P[i] = mu * F[i] + scale * (C[i] / J);
}
// The projected dP/dF of neo_hookean.py:70-131 in factored form:
//     dPdF = sum_{k,l} M[k][l] w_k w_l^T,   w_{a*3+b} = vec(u_a v_b^T)
// M holds the projected sigma-Hessian (pairs (a,a),(b,b)) and three projected
// 2x2 blocks coupling (a,b) with (b,a).
// dPsi/dsigma and d2Psi/dsigma2 (neo_hookean.py:19-58), H row-major 3x3
FI void nh_sigma_gradient(const double* s, double la, double mu, double* g)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    double lp = log(s[0] * s[1] * s[2]);
    for (int i = 0; i < 3; i++) g[i] = mu * (s[i] - 1.0 / s[i]) + la * (1.0 / s[i]) * lp;
}
FI void nh_sigma_hessian(const double* s, double la, double mu, double* H)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    double lp = log(s[0] * s[1] * s[2]);
    for (int i = 0; i < 3; i++) {
        double inv2 = 1.0 / s[i] / s[i];
        H[i * 4] = mu * (1.0 + inv2) - la * inv2 * (lp - 1.0);
    }
    H[1] = H[3] = la / s[0] / s[1];
    H[5] = H[7] = la / s[1] / s[2];
    H[2] = H[6] = la / s[0] / s[2];
}
struct NHFactor {
    double U[9], V[9];
    double S[3][3];   // projected d2Psi/dsigma2
    double B[3][4];   // make_pd'd blocks for pairs (0,1), (1,2), (2,0)
};
FI void nh_factor(const double* F, double la, double mu, NHFactor& o)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    double s[3];
    svd3(F, o.U, s, o.V);
    double lp = log(s[0] * s[1] * s[2]), middle = mu - la * lp;
    double g[3];
    nh_sigma_gradient(s, la, mu, g);
    double H[9];
    nh_sigma_hessian(s, la, mu, H);
    make_pd_full<3>(H);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o.S[i][j] = H[i * 3 + j];
    const int pi[3] = {0, 1, 2}, pj[3] = {1, 2, 0};
    for (int k = 0; k < 3; k++) {
        int i = pi[k], j = pj[k];
        double left = (mu + middle / s[i] / s[j]) / 2.0;
        double right = (g[i] + g[j]) / (2 * fmax(s[i] + s[j], 0.000001));
        o.B[k][0] = left + right; o.B[k][1] = left - right; o.B[k][2] = left - right; o.B[k][3] = left + right;
        make_pd2(o.B[k]);
    }
}

// The full projected dP/dF (neo_hookean.py:70-131) as a 9x9 row-major matrix
// whose rows and columns index the column-major vec of F (F[i][j] -> 3 j + i),
// expanded from the factored form: w_(a,b)[3 j + i] = U[i][a] V[j][b].
FI void nh_dPdF(const double* F, double la, double mu, double* D)
{
    NHFactor f;
    nh_factor(F, la, mu, f);
    double Mk[9][9];
    for (int k = 0; k < 9; k++)
        for (int l = 0; l < 9; l++) Mk[k][l] = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Mk[a * 4][b * 4] = f.S[a][b];
    const int PI[3] = {0, 1, 2}, PJ[3] = {1, 2, 0};
    for (int k = 0; k < 3; k++) {
        int ij = PI[k] * 3 + PJ[k], ji = PJ[k] * 3 + PI[k];
        Mk[ij][ij] = f.B[k][0];
        Mk[ij][ji] = f.B[k][1];
        Mk[ji][ij] = f.B[k][2];
        Mk[ji][ji] = f.B[k][3];
    }
    double w[9][9];  // w[k][vec index]
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) w[a * 3 + b][3 * j + i] = f.U[i * 3 + a] * f.V[j * 3 + b];
    for (int r = 0; r < 9; r++)
        for (int c = 0; c < 9; c++) {
            double acc = 0.0;
            for (int k = 0; k < 9; k++)
                for (int l = 0; l < 9; l++) acc += Mk[k][l] * w[k][r] * w[l][c];
            D[r * 9 + c] = acc;
        }
}

// 12x12 tet Hessian as Z M Z^T (see NHFactor), accumulated per 3x3 block in
// the arithmetic type T (double, or float where the blocks are stored in
// FP32 anyway: the factors U, V, S, B are computed in FP64 and rounded once).
// emit(a, c, blk) receives the block of vertex pair (a, c), c >= a, as T[9].
template <class T, class Emit>
FI void tet_hessian_blocks_t(const NHFactor& f, const double (&Bm_)[4][3], double sc_, Emit emit)
{
    // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
    T U[9], V[9], S[3][3], Bk[3][4], Bm[4][3];
#pragma unroll
    for (int k = 0; k < 9; k++) { U[k] = (T)f.U[k]; V[k] = (T)f.V[k]; }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] = (T)f.S[i][j];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 4; j++) Bk[k][j] = (T)f.B[k][j];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int j = 0; j < 3; j++) Bm[a][j] = (T)Bm_[a][j];
    const T sc = (T)sc_;
    // bv[a][b] = sum_j Bm[a][j] V[j][b]
    T bv[4][3];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) bv[a][b] = Bm[a][0] * V[0 * 3 + b] + Bm[a][1] * V[1 * 3 + b] + Bm[a][2] * V[2 * 3 + b];
    // M in the (ka,kb) pair basis: diagonal pairs via S, off pairs via B blocks
    // B block k couples idx (i,j) and (j,i) for (i,j) = (0,1), (1,2), (2,0)
    const int PI[3] = {0, 1, 2}, PJ[3] = {1, 2, 0};
#pragma unroll
    for (int a = 0; a < 4; a++) {
        // Y[i][l] for l = ka*3+kb: sum_k z_k[(a,i)] M[k][l]
        T Y[3][9];
#pragma unroll
        for (int i = 0; i < 3; i++) {
#pragma unroll
            for (int l = 0; l < 9; l++) Y[i][l] = (T)0;
#pragma unroll
            for (int be = 0; be < 3; be++) {
                T s = (T)0;
#pragma unroll
                for (int al = 0; al < 3; al++) s += bv[a][al] * U[i * 3 + al] * S[al][be];
                Y[i][be * 4] = s * sc;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int pi = PI[k], pj = PJ[k];
                // z_(pi,pj)[(a,i)] = bv[a][pj] U[i][pi]; z_(pj,pi)[(a,i)] = bv[a][pi] U[i][pj]
                T zij = bv[a][pj] * U[i * 3 + pi], zji = bv[a][pi] * U[i * 3 + pj];
                // block k on the basis ((pi,pj), (pj,pi)), row-major (neo_hookean.py:112-123;
                // for k = 2 the basis is (6, 2) in the reference's M indexing)
                const T B00 = Bk[k][0], B01 = Bk[k][1], B10 = Bk[k][2], B11 = Bk[k][3];
                Y[i][pi * 3 + pj] = (zij * B00 + zji * B10) * sc;
                Y[i][pj * 3 + pi] = (zij * B01 + zji * B11) * sc;
            }
        }
#pragma unroll
        for (int c = a; c < 4; c++) {
            // z_l[(c,j)] = bv[c][lb] U[j][la] for l = la*3+lb, so
            // blk[i][j] = sum_la U[j][la] Tm[i][la], Tm[i][la] = sum_lb Y[i][la*3+lb] bv[c][lb]
            T Tm[3][3];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int la = 0; la < 3; la++)
                    Tm[i][la] = Y[i][la * 3] * bv[c][0] + Y[i][la * 3 + 1] * bv[c][1] + Y[i][la * 3 + 2] * bv[c][2];
            T blk[9];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    blk[i * 3 + j] = Tm[i][0] * U[j * 3] + Tm[i][1] * U[j * 3 + 1] + Tm[i][2] * U[j * 3 + 2];
            emit(a, c, blk);
        }
    }
}

// the FP64 form (host drop-ins, FP64 mode)
template <class Emit>
FI void tet_hessian_blocks(const NHFactor& f, const double (&Bm)[4][3], double sc, Emit emit)
{
    tet_hessian_blocks_t<double>(f, Bm, sc, emit);
}

// ---------------------------------------------------------------------------
// roots (common/math/wrapper/wrapper.cpp:200-252) -- Cardano with complex
// arithmetic; the principal cube root is taken in polar form
// ---------------------------------------------------------------------------
struct cplx { double re, im; };
FI cplx cmul(cplx a, cplx b) { return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
FI cplx cdiv(cplx a, cplx b)
{
    // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, den = b.re + b.im * r;
        return cplx{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    } else {
        double r = b.re / b.im, den = b.re * r + b.im;
        return cplx{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
    }
}
FI double cabs_(cplx a) { return hypot(a.re, a.im); }
FI cplx csqrt_(cplx z)
{
    if (z.re == 0.0 && z.im == 0.0) return cplx{0.0, z.im};
    double m = cabs_(z);
    double t = sqrt(0.5 * (m + fabs(z.re)));
    if (z.re >= 0) return cplx{t, z.im / (2 * t)};
    return cplx{fabs(z.im) / (2 * t), copysign(t, z.im)};
}
FI cplx ccbrt_(cplx z)
{
    double m = cabs_(z);
    if (m == 0.0) return cplx{0.0, 0.0};
    double r = pow(m, 1.0 / 3.0), a = atan2(z.im, z.re) * (1.0 / 3.0);
    return cplx{r * cos(a), r * sin(a)};
}
FI double quad_root(double a, double b, double c, double tol)
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
FI double cubic_root(double a, double b, double c, double d, double tol)
{
    if (fabs(a) <= tol) return quad_root(b, c, d, tol);
    double t = -1;
    cplx d0 = cplx{b * b - 3 * a * c, 0.0};
    cplx d1 = cplx{2 * b * b * b - 9 * a * b * c + 27 * a * a * d, 0.0};
    cplx d03 = cmul(cmul(d0, d0), d0);
    cplx disc = csqrt_(cplx{d1.re * d1.re - 4.0 * d03.re, -4.0 * d03.im + 0.0});
    cplx C = ccbrt_(cplx{(d1.re + disc.re) / 2.0, (d1.im + disc.im) / 2.0});
    if (cabs_(C) == 0.0) C = ccbrt_(cplx{(d1.re - disc.re) / 2.0, (d1.im - disc.im) / 2.0});
    const double h = 0.8660254037844386;  // sqrt(3)/2
    cplx u[3] = {cplx{1.0, 0.0}, cplx{-0.5, h}, cplx{-0.5, -h}};
    for (int k = 0; k < 3; k++) {
        cplx uc = cmul(u[k], C);
        cplx q = cdiv(d0, uc);
        cplx num = cplx{b + uc.re + q.re, uc.im + q.im};
        cplx r = cplx{num.re / (-3.0 * a), num.im / (-3.0 * a)};
        if (fabs(r.im) < tol && r.re > 0 && (k == 0 || r.re < t || t < 0)) t = r.re;
    }
    return t;
}
// coefficients of det(x + t p) = a t^3 + b t^2 + c t + d of a tet's edge
// matrix (math_tools.py:160-181)
FI void tet_det_cubic(v3 x1, v3 x2, v3 x3, v3 x4, v3 p1, v3 p2, v3 p3, v3 p4, double& a, double& b, double& c,
                      double& d)
{
    v3 v01 = x2 - x1, v02 = x3 - x1, v03 = x4 - x1, p01 = p2 - p1, p02 = p3 - p1, p03 = p4 - p1;
    v3 pp = cross(p01, p02), s = cross(v01, p02) + cross(p01, v02), c0 = cross(v01, v02);
    a = dot(p03, pp); b = dot(v03, pp) + dot(p03, s); c = dot(p03, c0) + dot(v03, s); d = dot(v03, c0);
}
// inversion-free step bound of one tet (math_tools.py:160-181)
FI double tet_inversion_alpha(v3 x1, v3 x2, v3 x3, v3 x4, v3 p1, v3 p2, v3 p3, v3 p4, double eta)
{
    double a, b, c, d;
    tet_det_cubic(x1, x2, x3, x4, p1, p2, p3, p4, a, b, c, d);
    double r = cubic_root(a, b, c, d, 1e-8);
    return r < 0.0 ? 1.0 : r * (1.0 - eta);
}
// true when the tet's bound is certainly >= 1, so that it cannot lower a
// minimum that starts at 1: the bound is r (1 - eta) for the smallest
// positive root r, which matters only for r < T = 1 / (1 - eta); with d > 0
// and 2 (|a| T^3 + |b| T^2 + |c| T) < d, det(x + t p) > d / 2 on [0, T], so
// there is no real root there and no complex one close enough to the real
// axis to pass the root finder's tolerance
FI bool tet_inversion_clear(double a, double b, double c, double d, double eta)
{
    double T = 1.0 / (1.0 - eta);
    return d > 0.0 && 2.0 * (fabs(a) * T * T * T + fabs(b) * T * T + fabs(c) * T) < d;
}

// ---------------------------------------------------------------------------
// CCD (EVCTCD/CTCD.cpp:10-472, wrapper.cpp:254-328), evaluated twice: once
// with every multiply-add rounded twice (strict IEEE, as the rest of the
// discrete decisions here) and once with fused multiply-adds.  On exactly
// degenerate pairs -- parallel, lattice-aligned edges moving rigidly -- the
// polynomial coefficients vanish in exact arithmetic and the interval
// decisions follow the rounding of the products that cancel; the
// reference's own build (-mfma) fuses them.  The step bound is the smaller
// of the two results, so it is never later than either rounding's.
// ---------------------------------------------------------------------------
namespace ccd_strict {
#define MAD(a, b, c) ((a) * (b) + (c))
#include "fipc_ccd_impl.h"
#undef MAD
}  // namespace ccd_strict
namespace ccd_fused {
#define MAD(a, b, c) fma((a), (b), (c))
#include "fipc_ccd_impl.h"
#undef MAD
}  // namespace ccd_fused
// point_triangle_ccd (wrapper.cpp:254-290)
FI double pt_ccd(v3 p, v3 t0, v3 t1, v3 t2, v3 dp, v3 dt0, v3 dt1, v3 dt2, double eta, double dist2)
{
    return fmin(ccd_strict::pt_ccd(p, t0, t1, t2, dp, dt0, dt1, dt2, eta, dist2),
                ccd_fused::pt_ccd(p, t0, t1, t2, dp, dt0, dt1, dt2, eta, dist2));
}
// edge_edge_ccd (wrapper.cpp:292-328)
FI double ee_ccd(v3 a0, v3 a1, v3 b0, v3 b1, v3 da0, v3 da1, v3 db0, v3 db1, double eta, double dist2)
{
    return fmin(ccd_strict::ee_ccd(a0, a1, b0, b1, da0, da1, db0, db1, eta, dist2),
                ccd_fused::ee_ccd(a0, a1, b0, b1, da0, da1, db0, db1, eta, dist2));
}
// segment_intersect_triangle (common/math/ipc.py:511-526)
FI bool seg_tri(v3 P, v3 Q, v3 A, v3 B, v3 C)
{
    v3 R = Q - P;
    double RLen = sqrt(n2(R));
    R = v3{R.x / RLen, R.y / RLen, R.z / RLen};
    v3 E1 = B - A, E2 = C - A, N = cross(E1, E2);
    double det = -dot(R, N), invdet = 1.0 / det;
    v3 AO = P - A, DAO = cross(AO, R);
    double u = dot(E2, DAO) * invdet, v = -dot(E1, DAO) * invdet, t = dot(AO, N) * invdet;
    return det >= 1e-12 && t >= 0.0 && u >= 0.0 && v >= 0.0 && (u + v) <= 1.0 && t <= RLen;
}

}  // namespace fipc
