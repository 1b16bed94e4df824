// fipc_ccd_impl.h -- interval CCD with the semantics of EVCTCD/CTCD.cpp:10-472,
// included by fipc_math.h twice, inside the namespaces fipc::ccd_strict and
// fipc::ccd_fused, once under each floating-point contraction mode.  Every
// floating-point operation of the CCD is written in this file (the vector
// helpers below included) so that the mode of the inclusion governs all of
// them.  Polynomial coefficients are lowest degree first.  Real roots in
// [0,1] are isolated by a derivative cascade (roots of p' split [0,1] into
// monotone pieces, bisected to full precision).
struct Ivl { double l, u; };
struct IvlSet { int n; Ivl v[8]; };
// every product that feeds an addition is written MAD(a, b, c) = a b + c:
// two roundings in the strict inclusion, one (fma) in the fused one;
// strict gives the same bits as the plain expressions (a + b == b + a and
// negation are exact)
FI double dotm(v3 a, v3 b) { return MAD(a.z, b.z, MAD(a.y, b.y, a.x * b.x)); }
FI v3 crossm(v3 a, v3 b)
{
    return v3{MAD(a.y, b.z, -(a.z * b.y)), MAD(a.z, b.x, -(a.x * b.z)), MAD(a.x, b.y, -(a.y * b.x))};
}
FI double n2m(v3 a) { return dotm(a, a); }

FI double peval(const double* c, int deg, double t)
{
    double f = c[deg];
    for (int i = deg - 1; i >= 0; i--) f = MAD(f, t, c[i]);
    return f;
}
FI double bisect_root(const double* c, int deg, double a, double b, double fa)
{
    for (int it = 0; it < 200; it++) {
        double m = 0.5 * (a + b);
        if (m <= a || m >= b) break;
        double fm = peval(c, deg, m);
        if (fm == 0.0) return m;
        if ((fm < 0) == (fa < 0)) { a = m; fa = fm; } else b = m;
    }
    return a;
}
// all real roots of c (degree n <= 6) in [0,1], ascending; returns the count
FI int roots01(const double* c, int n, double* out)
{
    // derivative cascade: level k holds the (n-k)-th derivative (degree k)
    double D[7][7];
    for (int i = 0; i <= n; i++) D[n][i] = c[i];
    for (int k = n - 1; k >= 1; k--)
        for (int i = 0; i <= k; i++) D[k][i] = D[k + 1][i + 1] * (i + 1);
    double r[7];
    int nr = 0;
    // degree-1 level
    {
        double t = -D[1][0] / D[1][1];
        if (t >= 0.0 && t <= 1.0) r[nr++] = t;
    }
    for (int k = 2; k <= n; k++) {
        double br[9];
        int nb = 0;
        br[nb++] = 0.0;
        for (int i = 0; i < nr; i++) if (r[i] > br[nb - 1] && r[i] < 1.0) br[nb++] = r[i];
        br[nb++] = 1.0;
        double nrr[7];
        int m = 0;
        for (int s = 0; s + 1 < nb; s++) {
            double a = br[s], b = br[s + 1];
            double fa = peval(D[k], k, a), fb = peval(D[k], k, b);
            if (fa == 0.0) { if (m == 0 || nrr[m - 1] != a) nrr[m++] = a; continue; }
            if (fb == 0.0) continue;
            if ((fa < 0) == (fb < 0)) continue;
            nrr[m++] = bisect_root(D[k], k, a, b, fa);
        }
        if (peval(D[k], k, 1.0) == 0.0 && (m == 0 || nrr[m - 1] != 1.0)) nrr[m++] = 1.0;
        for (int i = 0; i < m; i++) r[i] = nrr[i];
        nr = m;
    }
    for (int i = 0; i < nr; i++) out[i] = r[i];
    return nr;
}
FI void push_ivl(IvlSet& s, double t1, double t2, const double* c, int deg, bool pos)
{
    t1 = fmin(1.0, fmax(0.0, t1));
    t2 = fmin(1.0, fmax(0.0, t2));
    double f = peval(c, deg, (t2 + t1) / 2);
    if ((pos && f >= 0) || (!pos && f <= 0)) { if (s.n < 8) { s.v[s.n].l = t1; s.v[s.n].u = t2; s.n++; } }
}
// intervals of [0,1] on which the polynomial is >= 0 (pos) or <= 0 (CTCD.cpp:92-163)
FI void sign_intervals(const double* cin, int n, bool pos, IvlSet& s)
{
    double c[7];
    s.n = 0;
    double mx = 0;
    for (int i = 0; i <= n; i++) mx = fmax(mx, fabs(cin[i]));
    for (int i = 0; i <= n; i++) c[i] = mx != 0 ? cin[i] / mx : cin[i];
    int deg = n;
    while (deg > 0 && c[deg] == 0) deg--;
    if (deg > 2) {
        // CTCD::couldHaveRoots (CTCD.cpp:77-90): bound of p on [0,1]
        double res = 0;
        for (int i = deg; i >= 1; i--) if ((pos && c[i] > 0) || (!pos && c[i] < 0)) res += c[i];
        res += c[0];
        if ((pos && res < 0) || (!pos && res > 0)) return;
        double r[7];
        int nr = roots01(c, deg, r);
        if (nr == 0) { push_ivl(s, 0.0, 1.0, c, deg, pos); return; }
        push_ivl(s, 0.0, r[0], c, deg, pos);
        for (int i = 0; i + 1 < nr; i++) push_ivl(s, r[i], r[i + 1], c, deg, pos);
        push_ivl(s, r[nr - 1], 1.0, c, deg, pos);
    } else if (deg == 2) {
        double a = c[2], b = c[1], cc = c[0], D = MAD(b, b, -(4 * a * cc));
        if (D < 0) { push_ivl(s, 0.0, 1.0, c, deg, pos); return; }
        double q = -0.5 * MAD((b < 0 ? -1.0 : 1.0), sqrt(D), b);
        double r0 = q / a, r1 = cc / q;
        if (r0 > r1) { double t = r0; r0 = r1; r1 = t; }
        if (r0 >= 0) push_ivl(s, 0, r0, c, deg, pos);
        if (!((r0 < 0 && r1 < 0) || (r0 > 1.0 && r1 > 1.0))) push_ivl(s, r0, r1, c, deg, pos);
        if (r1 <= 1.0) push_ivl(s, r1, 1.0, c, deg, pos);
    } else if (deg == 1) {
        double t = -c[0] / c[1];
        if (t >= 0) push_ivl(s, 0, t, c, deg, pos);
        if (t <= 1.0) push_ivl(s, t, 1.0, c, deg, pos);
    } else {
        if ((!pos && c[0] <= 0) || (pos && c[0] >= 0)) { s.v[0].l = 0; s.v[0].u = 1.0; s.n = 1; }
    }
}
// linear vector polynomial x + t v; products accumulate into coefficient arrays
FI void lin_cross(v3 x2, v3 v2, v3 x3, v3 v3_, v3* n)
{
    n[0] = crossm(x2, x3);
    n[1] = crossm(x2, v3_) + crossm(v2, x3);
    n[2] = crossm(v2, v3_);
}
// (x1 + t v1) . ((x2 + t v2) x (x3 + t v3)), cubic
FI void triple_poly(v3 x1, v3 v1, const v3* n, double* f)
{
    f[0] = dotm(x1, n[0]);
    f[1] = dotm(x1, n[1]) + dotm(v1, n[0]);
    f[2] = dotm(x1, n[2]) + dotm(v1, n[1]);
    f[3] = dotm(v1, n[2]);
}
// >= 0 set of the triple product (CTCD.cpp:201-215)
FI void plane_set(v3 x1, v3 x2, v3 x3, v3 v1, v3 v2, v3 v3_, IvlSet& s)
{
    v3 n[3];
    lin_cross(x2, v2, x3, v3_, n);
    double f[4];
    triple_poly(x1, v1, n, f);
    sign_intervals(f, 3, true, s);
}
// f^2 - minD |n|^2 <= 0 set (CTCD.cpp:217-243)
FI void dist_set(v3 x1, v3 x2, v3 x3, v3 v1, v3 v2, v3 v3_, double minD, IvlSet& s)
{
    v3 n[3];
    lin_cross(x2, v2, x3, v3_, n);
    double f[4];
    triple_poly(x1, v1, n, f);
    double p[7];
    for (int i = 0; i < 7; i++) p[i] = 0.0;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) p[i + j] = MAD(f[i], f[j], p[i + j]);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p[i + j] = MAD(-minD, dotm(n[i], n[j]), p[i + j]);
    sign_intervals(p, 6, false, s);
}
// (e1.e0)(q.e0) - |e0|^2 (q.e1) >= 0 with e0 = x1+tv1, e1 = x2+tv2, q = x3+tv3 (CTCD.cpp:165-199)
FI void bary_set(v3 x1, v3 x2, v3 x3, v3 v1, v3 v2, v3 v3_, IvlSet& s)
{
    double aa[3] = {dotm(x1, x1), 2 * dotm(x1, v1), dotm(v1, v1)};
    double ba[3] = {dotm(x2, x1), dotm(x2, v1) + dotm(v2, x1), dotm(v2, v1)};
    double cb[3] = {dotm(x3, x2), dotm(x3, v2) + dotm(v3_, x2), dotm(v3_, v2)};
    double ca[3] = {dotm(x3, x1), dotm(x3, v1) + dotm(v3_, x1), dotm(v3_, v1)};
    double p[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p[i + j] += MAD(ba[i], ca[j], -(aa[i] * cb[j]));
    sign_intervals(p, 4, true, s);
}
FI bool ivl_overlap(Ivl a, Ivl b) { return !(a.l > b.u || b.l > a.u); }

// vertexFaceCTCD (CTCD.cpp:383-472)
FI bool vf_ctcd(v3 q0s, v3 q1s, v3 q2s, v3 q3s, v3 q0e, v3 q1e, v3 q2e, v3 q3e, double eta, double& t)
{
    double minD = eta * eta;
    v3 v0 = q0e - q0s, v1 = q1e - q1s, v2 = q2e - q2s, v3_ = q3e - q3s;
    IvlSet e1, e2, e3, cop;
    plane_set(q0s - q1s, crossm(q3s - q1s, q2s - q1s), q3s - q1s, v0 - v1, crossm(v3_ - v1, v2 - v1), v3_ - v1, e1);
    if (!e1.n) return false;
    plane_set(q0s - q2s, crossm(q1s - q2s, q3s - q2s), q1s - q2s, v0 - v2, crossm(v1 - v2, v3_ - v2), v1 - v2, e2);
    if (!e2.n) return false;
    plane_set(q0s - q3s, crossm(q2s - q3s, q1s - q3s), q2s - q3s, v0 - v3_, crossm(v2 - v3_, v1 - v3_), v2 - v3_, e3);
    if (!e3.n) return false;
    dist_set(q0s - q1s, q2s - q1s, q3s - q1s, v0 - v1, v2 - v1, v3_ - v1, minD, cop);
    if (!cop.n) return false;
    bool col = false;
    double mint = 1.0;
    for (int i = 0; i < cop.n; i++) for (int j = 0; j < e1.n; j++) for (int k = 0; k < e2.n; k++) for (int l = 0; l < e3.n; l++) {
        Ivl iv[4] = {cop.v[i], e1.v[j], e2.v[k], e3.v[l]};
        bool ok = true;
        for (int p = 0; p < 4; p++) for (int q = p + 1; q < 4; q++) ok = ok && ivl_overlap(iv[p], iv[q]);
        if (ok) {
            double lo = 0.0;
            for (int p = 0; p < 4; p++) lo = fmax(lo, iv[p].l);
            mint = fmin(mint, lo);
            col = true;
        }
    }
    if (col) t = mint;
    return col;
}
// edgeEdgeCTCD (CTCD.cpp:245-381), arguments (q0, p0, q1, p1) = (a0, a1, b0, b1)
FI bool ee_ctcd(v3 q0s, v3 p0s, v3 q1s, v3 p1s, v3 q0e, v3 p0e, v3 q1e, v3 p1e, double eta, double& t)
{
    double minD = eta * eta;
    v3 vp0 = p0e - p0s, vp1 = p1e - p1s, vq0 = q0e - q0s, vq1 = q1e - q1s;
    IvlSet raw, a0, a1, b0, b1;
    dist_set(p0s - p1s, p0s - q0s, p1s - q1s, vp0 - vp1, vp0 - vq0, vp1 - vq1, minD, raw);
    Ivl cop[8], par[8];
    int nc = 0, np = 0;
    for (int i = 0; i < raw.n; i++) {
        double midt = (raw.v[i].u + raw.v[i].l) / 2;
        v3 b1 = q0s - p0s, d1 = vq0 - vp0, b2 = q1s - p1s, d2 = vq1 - vp1;
        v3 y1 = mk(MAD(midt, d1.x, b1.x), MAD(midt, d1.y, b1.y), MAD(midt, d1.z, b1.z));
        v3 y2 = mk(MAD(midt, d2.x, b2.x), MAD(midt, d2.y, b2.y), MAD(midt, d2.z, b2.z));
        if (sqrt(n2m(crossm(y1, y2))) < 1e-8) par[np++] = raw.v[i]; else cop[nc++] = raw.v[i];
    }
    if (!nc) return false;
    v3 x10 = p1s - q1s, v10 = vp1 - vq1;
    bary_set(x10, p0s - q0s, q0s - q1s, v10, vp0 - vq0, vq0 - vq1, a0);
    if (!a0.n) return false;
    bary_set(x10, q0s - p0s, p0s - q1s, v10, vq0 - vp0, vp0 - vq1, a1);
    if (!a1.n) return false;
    x10 = p0s - q0s; v10 = vp0 - vq0;
    bary_set(x10, p1s - q1s, q1s - q0s, v10, vp1 - vq1, vq1 - vq0, b0);
    if (!b0.n) return false;
    bary_set(x10, q1s - p1s, p1s - q0s, v10, vq1 - vp1, vp1 - vq0, b1);
    if (!b1.n) return false;
    bool col = false;
    double mint = 1.0;
    for (int i = 0; i < nc; i++) for (int j = 0; j < a0.n; j++) for (int k = 0; k < a1.n; k++)
    for (int l = 0; l < b0.n; l++) for (int m = 0; m < b1.n; m++) {
        Ivl iv[5] = {cop[i], a0.v[j], a1.v[k], b0.v[l], b1.v[m]};
        bool ok = true;
        for (int p = 0; p < 5; p++) for (int q = p + 1; q < 5; q++) ok = ok && ivl_overlap(iv[p], iv[q]);
        if (!ok) continue;
        Ivl is = {0.0, 1.0};
        for (int p = 0; p < 5; p++) { is.l = fmax(iv[p].l, is.l); is.u = fmin(iv[p].u, is.u); }
        bool skip = false;
        for (int p = 0; p < np; p++) skip = skip || ivl_overlap(is, par[p]);
        if (!skip) { mint = fmin(mint, is.l); col = true; }
    }
    if (col) t = mint;
    return col;
}
// point_triangle_ccd (wrapper.cpp:254-290)
FI double pt_ccd(v3 p, v3 t0, v3 t1, v3 t2, v3 dp, v3 dt0, v3 dt1, v3 dt2, double eta, double dist2)
{
    double toc = 1.0;
    double maxDisp = sqrt(n2m(dp)) + sqrt(fmax(fmax(n2m(dt0), n2m(dt1)), n2m(dt2)));
    if (maxDisp != 0) {
        double lb = (1 - eta) * sqrt(dist2) / maxDisp;
        if (lb < 1) {
            double t;
            if (vf_ctcd(p, t0, t1, t2, p + dp, t0 + dt0, t1 + dt1, t2 + dt2, eta * sqrt(dist2), t)) {
                toc = t;
                if (toc < lb) toc = lb;
            }
        }
    }
    return toc;
}
// edge_edge_ccd (wrapper.cpp:292-328)
FI double ee_ccd(v3 a0, v3 a1, v3 b0, v3 b1, v3 da0, v3 da1, v3 db0, v3 db1, double eta, double dist2)
{
    double toc = 1.0;
    double maxDisp = sqrt(fmax(n2m(da0), n2m(da1))) + sqrt(fmax(n2m(db0), n2m(db1)));
    if (maxDisp != 0) {
        double lb = (1 - eta) * sqrt(dist2) / maxDisp;
        if (lb < 1) {
            double t;
            if (ee_ctcd(a0, a1, b0, b1, a0 + da0, a1 + da1, b0 + db0, b1 + db1, eta * sqrt(dist2), t)) {
                toc = t;
                if (toc < lb) toc = lb;
            }
        }
    }
    return toc;
}
