/* ccd_body.h -- the CCD of the oracle, restated from EVCTCD/CTCD.cpp:10-472
 * and wrapper.cpp:254-328, included twice by ipc_oracle.c: with
 * MAD(a, b, c) = a * b + c (two roundings, strict IEEE) and with
 * MAD(a, b, c) = fma(a, b, c) (one rounding), the names suffixed by CCDF.
 * Every product that feeds an addition in the CCD is written MAD, including
 * in the dot and cross products below, which stand in for vdot / vcross /
 * vnorm2 here.  Real roots on [0,1] are isolated by derivative recursion and
 * bisection instead of Jenkins-Traub; polynomial coefficients come from
 * polynomial arithmetic on vector polynomials.  Test infrastructure. */
static inline double CCDF(dotm)(const double* a, const double* b) { return MAD(a[2], b[2], MAD(a[1], b[1], a[0] * b[0])); }
static inline void CCDF(crossm)(const double* a, const double* b, double* o)
{
    double t0 = MAD(a[1], b[2], -(a[2] * b[1]));
    double t1 = MAD(a[2], b[0], -(a[0] * b[2]));
    double t2 = MAD(a[0], b[1], -(a[1] * b[0]));
    o[0] = t0; o[1] = t1; o[2] = t2;
}
static inline double CCDF(n2m)(const double* a) { return CCDF(dotm)(a, a); }
#define vdot CCDF(dotm)
#define vcross CCDF(crossm)
#define vnorm2 CCDF(n2m)
#define peval CCDF(peval)
#define roots01 CCDF(roots01)
#define could_have_roots CCDF(could_have_roots)
#define check_interval CCDF(check_interval)
#define find_intervals CCDF(find_intervals)
#define lin_cross3 CCDF(lin_cross3)
#define triple3 CCDF(triple3)
#define plane_set CCDF(plane_set)
#define dist_set CCDF(dist_set)
#define bary_set CCDF(bary_set)
#define ivl_overlap CCDF(ivl_overlap)
#define vf_ctcd CCDF(vf_ctcd)
#define ee_ctcd CCDF(ee_ctcd)


/* polynomials: c[0] constant term ... c[deg] leading */
static double peval(const double* c, int deg, double t)
{
    double f = c[deg];
    for (int i = deg - 1; i >= 0; i--) f = MAD(f, t, c[i]);
    return f;
}
/* real roots of c (degree deg) in [0,1], ascending, deduplicated */
static int roots01(const double* c, int deg, double* out)
{
    if (deg <= 0) return 0;
    if (deg == 1) {
        double r = -c[0] / c[1];
        if (r >= 0.0 && r <= 1.0) { out[0] = r; return 1; }
        return 0;
    }
    double dc[8], crit[8];
    for (int i = 1; i <= deg; i++) dc[i - 1] = c[i] * i;
    int nc = roots01(dc, deg - 1, crit);
    double br[10];
    int nb = 0;
    br[nb++] = 0.0;
    for (int i = 0; i < nc; i++) if (crit[i] > br[nb - 1] && crit[i] < 1.0) br[nb++] = crit[i];
    br[nb++] = 1.0;
    int n = 0;
    for (int k = 0; k + 1 < nb; k++) {
        double a = br[k], b = br[k + 1], fa = peval(c, deg, a), fb = peval(c, deg, b);
        if (fa == 0.0) { if (n == 0 || out[n - 1] != a) out[n++] = a; continue; }
        if (fb == 0.0) continue;
        if ((fa < 0) == (fb < 0)) continue;
        for (int it = 0; it < 200; it++) {
            double m = 0.5 * (a + b);
            if (m <= a || m >= b) break;
            double fm = peval(c, deg, m);
            if (fm == 0.0) { a = b = m; break; }
            if ((fm < 0) == (fa < 0)) { a = m; fa = fm; } else b = m;
        }
        out[n++] = a;
    }
    if (peval(c, deg, 1.0) == 0.0 && (n == 0 || out[n - 1] != 1.0)) out[n++] = 1.0;
    return n;
}
/* could the polynomial take the wanted sign on [0,1]? (CTCD.cpp:77-90), hi
 * is coefficients highest degree first */
static int could_have_roots(const double* hi, int degree, int pos)
{
    double result = 0;
    if ((pos && hi[0] > 0) || (!pos && hi[0] < 0)) result = hi[0];
    for (int i = 1; i < degree; i++) if ((pos && hi[i] > 0) || (!pos && hi[i] < 0)) result += hi[i];
    result += hi[degree];
    return !((pos && result < 0) || (!pos && result > 0));
}
static void check_interval(double t1, double t2, const double* c, int deg, ivlset* s, int pos)
{
    t1 = dmin(1.0, dmax(0.0, t1));
    t2 = dmin(1.0, dmax(0.0, t2));
    double f = peval(c, deg, (t2 + t1) / 2);
    if ((pos && f >= 0) || (!pos && f <= 0)) { s->v[s->n].l = t1; s->v[s->n].u = t2; s->n++; }
}
/* intervals of [0,1] where the polynomial is >= 0 (pos) or <= 0 (CTCD.cpp:92-163) */
static void find_intervals(const double* cin, int n, int pos, ivlset* s)
{
    double c[8];
    s->n = 0;
    double mx = 0;
    for (int i = 0; i <= n; i++) mx = dmax(mx, fabs(cin[i]));
    for (int i = 0; i <= n; i++) c[i] = mx != 0 ? cin[i] / mx : cin[i];
    int deg = n;
    while (deg > 0 && c[deg] == 0) deg--;
    double r[8];
    int nr;
    if (deg > 2) {
        double hi[8];
        for (int i = 0; i <= deg; i++) hi[i] = c[deg - i];
        if (!could_have_roots(hi, deg, pos)) return;
        nr = roots01(c, deg, r);
    } else if (deg == 2) {
        double a = c[2], b = c[1], cc = c[0];
        double D = MAD(b, b, -(4 * a * cc));
        nr = 0;
        if (D >= 0) {
            double q = -0.5 * MAD((b < 0 ? -1.0 : 1.0), sqrt(D), b);
            double r0 = q / a, r1 = cc / q;
            if (r0 > r1) { double t = r0; r0 = r1; r1 = t; }
            /* keep the roots that split [0,1]; outside roots clamp away */
            double all[2] = {r0, r1};
            for (int i = 0; i < 2; i++) r[nr++] = all[i];
        }
        /* mirror the generic splitting below with clamping */
        if (nr == 0) { check_interval(0.0, 1.0, c, deg, s, pos); return; }
        if (r[0] >= 0) check_interval(0, r[0], c, deg, s, pos);
        if (!((r[0] < 0 && r[1] < 0) || (r[0] > 1.0 && r[1] > 1.0))) check_interval(r[0], r[1], c, deg, s, pos);
        if (r[1] <= 1.0) check_interval(r[1], 1.0, c, deg, s, pos);
        return;
    } else if (deg == 1) {
        double t = -c[0] / c[1];
        if (t >= 0) check_interval(0, t, c, deg, s, pos);
        if (t <= 1.0) check_interval(t, 1.0, c, deg, s, pos);
        return;
    } else {
        if ((!pos && c[0] <= 0) || (pos && c[0] >= 0)) { s->v[0].l = 0; s->v[0].u = 1.0; s->n = 1; }
        return;
    }
    if (nr == 0) { check_interval(0.0, 1.0, c, deg, s, pos); return; }
    check_interval(0, r[0], c, deg, s, pos);
    for (int i = 0; i + 1 < nr; i++) check_interval(r[i], r[i + 1], c, deg, s, pos);
    check_interval(r[nr - 1], 1.0, c, deg, s, pos);
}
/* Polynomial coefficients of the three CTCD sign tests along the linear
 * trajectory y(t) = x + t v.  The expressions are written out term by term
 * (cross products of the start/velocity parts, then dot products) in the
 * same order as the product library, so both strict-IEEE implementations
 * round identically on degenerate (parallel, touching) configurations. */
static void lin_cross3(const double* x2, const double* v2, const double* x3, const double* v3, double n[3][3])
{
    double a[3], b[3];
    vcross(x2, x3, n[0]);
    vcross(x2, v3, a); vcross(v2, x3, b); vadd(a, b, n[1]);
    vcross(v2, v3, n[2]);
}
static void triple3(const double* x1, const double* v1, double n[3][3], double* f)
{
    f[0] = vdot(x1, n[0]);
    f[1] = vdot(x1, n[1]) + vdot(v1, n[0]);
    f[2] = vdot(x1, n[2]) + vdot(v1, n[1]);
    f[3] = vdot(v1, n[2]);
}
/* sign-set of (x1 + t v1) . ((x2 + t v2) x (x3 + t v3)) >= 0 (CTCD.cpp:201-215) */
static void plane_set(const double* x1, const double* x2, const double* x3, const double* v1, const double* v2, const double* v3, ivlset* s)
{
    double n[3][3], f[4];
    lin_cross3(x2, v2, x3, v3, n);
    triple3(x1, v1, n, f);
    find_intervals(f, 3, 1, s);
}
/* distance-to-plane^2 <= minD2 set (CTCD.cpp:217-243):
 * (x1.(x2 x x3))^2 - minD2 |x2 x x3|^2 <= 0 along the linear trajectory */
static void dist_set(const double* x1, const double* x2, const double* x3, const double* v1, const double* v2, const double* v3, double minD2, ivlset* s)
{
    double n[3][3], f[4], p[7] = {0, 0, 0, 0, 0, 0, 0};
    lin_cross3(x2, v2, x3, v3, n);
    triple3(x1, v1, n, f);
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) p[i + j] = MAD(f[i], f[j], p[i + j]);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p[i + j] = MAD(-minD2, vdot(n[i], n[j]), p[i + j]);
    find_intervals(p, 6, 0, s);
}
/* barycentric side set (CTCD.cpp:165-199): with e0 = x1 + t v1,
 * e1 = x2 + t v2 and q = x3 + t v3 the polynomial is
 *   (e1.e0)(q.e0) - |e0|^2 (q.e1)  >= 0 */
static void bary_set(const double* x1, const double* x2, const double* x3, const double* v1, const double* v2, const double* v3, ivlset* s)
{
    double aa[3] = {vdot(x1, x1), 2 * vdot(x1, v1), vdot(v1, v1)};
    double ba[3] = {vdot(x2, x1), vdot(x2, v1) + vdot(v2, x1), vdot(v2, v1)};
    double cb[3] = {vdot(x3, x2), vdot(x3, v2) + vdot(v3, x2), vdot(v3, v2)};
    double ca[3] = {vdot(x3, x1), vdot(x3, v1) + vdot(v3, x1), vdot(v3, v1)};
    double p[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p[i + j] += MAD(ba[i], ca[j], -(aa[i] * cb[j]));
    find_intervals(p, 4, 1, s);
}
static int ivl_overlap(ivl a, ivl b) { return !(a.l > b.u || b.l > a.u); }

/* vertexFaceCTCD (CTCD.cpp:383-472): earliest t at which q0 is within eta of
 * the (moving) triangle's plane while inside its three edge slabs */
static int vf_ctcd(const double* q0s, const double* q1s, const double* q2s, const double* q3s,
                   const double* q0e, const double* q1e, const double* q2e, const double* q3e, double eta, double* t)
{
    double minD = eta * eta;
    double v0[3], v1[3], v2[3], v3[3];
    vsub(q0e, q0s, v0); vsub(q1e, q1s, v1); vsub(q2e, q2s, v2); vsub(q3e, q3s, v3);
    ivlset e1, e2, e3, cop;
    double x10[3], v10[3], x20[3], v20[3], x30[3], v30[3], a[3], b[3];
    vsub(q0s, q1s, x10); vsub(v0, v1, v10);
    vsub(q3s, q1s, a); vsub(q2s, q1s, b); vcross(a, b, x20);
    vsub(v3, v1, a); vsub(v2, v1, b); vcross(a, b, v20);
    vsub(q3s, q1s, x30); vsub(v3, v1, v30);
    plane_set(x10, x20, x30, v10, v20, v30, &e1);
    if (!e1.n) return 0;
    vsub(q0s, q2s, x10); vsub(v0, v2, v10);
    vsub(q1s, q2s, a); vsub(q3s, q2s, b); vcross(a, b, x20);
    vsub(v1, v2, a); vsub(v3, v2, b); vcross(a, b, v20);
    vsub(q1s, q2s, x30); vsub(v1, v2, v30);
    plane_set(x10, x20, x30, v10, v20, v30, &e2);
    if (!e2.n) return 0;
    vsub(q0s, q3s, x10); vsub(v0, v3, v10);
    vsub(q2s, q3s, a); vsub(q1s, q3s, b); vcross(a, b, x20);
    vsub(v2, v3, a); vsub(v1, v3, b); vcross(a, b, v20);
    vsub(q2s, q3s, x30); vsub(v2, v3, v30);
    plane_set(x10, x20, x30, v10, v20, v30, &e3);
    if (!e3.n) return 0;
    vsub(q0s, q1s, x10); vsub(q2s, q1s, x20); vsub(q3s, q1s, x30);
    vsub(v0, v1, v10); vsub(v2, v1, v20); vsub(v3, v1, v30);
    dist_set(x10, x20, x30, v10, v20, v30, minD, &cop);
    if (!cop.n) return 0;
    int col = 0;
    double mint = 1.0;
    for (int i = 0; i < cop.n; i++) for (int j = 0; j < e1.n; j++) for (int k = 0; k < e2.n; k++) for (int l = 0; l < e3.n; l++) {
        ivl iv[4] = {cop.v[i], e1.v[j], e2.v[k], e3.v[l]};
        int ok = 1;
        for (int p = 0; p < 4 && ok; p++) for (int q = p + 1; q < 4; q++) if (!ivl_overlap(iv[p], iv[q])) { ok = 0; break; }
        if (ok) {
            double lo = 0.0;
            for (int p = 0; p < 4; p++) lo = dmax(lo, iv[p].l);
            mint = dmin(mint, lo);
            col = 1;
        }
    }
    if (col) *t = mint;
    return col;
}
/* edgeEdgeCTCD (CTCD.cpp:245-381); arguments in the reference's order
 * (q0, p0, q1, p1) = (a0, a1, b0, b1) */
static int ee_ctcd(const double* q0s, const double* p0s, const double* q1s, const double* p1s,
                   const double* q0e, const double* p0e, const double* q1e, const double* p1e, double eta, double* t)
{
    double minD = eta * eta;
    double x10[3], x20[3], x30[3], vp0[3], vp1[3], vq0[3], vq1[3], v10[3], v20[3], v30[3];
    vsub(p0s, p1s, x10); vsub(p0s, q0s, x20); vsub(p1s, q1s, x30);
    vsub(p0e, p0s, vp0); vsub(p1e, p1s, vp1); vsub(q0e, q0s, vq0); vsub(q1e, q1s, vq1);
    vsub(vp0, vp1, v10); vsub(vp0, vq0, v20); vsub(vp1, vq1, v30);
    ivlset raw, cop, par, a0, a1, b0, b1;
    dist_set(x10, x20, x30, v10, v20, v30, minD, &raw);
    cop.n = par.n = 0;
    for (int i = 0; i < raw.n; i++) {
        double midt = (raw.v[i].u + raw.v[i].l) / 2, y1[3], y2[3], d1[3], d2[3], cr[3];
        vsub(q0s, p0s, y1); vsub(vq0, vp0, d1);
        vsub(q1s, p1s, y2); vsub(vq1, vp1, d2);
        for (int d = 0; d < 3; d++) { y1[d] = MAD(midt, d1[d], y1[d]); y2[d] = MAD(midt, d2[d], y2[d]); }
        vcross(y1, y2, cr);
        if (sqrt(vnorm2(cr)) < 1e-8) par.v[par.n++] = raw.v[i]; else cop.v[cop.n++] = raw.v[i];
    }
    if (!cop.n) return 0;
    vsub(p1s, q1s, x10); vsub(vp1, vq1, v10);
    vsub(p0s, q0s, x20); vsub(vp0, vq0, v20);
    vsub(q0s, q1s, x30); vsub(vq0, vq1, v30);
    bary_set(x10, x20, x30, v10, v20, v30, &a0);
    if (!a0.n) return 0;
    vsub(q0s, p0s, x20); vsub(vq0, vp0, v20);
    vsub(p0s, q1s, x30); vsub(vp0, vq1, v30);
    bary_set(x10, x20, x30, v10, v20, v30, &a1);
    if (!a1.n) return 0;
    vsub(p0s, q0s, x10); vsub(vp0, vq0, v10);
    vsub(p1s, q1s, x20); vsub(vp1, vq1, v20);
    vsub(q1s, q0s, x30); vsub(vq1, vq0, v30);
    bary_set(x10, x20, x30, v10, v20, v30, &b0);
    if (!b0.n) return 0;
    vsub(q1s, p1s, x20); vsub(vq1, vp1, v20);
    vsub(p1s, q0s, x30); vsub(vp1, vq0, v30);
    bary_set(x10, x20, x30, v10, v20, v30, &b1);
    if (!b1.n) return 0;
    int col = 0;
    double mint = 1.0;
    for (int i = 0; i < cop.n; i++) for (int j = 0; j < a0.n; j++) for (int k = 0; k < a1.n; k++)
    for (int l = 0; l < b0.n; l++) for (int m = 0; m < b1.n; m++) {
        ivl iv[5] = {cop.v[i], a0.v[j], a1.v[k], b0.v[l], b1.v[m]};
        int ok = 1;
        for (int p = 0; p < 5 && ok; p++) for (int q = p + 1; q < 5; q++) if (!ivl_overlap(iv[p], iv[q])) { ok = 0; break; }
        if (!ok) continue;
        ivl is = {0.0, 1.0};
        for (int p = 0; p < 5; p++) { is.l = dmax(iv[p].l, is.l); is.u = dmin(iv[p].u, is.u); }
        int skip = 0;
        for (int p = 0; p < par.n; p++) if (ivl_overlap(is, par.v[p])) { skip = 1; break; }
        if (!skip) { mint = dmin(mint, is.l); col = 1; }
    }
    if (col) *t = mint;
    return col;
}
/* point_triangle_ccd (wrapper.cpp:254-290) */
static double CCDF(pt_ccd)(const double* p, const double* t0, const double* t1, const double* t2,
                          const double* dp, const double* dt0, const double* dt1, const double* dt2, double eta, double dist2)
{
    double toc = 1.0;
    double m = dmax(dmax(vnorm2(dt0), vnorm2(dt1)), vnorm2(dt2));
    double maxDisp = sqrt(vnorm2(dp)) + sqrt(m);
    if (maxDisp != 0) {
        double lb = (1 - eta) * sqrt(dist2) / maxDisp;
        if (lb < 1) {
            double pe[3], a[3], b[3], c[3], t;
            vadd(p, dp, pe); vadd(t0, dt0, a); vadd(t1, dt1, b); vadd(t2, dt2, c);
            if (vf_ctcd(p, t0, t1, t2, pe, a, b, c, eta * sqrt(dist2), &t)) {
                toc = t;
                if (toc < lb) toc = lb;
            }
        }
    }
    return toc;
}
/* edge_edge_ccd (wrapper.cpp:292-328) */
static double CCDF(ee_ccd)(const double* a0, const double* a1, const double* b0, const double* b1,
                          const double* da0, const double* da1, const double* db0, const double* db1, double eta, double dist2)
{
    double toc = 1.0;
    double maxDisp = sqrt(dmax(vnorm2(da0), vnorm2(da1))) + sqrt(dmax(vnorm2(db0), vnorm2(db1)));
    if (maxDisp != 0) {
        double lb = (1 - eta) * sqrt(dist2) / maxDisp;
        if (lb < 1) {
            double A0[3], A1[3], B0[3], B1[3], t;
            vadd(a0, da0, A0); vadd(a1, da1, A1); vadd(b0, db0, B0); vadd(b1, db1, B1);
            if (ee_ctcd(a0, a1, b0, b1, A0, A1, B0, B1, eta * sqrt(dist2), &t)) {
                toc = t;
                if (toc < lb) toc = lb;
            }
        }
    }
    return toc;
}

#undef vdot
#undef vcross
#undef vnorm2
#undef peval
#undef roots01
#undef could_have_roots
#undef check_interval
#undef find_intervals
#undef lin_cross3
#undef triple3
#undef plane_set
#undef dist_set
#undef bary_set
#undef ivl_overlap
#undef vf_ctcd
#undef ee_ctcd
