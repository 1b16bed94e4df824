// fipc_broad.hip -- spatial-hash broad phase and everything built on it:
// constraint search (NEWTON_IPC.py:159-357), CCD step filters (:377-468) and
// the collision check (:855-886).
//
// Grid: every inserted item is entered in each cell its box overlaps
// (cell size = 1 / inv_dx, as the reference's hash).  Entries are grouped by
// a hashed bucket; each entry keeps its exact 63-bit cell key and three bits
// telling whether the cell is the item's lowest cell along x, y, z.  A query
// walks the cells of its own box in chunks of QCHUNK cells (one thread per
// chunk, so huge boxes do not serialise a wave) and accepts an entry only in
// the first cell common to both boxes, so every pair is tested exactly once.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <vector>
#include "fipc_ctx.h"

using namespace fipc;

#define TPB 256
#define QCHUNK 32
static inline int nblk(long n, int t = TPB) { return (int)((n + t - 1) / t); }

__device__ __forceinline__ v3 ldv(const double* a, int i) { return v3{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

struct CellRange { int lo[3], hi[3]; };
__device__ __forceinline__ CellRange cells_of(v3 lo, v3 hi, double inv)
{
    CellRange r;
    r.lo[0] = (int)floor(lo.x * inv); r.lo[1] = (int)floor(lo.y * inv); r.lo[2] = (int)floor(lo.z * inv);
    r.hi[0] = (int)floor(hi.x * inv); r.hi[1] = (int)floor(hi.y * inv); r.hi[2] = (int)floor(hi.z * inv);
    return r;
}
__device__ __forceinline__ long long ncells(const CellRange& r)
{
    return (long long)(r.hi[0] - r.lo[0] + 1) * (r.hi[1] - r.lo[1] + 1) * (r.hi[2] - r.lo[2] + 1);
}
__device__ __forceinline__ unsigned long long cell_key(int i, int j, int k)
{
    const long long B = 1 << 20;
    return ((unsigned long long)(i + B) << 42) | ((unsigned long long)(j + B) << 21) | (unsigned long long)(k + B);
}
// bucket of a cell: splitmix64 finaliser of the exact cell key, so that the
// axis-aligned rows and planes of cells that lattice meshes produce spread
// evenly over a power-of-two table
__device__ __forceinline__ unsigned bucket_of(int i, int j, int k, unsigned M)
{
    unsigned long long z = cell_key(i, j, k) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (unsigned)z & (M - 1);
}

// ---- boxes ---------------------------------------------------------------
// kind of primitive: 0 boundary point, 1 boundary edge, 2 boundary triangle.
// sol != nullptr gives the swept box of x and x + sol.
struct BoxFn {
    const double* x;
    const double* sol;
    const int* idx;
    int kind;
    double infl;
    __device__ __forceinline__ void operator()(int i, v3& lo, v3& hi) const
    {
        int nvx = kind + 1;
        lo = mk(1e300, 1e300, 1e300);
        hi = mk(-1e300, -1e300, -1e300);
        for (int a = 0; a < nvx; a++) {
            int v = idx[nvx * i + a];
            v3 p = ldv(x, v);
            lo = vmin(lo, p);
            hi = vmax(hi, p);
            if (sol) {
                v3 q = p + ldv(sol, v);
                lo = vmin(lo, q);
                hi = vmax(hi, q);
            }
        }
        lo = lo - mk(infl, infl, infl);
        hi = hi + mk(infl, infl, infl);
    }
};

__global__ void k_grid_count(int n, BoxFn box, double inv, int* cnt, int* overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { cnt[n] = 0; return; }
    v3 lo, hi;
    box(i, lo, hi);
    long long c = ncells(cells_of(lo, hi, inv));
    if (c > (1 << 24) || c < 0) { atomicOr(overflow, 1); c = 0; }
    cnt[i] = (int)c;
}
// one thread per grid entry: the owning item is found by binary search in
// the scanned counts, so an item covering thousands of cells (a long plate
// edge) does not serialise a wave
__global__ void k_grid_fill(int n, int ne, BoxFn box, double inv, unsigned M, const int* __restrict__ off,
                            unsigned* ekey, int* eitem, unsigned long long* ecell)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid; else hi = mid;
    }
    int i = lo;
    v3 blo, bhi;
    box(i, blo, bhi);
    CellRange r = cells_of(blo, bhi, inv);
    int ny = r.hi[1] - r.lo[1] + 1, nz = r.hi[2] - r.lo[2] + 1;
    int l = e - off[i];
    int a = r.lo[0] + l / (ny * nz), b = r.lo[1] + (l / nz) % ny, c = r.lo[2] + l % nz;
    unsigned flags = (a == r.lo[0] ? 1u : 0u) | (b == r.lo[1] ? 2u : 0u) | (c == r.lo[2] ? 4u : 0u);
    ekey[e] = bucket_of(a, b, c, M);
    eitem[e] = (int)((unsigned)i | (flags << 28));
    ecell[e] = cell_key(a, b, c);
}
__global__ void k_grid_hist(int ne, const unsigned* ekey, int* bcount)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne) atomicAdd(&bcount[ekey[e]], 1);
}
__global__ void k_grid_scatter(int ne, const unsigned* ekey, const int* eitem, const unsigned long long* ecell,
                               const int* bstart, int* bfill, int* items, unsigned long long* cells)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    unsigned k = ekey[e];
    int pos = bstart[k] + atomicAdd(&bfill[k], 1);
    items[pos] = eitem[e];
    cells[pos] = ecell[e];
}
// number of QCHUNK-cell chunks of every query box
__global__ void k_query_chunks(int n, BoxFn box, double inv, int* cnt, int* overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { cnt[n] = 0; return; }
    v3 lo, hi;
    box(i, lo, hi);
    long long c = ncells(cells_of(lo, hi, inv));
    if (c > (1 << 26) || c < 0) { atomicOr(overflow, 1); c = 0; }
    cnt[i] = (int)((c + QCHUNK - 1) / QCHUNK);
}

// generic query: visit(q, item) is called once per (query, item) pair whose
// cell ranges overlap
template <class Visit>
__global__ void __launch_bounds__(TPB) k_grid_query(int nq, int nchunks, const int* __restrict__ qoff, BoxFn qbox,
                                                    double inv, unsigned M, const int* __restrict__ bstart,
                                                    const int* __restrict__ items,
                                                    const unsigned long long* __restrict__ cells, Visit visit)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nchunks) return;
    // query owning chunk t
    int lo = 0, hi = nq;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (qoff[mid] <= t) lo = mid; else hi = mid;
    }
    int q = lo, chunk = t - qoff[q];
    v3 blo, bhi;
    qbox(q, blo, bhi);
    CellRange r = cells_of(blo, bhi, inv);
    int ny = r.hi[1] - r.lo[1] + 1, nz = r.hi[2] - r.lo[2] + 1;
    long long total = ncells(r);
    long long l0 = (long long)chunk * QCHUNK, l1 = l0 + QCHUNK < total ? l0 + QCHUNK : total;
    for (long long l = l0; l < l1; l++) {
        int a = r.lo[0] + (int)(l / ((long long)ny * nz));
        int b = r.lo[1] + (int)((l / nz) % ny);
        int c = r.lo[2] + (int)(l % nz);
        unsigned long long key = cell_key(a, b, c);
        unsigned bk = bucket_of(a, b, c, M);
        int e0 = bstart[bk], e1 = bstart[bk + 1];
        for (int e = e0; e < e1; e++) {
            if (cells[e] != key) continue;
            unsigned it = (unsigned)items[e];
            unsigned fl = it >> 28;
            // first common cell: along each axis the cell is the lowest cell of the query or of the item
            if (!((a == r.lo[0] || (fl & 1u)) && (b == r.lo[1] || (fl & 2u)) && (c == r.lo[2] || (fl & 4u)))) continue;
            visit(q, (int)(it & 0x0fffffffu));
        }
    }
}

// build the grid over n items
static int grid_build(fipc_ctx* c, int n, const BoxFn& box)
{
    Grid& g = c->grid;
    STCHK(fipc_tm_begin(c, TM_GRID));
    STCHK(ensure(g.cnt, (size_t)n + 2));
    HIPCHK(hipMemsetAsync(c->flag.p, 0, 4, c->s));
    k_grid_count<<<nblk(n + 1), TPB, 0, c->s>>>(n, box, c->inv_dx, g.cnt.p, c->flag.p);
    HIPCHK(hipGetLastError());
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, g.cnt.p, g.cnt.p, n + 1, c->s));
    if (bytes > c->cub_tmp.cap) STCHK(ensure(c->cub_tmp, bytes));
    bytes = c->cub_tmp.cap;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, bytes, g.cnt.p, g.cnt.p, n + 1, c->s));
    int hv[2];
    HIPCHK(hipMemcpyAsync(&hv[0], g.cnt.p + n, 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(&hv[1], c->flag.p, 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (hv[1]) return FIPC_ERR_CAPACITY;
    int ne = hv[0];
    g.nentries = ne;
    unsigned M = 1024;
    while (M < (unsigned)ne * 2u && M < (1u << 30)) M <<= 1;
    g.M = M;
    STCHK(ensure(g.ekey, ne)); STCHK(ensure(g.eitem, ne));
    STCHK(ensure(g.bstart, (size_t)M + 1)); STCHK(ensure(g.bfill, M)); STCHK(ensure(g.items, ne));
    static_assert(sizeof(unsigned long long) == 8, "");
    // cell keys: entries and sorted copy share one allocation of 2*ne
    DVec<unsigned long long>& ck = c->keys_tmp;
    STCHK(ensure(ck, (size_t)ne * 2 + 2));
    unsigned long long* ecell = ck.p;
    unsigned long long* cells = ck.p + ne;
    if (ne) k_grid_fill<<<nblk(ne), TPB, 0, c->s>>>(n, ne, box, c->inv_dx, M, g.cnt.p, g.ekey.p, g.eitem.p, ecell);
    HIPCHK(hipMemsetAsync(g.bstart.p, 0, ((size_t)M + 1) * 4, c->s));
    HIPCHK(hipMemsetAsync(g.bfill.p, 0, (size_t)M * 4, c->s));
    if (ne) k_grid_hist<<<nblk(ne), TPB, 0, c->s>>>(ne, g.ekey.p, g.bstart.p);
    HIPCHK(hipGetLastError());
    bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, g.bstart.p, g.bstart.p, M + 1, c->s));
    if (bytes > c->cub_tmp.cap) STCHK(ensure(c->cub_tmp, bytes));
    bytes = c->cub_tmp.cap;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, bytes, g.bstart.p, g.bstart.p, M + 1, c->s));
    if (ne) k_grid_scatter<<<nblk(ne), TPB, 0, c->s>>>(ne, g.ekey.p, g.eitem.p, ecell, g.bstart.p, g.bfill.p, g.items.p, cells);
    HIPCHK(hipGetLastError());
    STCHK(fipc_tm_end(c, TM_GRID));
    return FIPC_OK;
}

// chunk offsets of nq query boxes in qoff (n+1), returns the chunk count
static int query_prep(fipc_ctx* c, int nq, const BoxFn& qbox, DVec<int>& qoff, int* nchunks)
{
    STCHK(ensure(qoff, (size_t)nq + 2));
    HIPCHK(hipMemsetAsync(c->flag.p, 0, 4, c->s));
    k_query_chunks<<<nblk(nq + 1), TPB, 0, c->s>>>(nq, qbox, c->inv_dx, qoff.p, c->flag.p);
    HIPCHK(hipGetLastError());
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, qoff.p, qoff.p, nq + 1, c->s));
    if (bytes > c->cub_tmp.cap) STCHK(ensure(c->cub_tmp, bytes));
    bytes = c->cub_tmp.cap;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, bytes, qoff.p, qoff.p, nq + 1, c->s));
    int hv[2];
    HIPCHK(hipMemcpyAsync(&hv[0], qoff.p + nq, 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(&hv[1], c->flag.p, 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (hv[1]) return FIPC_ERR_CAPACITY;
    *nchunks = hv[0];
    return FIPC_OK;
}

template <class Visit>
static int run_query(fipc_ctx* c, int nq, const BoxFn& qbox, Visit visit, int timer)
{
    if (nq == 0 || c->grid.nentries == 0) return FIPC_OK;
    int nchunks = 0;
    STCHK(query_prep(c, nq, qbox, c->grid.qoff, &nchunks));
    if (!nchunks) return FIPC_OK;
    Grid& g = c->grid;
    STCHK(fipc_tm_begin(c, timer));
    k_grid_query<<<nblk(nchunks), TPB, 0, c->s>>>(nq, nchunks, c->grid.qoff.p, qbox, c->inv_dx, g.M, g.bstart.p, g.items.p,
                                                   c->keys_tmp.p + g.nentries, visit);
    HIPCHK(hipGetLastError());
    STCHK(fipc_tm_end(c, timer));
    return FIPC_OK;
}

// ---- constraint search ----------------------------------------------------
struct RowSink {
    int* cnt;                       // 7 raw counters
    long long cap[NTYPES];
    unsigned long long* kpp;        // PP keys (a << 21 | b)
    unsigned long long* kpe;        // PE keys (p << 42 | e0 << 21 | e1)
    int* rows[NTYPES];
    __device__ __forceinline__ void pp(int a, int b) const
    {
        int s = atomicAdd(&cnt[0], 1);
        if (s < cap[0]) kpp[s] = ((unsigned long long)a << 21) | (unsigned long long)b;
    }
    __device__ __forceinline__ void pe(int p, int e0, int e1) const
    {
        int s = atomicAdd(&cnt[1], 1);
        if (s < cap[1]) kpe[s] = ((unsigned long long)p << 42) | ((unsigned long long)e0 << 21) | (unsigned long long)e1;
    }
    __device__ __forceinline__ void r4(int t, int a, int b, int c, int d) const
    {
        int s = atomicAdd(&cnt[t], 1);
        if (s < cap[t]) { int* r = rows[t] + 4 * s; r[0] = a; r[1] = b; r[2] = c; r[3] = d; }
    }
};

// attempt_PT (NEWTON_IPC.py:181-212); query = triangle, item = boundary point
struct VisitPT {
    const double* x;
    const int *bp, *bt;
    double dHat, dHat2;
    RowSink out;
    __device__ __forceinline__ void operator()(int tri, int pi) const
    {
        int p = bp[pi], t0 = bt[3 * tri], t1 = bt[3 * tri + 1], t2 = bt[3 * tri + 2];
        if (p == t0 || p == t1 || p == t2) return;
        v3 P = ldv(x, p), T0 = ldv(x, t0), T1 = ldv(x, t1), T2 = ldv(x, t2);
        // point_triangle_ccd_broadphase (ipc.py:332-336)
        v3 mx = vmax(vmax(T0, T1), T2), mn = vmin(vmin(T0, T1), T2);
        if (!(P.x < mx.x + dHat && P.y < mx.y + dHat && P.z < mx.z + dHat)) return;
        if (!(mn.x - dHat < P.x && mn.y - dHat < P.y && mn.z - dHat < P.z)) return;
        int k = PT_type(P, T0, T1, T2);
        if (!(PT_dist2(P, T0, T1, T2, k) < dHat2)) return;
        switch (k) {
        case 0: out.pp(p, t0); break;
        case 1: out.pp(p, t1); break;
        case 2: out.pp(p, t2); break;
        case 3: out.pe(p, t0, t1); break;
        case 4: out.pe(p, t1, t2); break;
        case 5: out.pe(p, t2, t0); break;
        default: out.r4(2, p, t0, t1, t2); break;
        }
    }
};
// attempt_EE (NEWTON_IPC.py:235-312); query = edge i, item = edge j, keep i < j
struct VisitEE {
    const double *x, *x0;
    const int* be;
    double dHat, dHat2;
    RowSink out;
    __device__ __forceinline__ void operator()(int i, int j) const
    {
        if (!(i < j)) return;
        int a0 = be[2 * i], a1 = be[2 * i + 1], b0 = be[2 * j], b1 = be[2 * j + 1];
        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
        v3 A0 = ldv(x, a0), A1 = ldv(x, a1), B0 = ldv(x, b0), B1 = ldv(x, b1);
        // edge_edge_ccd_broadphase (ipc.py:348-354)
        v3 mxa = vmax(A0, A1), mna = vmin(A0, A1), mxb = vmax(B0, B1), mnb = vmin(B0, B1);
        if (!(mna.x < mxb.x + dHat && mna.y < mxb.y + dHat && mna.z < mxb.z + dHat)) return;
        if (!(mnb.x - dHat < mxa.x && mnb.y - dHat < mxa.y && mnb.z - dHat < mxa.z)) return;
        int k = EE_type(A0, A1, B0, B1);
        if (!(EE_dist2(A0, A1, B0, B1, k) < dHat2)) return;
        double cn2 = EECN2(A0, A1, B0, B1);
        double eps = M_threshold(ldv(x0, a0), ldv(x0, a1), ldv(x0, b0), ldv(x0, b1));
        bool moll = cn2 < eps;
        switch (k) {
        case 0: if (moll) out.r4(5, a0, a1, b0, b1); else out.pp(a0, b0); break;
        case 1: if (moll) out.r4(5, a0, a1, b1, b0); else out.pp(a0, b1); break;
        case 3: if (moll) out.r4(5, a1, a0, b0, b1); else out.pp(a1, b0); break;
        case 4: if (moll) out.r4(5, a1, a0, b1, b0); else out.pp(a1, b1); break;
        case 2: if (moll) out.r4(6, a0, a1, b0, b1); else out.pe(a0, b0, b1); break;
        case 5: if (moll) out.r4(6, a1, a0, b0, b1); else out.pe(a1, b0, b1); break;
        case 6: if (moll) out.r4(6, b0, b1, a0, a1); else out.pe(b0, a0, a1); break;
        case 7: if (moll) out.r4(6, b1, b0, a0, a1); else out.pe(b1, a0, a1); break;
        default: out.r4(moll ? 4 : 3, a0, a1, b0, b1); break;
        }
    }
};

__global__ void k_decode_pp(int n, const unsigned long long* k, int* rows)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long v = k[i];
    rows[2 * i] = (int)(v >> 21);
    rows[2 * i + 1] = (int)(v & ((1ull << 21) - 1));
}
__global__ void k_decode_pe(int n, const unsigned long long* k, int* rows)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long v = k[i], m = (1ull << 21) - 1;
    rows[3 * i] = (int)(v >> 42);
    rows[3 * i + 1] = (int)((v >> 21) & m);
    rows[3 * i + 2] = (int)(v & m);
}


static int dedup(fipc_ctx* c, unsigned long long* keys, int n, int end_bit, int* nuniq)
{
    if (n == 0) { *nuniq = 0; return FIPC_OK; }
    unsigned long long* sorted = c->keys_tmp.p;
    STCHK(ensure(c->keys_tmp, (size_t)n * 2 + 2));
    sorted = c->keys_tmp.p;
    unsigned long long* uniq = sorted + n;
    size_t b1 = 0, b2 = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, b1, keys, sorted, n, 0, end_bit, c->s));
    HIPCHK(hipcub::DeviceSelect::Unique(nullptr, b2, sorted, uniq, c->nuniq.p, n, c->s));
    size_t need = std::max(b1, b2);
    if (need > c->cub_tmp.cap) STCHK(ensure(c->cub_tmp, need));
    b1 = c->cub_tmp.cap;
    HIPCHK(hipcub::DeviceRadixSort::SortKeys(c->cub_tmp.p, b1, keys, sorted, n, 0, end_bit, c->s));
    b2 = c->cub_tmp.cap;
    HIPCHK(hipcub::DeviceSelect::Unique(c->cub_tmp.p, b2, sorted, uniq, c->nuniq.p, n, c->s));
    HIPCHK(hipMemcpyAsync(nuniq, c->nuniq.p, 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpyAsync(keys, uniq, (size_t)*nuniq * 8, hipMemcpyDeviceToDevice, c->s));
    return FIPC_OK;
}

static int bits_for(int n)
{
    int b = 1;
    while ((1 << b) < n) b++;
    return b;
}

static int find_once(fipc_ctx* c, int* counts)
{
    HIPCHK(hipMemsetAsync(c->counters.p, 0, NTYPES * 4, c->s));
    RowSink sink;
    sink.cnt = c->counters.p;
    for (int t = 0; t < NTYPES; t++) { sink.cap[t] = c->cap[t]; sink.rows[t] = c->rows[t].p; }
    sink.kpp = c->keys_pp.p;
    sink.kpe = c->keys_pe.p;
    // PT: points inserted with +-dHat, triangles query their box
    BoxFn pts{c->x.p, nullptr, c->bp.p, 0, c->dHat};
    // point boxes index bp directly: kind 0 reads idx[i]
    STCHK(grid_build(c, c->nbp, pts));
    BoxFn tris{c->x.p, nullptr, c->bt.p, 2, 0.0};
    VisitPT vpt{c->x.p, c->bp.p, c->bt.p, c->dHat, c->dHat2, sink};
    STCHK(run_query(c, c->nbt, tris, vpt, TM_FIND));
    // EE: edges inserted with their box, queried with box +- dHat
    BoxFn edges{c->x.p, nullptr, c->be.p, 1, 0.0};
    STCHK(grid_build(c, c->nbe, edges));
    BoxFn qedges{c->x.p, nullptr, c->be.p, 1, c->dHat};
    VisitEE vee{c->x.p, c->x0.p, c->be.p, c->dHat, c->dHat2, sink};
    STCHK(run_query(c, c->nbe, qedges, vee, TM_FIND));
    HIPCHK(hipMemcpyAsync(counts, c->counters.p, NTYPES * 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

extern "C" int fipc_find_constraints(fipc_ctx* c, int* counts_out)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    int raw[NTYPES];
    for (int attempt = 0; attempt < 4; attempt++) {
        STCHK(ensure(c->keys_pp, c->cap[0]));
        STCHK(ensure(c->keys_pe, c->cap[1]));
        for (int t = 2; t < NTYPES; t++) STCHK(ensure(c->rows[t], c->cap[t] * 4));
        STCHK(find_once(c, raw));
        bool grow = false;
        for (int t = 0; t < NTYPES; t++)
            if (raw[t] > c->cap[t]) { c->cap[t] = (int64_t)raw[t] * 2; grow = true; }
        if (!grow) break;
        if (attempt == 3) return FIPC_ERR_CAPACITY;
    }
    for (int t = 0; t < NTYPES; t++) c->raw_count[t] = raw[t];
    int bits = bits_for(c->nv);
    int npp = 0, npe = 0;
    STCHK(dedup(c, c->keys_pp.p, raw[0], std::min(64, 21 + bits), &npp));
    STCHK(dedup(c, c->keys_pe.p, raw[1], std::min(64, 42 + bits), &npe));
    STCHK(ensure(c->rows[0], (size_t)std::max(npp, 1) * 2));
    STCHK(ensure(c->rows[1], (size_t)std::max(npe, 1) * 3));
    if (npp) k_decode_pp<<<nblk(npp), TPB, 0, c->s>>>(npp, c->keys_pp.p, c->rows[0].p);
    if (npe) k_decode_pe<<<nblk(npe), TPB, 0, c->s>>>(npe, c->keys_pe.p, c->rows[1].p);
    HIPCHK(hipGetLastError());
    c->count[0] = npp;
    c->count[1] = npe;
    for (int t = 2; t < NTYPES; t++) c->count[t] = raw[t];
    for (int t = 0; t < NTYPES; t++) {
        int N = 3 * (t == 0 ? 2 : (t == 1 ? 3 : 4));
        STCHK(ensure(c->cH[t], (size_t)std::max(c->count[t], 1) * N * N));
        if (counts_out) counts_out[t] = c->count[t];
    }
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

extern "C" int fipc_get_constraints(fipc_ctx* c, int type, int* rows_out)
{
    if (!c || type < 0 || type >= NTYPES || !rows_out) return FIPC_ERR_ARG;
    int w = type == 0 ? 2 : (type == 1 ? 3 : 4);
    if (c->count[type])
        HIPCHK(hipMemcpyAsync(rows_out, c->rows[type].p, (size_t)c->count[type] * w * 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

// ---- CCD filters ----------------------------------------------------------
__device__ __forceinline__ void amin_d(unsigned long long* a, double v)
{
    atomicMin(a, (unsigned long long)__double_as_longlong(v));
}
// compute_filter_3D_PT (NEWTON_IPC.py:377-406); query = triangle, item = point
struct VisitCCDPT {
    const double *x, *p;
    const int *bp, *bt;
    double dHat, eta;
    unsigned long long* amin;
    __device__ __forceinline__ void operator()(int tri, int pi) const
    {
        int q = bp[pi], t0 = bt[3 * tri], t1 = bt[3 * tri + 1], t2 = bt[3 * tri + 2];
        if (q == t0 || q == t1 || q == t2) return;
        v3 P = ldv(x, q), T0 = ldv(x, t0), T1 = ldv(x, t1), T2 = ldv(x, t2);
        v3 dP = ldv(p, q), d0 = ldv(p, t0), d1 = ldv(p, t1), d2 = ldv(p, t2);
        // moving_point_triangle_ccd_broadphase (ipc.py:339-345)
        v3 mxp = vmax(P, P + dP), mnp = vmin(P, P + dP);
        v3 mxt = vmax(vmax(T0, T0 + d0), vmax(vmax(T1, T1 + d1), vmax(T2, T2 + d2)));
        v3 mnt = vmin(vmin(T0, T0 + d0), vmin(vmin(T1, T1 + d1), vmin(T2, T2 + d2)));
        if (!(mnp.x < mxt.x + dHat && mnp.y < mxt.y + dHat && mnp.z < mxt.z + dHat)) return;
        if (!(mnt.x - dHat < mxp.x && mnt.y - dHat < mxp.y && mnt.z - dHat < mxp.z)) return;
        double d2c = PT_dist2(P, T0, T1, T2, PT_type(P, T0, T1, T2));
        double toc = pt_ccd(P, T0, T1, T2, dP, d0, d1, d2, eta, d2c);
        if (toc < 1.0) amin_d(amin, toc);
    }
};
// compute_filter_3D_EE (NEWTON_IPC.py:409-440); query edge i, item edge j
struct VisitCCDEE {
    const double *x, *p;
    const int* be;
    double dHat, eta;
    unsigned long long* amin;
    __device__ __forceinline__ void operator()(int i, int j) const
    {
        if (!(i < j)) return;
        int a0 = be[2 * i], a1 = be[2 * i + 1], b0 = be[2 * j], b1 = be[2 * j + 1];
        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
        v3 A0 = ldv(x, a0), A1 = ldv(x, a1), B0 = ldv(x, b0), B1 = ldv(x, b1);
        v3 dA0 = ldv(p, a0), dA1 = ldv(p, a1), dB0 = ldv(p, b0), dB1 = ldv(p, b1);
        // moving_edge_edge_ccd_broadphase (ipc.py:357-363)
        v3 mxa = vmax(vmax(A0, A0 + dA0), vmax(A1, A1 + dA1)), mna = vmin(vmin(A0, A0 + dA0), vmin(A1, A1 + dA1));
        v3 mxb = vmax(vmax(B0, B0 + dB0), vmax(B1, B1 + dB1)), mnb = vmin(vmin(B0, B0 + dB0), vmin(B1, B1 + dB1));
        if (!(mna.x < mxb.x + dHat && mna.y < mxb.y + dHat && mna.z < mxb.z + dHat)) return;
        if (!(mnb.x - dHat < mxa.x && mnb.y - dHat < mxa.y && mnb.z - dHat < mxa.z)) return;
        double d2c = EE_dist2(A0, A1, B0, B1, EE_type(A0, A1, B0, B1));
        double toc = ee_ccd(A0, A1, B0, B1, dA0, dA1, dB0, dB1, eta, d2c);
        if (toc < 1.0) amin_d(amin, toc);
    }
};

extern int fipc_inversion_alpha(fipc_ctx* c, unsigned long long* dst);

extern "C" int fipc_ccd_alpha(fipc_ctx* c, double eta, int inversion, double* alpha, double* parts)
{
    if (!c || !c->nv || !alpha) return FIPC_ERR_STATE;
    double one = 1.0;
    unsigned long long init[3];
    for (int i = 0; i < 3; i++) memcpy(&init[i], &one, 8);
    HIPCHK(hipMemcpyAsync(c->amin.p, init, 24, hipMemcpyHostToDevice, c->s));
    BoxFn pts{c->x.p, c->sol.p, c->bp.p, 0, c->dHat};
    STCHK(grid_build(c, c->nbp, pts));
    BoxFn tris{c->x.p, c->sol.p, c->bt.p, 2, 0.0};
    VisitCCDPT vpt{c->x.p, c->sol.p, c->bp.p, c->bt.p, c->dHat, eta, c->amin.p};
    STCHK(run_query(c, c->nbt, tris, vpt, TM_CCD));
    BoxFn edges{c->x.p, c->sol.p, c->be.p, 1, 0.0};
    STCHK(grid_build(c, c->nbe, edges));
    BoxFn qedges{c->x.p, c->sol.p, c->be.p, 1, c->dHat};
    VisitCCDEE vee{c->x.p, c->sol.p, c->be.p, c->dHat, eta, c->amin.p + 1};
    STCHK(run_query(c, c->nbe, qedges, vee, TM_CCD));
    if (inversion) STCHK(fipc_inversion_alpha(c, c->amin.p + 2));
    unsigned long long out[3];
    HIPCHK(hipMemcpyAsync(out, c->amin.p, 24, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    double a[3];
    memcpy(a, out, 24);
    if (parts) { parts[0] = a[0]; parts[1] = a[1]; parts[2] = a[2]; }
    *alpha = std::min(std::min(a[0], a[1]), a[2]);
    return FIPC_OK;
}

// ---- collision check --------------------------------------------------------
struct VisitCollide {
    const double* x;
    const int *be, *bt;
    int* hit;
    __device__ __forceinline__ void operator()(int tri, int ei) const
    {
        int a0 = be[2 * ei], a1 = be[2 * ei + 1], t0 = bt[3 * tri], t1 = bt[3 * tri + 1], t2 = bt[3 * tri + 2];
        if (a0 == t0 || a0 == t1 || a0 == t2 || a1 == t0 || a1 == t1 || a1 == t2) return;
        if (seg_tri(ldv(x, a0), ldv(x, a1), ldv(x, t0), ldv(x, t1), ldv(x, t2))) atomicOr(hit, 1);
    }
};

extern "C" int fipc_check_collision(fipc_ctx* c, int* hit)
{
    if (!c || !c->nv || !hit) return FIPC_ERR_STATE;
    BoxFn edges{c->x.p, nullptr, c->be.p, 1, c->dHat};
    STCHK(grid_build(c, c->nbe, edges));
    HIPCHK(hipMemsetAsync(c->flag.p + 1, 0, 4, c->s));
    BoxFn tris{c->x.p, nullptr, c->bt.p, 2, 0.0};
    VisitCollide v{c->x.p, c->be.p, c->bt.p, c->flag.p + 1};
    STCHK(run_query(c, c->nbt, tris, v, TM_FIND));
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, c->flag.p + 1, 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    *hit = h;
    return FIPC_OK;
}

// ---- batched CCD for tests -------------------------------------------------
__global__ void k_batch_ccd(int type, int n, const double* pts, const double* disp, double eta, const double* d2, double* toc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* P = pts + 12 * i;
    const double* D = disp + 12 * i;
    v3 a = ld3(P), b = ld3(P + 3), cc = ld3(P + 6), d = ld3(P + 9);
    v3 da = ld3(D), db = ld3(D + 3), dc = ld3(D + 6), dd = ld3(D + 9);
    toc[i] = type == 0 ? pt_ccd(a, b, cc, d, da, db, dc, dd, eta, d2[i]) : ee_ccd(a, b, cc, d, da, db, dc, dd, eta, d2[i]);
}
extern "C" int fipc_batch_ccd(int type, int n, const double* pts, const double* disp, double eta, const double* dist2, double* toc)
{
    if (n < 0 || (type != 0 && type != 1)) return FIPC_ERR_ARG;
    if (n == 0) return FIPC_OK;
    double *dp, *dd, *d2, *dt;
    HIPCHK(hipMalloc(&dp, (size_t)n * 96)); HIPCHK(hipMalloc(&dd, (size_t)n * 96));
    HIPCHK(hipMalloc(&d2, (size_t)n * 8)); HIPCHK(hipMalloc(&dt, (size_t)n * 8));
    HIPCHK(hipMemcpy(dp, pts, (size_t)n * 96, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dd, disp, (size_t)n * 96, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d2, dist2, (size_t)n * 8, hipMemcpyHostToDevice));
    k_batch_ccd<<<nblk(n), TPB>>>(type, n, dp, dd, eta, d2, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(toc, dt, (size_t)n * 8, hipMemcpyDeviceToHost));
    (void)hipFree(dp); (void)hipFree(dd); (void)hipFree(d2); (void)hipFree(dt);
    return FIPC_OK;
}
