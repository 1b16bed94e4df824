// fipc_broad.hip -- spatial-hash broad phase and everything built on it:
// constraint search (NEWTON_IPC.py:159-357), CCD step filters (:377-468) and
// the collision check (:855-886).
//
// Grid: every inserted item is entered in each cell its box overlaps
// (cell size = 1 / inv_dx, as the reference's hash).  Entries are grouped by
// a hashed bucket; each entry keeps its exact 63-bit cell key and three bits
// telling whether the cell is the item's lowest cell along x, y, z.  A query
// runs one thread per (query box, cell) pair, so huge boxes do not serialise
// a wave, and accepts an entry only in the first cell common to both boxes,
// so every pair is tested exactly once.  Builds and queries size their
// buffers from capacities and never wait for the host; each operation
// synchronises once at its end and is redone when a capacity was exceeded.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <algorithm>
#include <type_traits>
#include <vector>
#include "fipc_ctx.h"
#include "fipc_rows.h"

using namespace fipc;

#define TPB 256
static inline int nblk(long n, int t = TPB) { return (int)((n + t - 1) / t); }


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

#define GRID_MAXB 2048
static inline int stride_grid(long n) { return std::max(1, std::min(nblk(n), GRID_MAXB)); }

// ---- two-level offsets ----------------------------------------------------
// The entries of n boxes (one per overlapped cell) are numbered without a
// device-wide scan: k_count_cells stores each box's offset within its block
// of TPB boxes (loc) and the block's total (part); a consumer workgroup
// prefix-sums the ≤ OFF_PMAX block totals in LDS, and the offset of box i is
// P[i / TPB] + loc[i].  Saves a scan launch pair per grid build and query.
constexpr int OFF_PMAX = 8192;  // block totals per table: up to 2M boxes (32 KB of LDS)

// exclusive prefix of v over the workgroup (TPB threads), *total = the sum
__device__ __forceinline__ long long block_exscan(long long v, long long* total)
{
    __shared__ long long wsum[TPB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        long long y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    long long before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < TPB / 64; k++) {
        if (k < w) before += wsum[k];
        all += wsum[k];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}
// cells of every box: loc[i] = offset of box i within its block, part[b] =
// the block's total; a box over more than 2^24 cells, or a block over 2^30,
// flags GF_HUGE (the operation fails with FIPC_ERR_CAPACITY)
__global__ void __launch_bounds__(TPB) k_count_cells(int n, BoxFn box, double inv, int* loc, int* part, int* flag,
                                                     int* zero = nullptr, long nzero = 0)
{
    const int i = blockIdx.x * TPB + threadIdx.x;
    // a grid build clears its bucket counts here (the fill adds to them)
    for (long k = i; k < nzero; k += (long)gridDim.x * TPB) zero[k] = 0;
    long long c = 0;
    if (i < n) {
        v3 lo, hi;
        box(i, lo, hi);
        c = ncells(cells_of(lo, hi, inv));
        if (c > (1 << 24) || c < 0) { atomicOr(flag, GF_HUGE); c = 0; }
    }
    long long tot;
    const long long pre = block_exscan(c, &tot);
    if (tot >= (1LL << 30)) {
        if (threadIdx.x == 0) atomicOr(flag, GF_HUGE);
        tot = 0;
    }
    if (i < n) loc[i] = (int)pre;
    if (threadIdx.x == 0) part[blockIdx.x] = (int)tot;
}
// sP[0..np] = exclusive prefix of part[0..np) in LDS (every thread of the
// workgroup calls it); returns the total, or -1 past 2^31 - 1
__device__ int lds_prefix(const int* __restrict__ part, int np, int* sP)
{
    for (int k = threadIdx.x; k < np; k += TPB) sP[k] = part[k];
    __syncthreads();
    const int per = (np + TPB - 1) / TPB, k0 = min(np, (int)threadIdx.x * per), k1 = min(np, k0 + per);
    long long sum = 0;
    for (int k = k0; k < k1; k++) sum += sP[k];
    long long tot;
    long long pre = block_exscan(sum, &tot);
    if (tot > 0x7fffffffLL) tot = -1, pre = 0;
    for (int k = k0; k < k1; k++) {
        const int v = sP[k];
        sP[k] = (int)pre;
        pre += v;
    }
    __syncthreads();
    return (int)tot;
}
// the box whose entries hold entry e (0 <= e < total): the last box with
// offset <= e (boxes with no cells share the next box's offset); *first = its
// first entry
__device__ __forceinline__ int find_box(int e, const int* sP, int np, const int* __restrict__ loc, int n, int* first)
{
    int lo = 0, hi = np;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (sP[mid] <= e) lo = mid; else hi = mid;
    }
    const int r = e - sP[lo];
    int l = lo * TPB, h = min(n, l + TPB);
    while (h - l > 1) {
        const int mid = (l + h) >> 1;
        if (loc[mid] <= r) l = mid; else h = mid;
    }
    *first = sP[lo] + loc[l];
    return l;
}

// Grid-stride over the entries (their number is the sum of the block totals,
// so the host never waits for it).  A count above the capacity flags a retry
// and records the count needed; workgroup 0 stores the count to part[np]
// for the scatter.  The owning item of an entry is found by binary search
// in the offsets, so an item covering thousands of cells (a long plate edge)
// does not serialise a wave.  Dynamic LDS: (np + 1) ints.
__global__ void __launch_bounds__(TPB) k_grid_fill(int n, int np, int cap, BoxFn box, double inv, unsigned M,
                                                   const int* __restrict__ loc, int* part, unsigned* ekey,
                                                   int* eitem, unsigned long long* ecell, int* bcount, int* flag)
{
    extern __shared__ int sP[];
    const int ne = lds_prefix(part, np, sP);
    if (blockIdx.x == 0 && threadIdx.x == 0) part[np] = ne;
    if (ne > cap || ne < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(flag, ne < 0 ? GF_HUGE : GF_RETRY);
            atomicMax(flag + 2, ne);
        }
        return;
    }
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
        int first;
        const int i = find_box(e, sP, np, loc, n, &first);
        v3 blo, bhi;
        box(i, blo, bhi);
        CellRange r = cells_of(blo, bhi, inv);
        int ny = r.hi[1] - r.lo[1] + 1, nz = r.hi[2] - r.lo[2] + 1;
        int l = e - first;
        int a = r.lo[0] + l / (ny * nz), b = r.lo[1] + (l / nz) % ny, c = r.lo[2] + l % nz;
        unsigned flags = (a == r.lo[0] ? 1u : 0u) | (b == r.lo[1] ? 2u : 0u) | (c == r.lo[2] ? 4u : 0u);
        unsigned k = bucket_of(a, b, c, M);
        ekey[e] = k;
        eitem[e] = (int)((unsigned)i | (flags << 28));
        ecell[e] = cell_key(a, b, c);
        atomicAdd(&bcount[k], 1);
    }
}
// bstart holds the scanned bucket starts; each entry takes the next slot of
// its bucket, which leaves bstart[k] at the end of bucket k (the start of
// k + 1): bucket k is [k ? bstart[k - 1] : 0, bstart[k]) afterwards
__global__ void k_grid_scatter(const int* __restrict__ nep, int cap, const unsigned* ekey, const int* eitem,
                               const unsigned long long* ecell, int* bstart, int* items, unsigned long long* cells)
{
    int ne = *nep;
    if (ne > cap || ne < 0) return;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
        unsigned k = ekey[e];
        int pos = atomicAdd(&bstart[k], 1);
        items[pos] = eitem[e];
        cells[pos] = ecell[e];
    }
}

// A visitor with a prep(q) member loads what it needs of the query once per
// (query, cell) pair and takes it with each item: visit.item(st, q, item)
template <class V, class = void>
struct has_prep : std::false_type {};
template <class V>
struct has_prep<V, std::void_t<decltype(&V::prep)>> : std::true_type {};

// Generic query, one thread per (query, cell) pair, grid-stride: visit(q,
// item) is called once per (query, item) pair whose cell ranges overlap,
// in the first cell common to both.  Does nothing when the grid build
// flagged a retry (the caller then redoes the whole operation).
template <class Visit>
__global__ void __launch_bounds__(TPB) k_grid_query(int nq, int np, const int* __restrict__ qloc,
                                                    const int* __restrict__ qpart, BoxFn qbox, double inv,
                                                    unsigned M, const int* __restrict__ bstart,
                                                    const int* __restrict__ items,
                                                    const unsigned long long* __restrict__ cells, int* flag,
                                                    Visit visit)
{
    extern __shared__ int sP[];
    // the whole workgroup leaves or stays (lds_prefix synchronises it)
    __shared__ int stop;
    if (threadIdx.x == 0) stop = flag[0] & (GF_RETRY | GF_HUGE);
    __syncthreads();
    if (stop) return;
    const int total = lds_prefix(qpart, np, sP);
    if (total < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flag, GF_HUGE);
        return;
    }
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        int first;
        const int q = find_box(t, sP, np, qloc, nq, &first);
        const int l = t - first;
        v3 blo, bhi;
        qbox(q, blo, bhi);
        CellRange r = cells_of(blo, bhi, inv);
        int ny = r.hi[1] - r.lo[1] + 1, nz = r.hi[2] - r.lo[2] + 1;
        int a = r.lo[0] + l / (ny * nz);
        int b = r.lo[1] + (l / nz) % ny;
        int c = r.lo[2] + l % nz;
        unsigned long long key = cell_key(a, b, c);
        unsigned bk = bucket_of(a, b, c, M);
        int e0 = bk ? bstart[bk - 1] : 0, e1 = bstart[bk];
        // first common cell: along each axis the cell is the lowest cell of the query or of the item
        auto first_common = [&](unsigned fl) {
            return (a == r.lo[0] || (fl & 1u)) && (b == r.lo[1] || (fl & 2u)) && (c == r.lo[2] || (fl & 4u));
        };
        if constexpr (has_prep<Visit>::value) {
            const auto st = visit.prep(q);
            for (int e = e0; e < e1; e++) {
                if (cells[e] != key) continue;
                unsigned it = (unsigned)items[e];
                if (!first_common(it >> 28)) continue;
                visit.item(st, q, (int)(it & 0x0fffffffu));
            }
        } else {
            for (int e = e0; e < e1; e++) {
                if (cells[e] != key) continue;
                unsigned it = (unsigned)items[e];
                if (!first_common(it >> 28)) continue;
                visit(q, (int)(it & 0x0fffffffu));
            }
        }
    }
}

// st: the stream (c->s if null); each stream has its own scan scratch
static int scan_inplace(fipc_ctx* c, int* d, int n, hipStream_t st = nullptr)
{
    const hipStream_t s = st ? st : c->s;
    DVec<unsigned char>& tmp = s == c->s ? c->cub_tmp : c->cub_tmp2;
    size_t bytes = 0;
    HIPCHK(rocprim::exclusive_scan(nullptr, bytes, d, d, 0, (size_t)n, rocprim::plus<int>(), s));
    if (bytes > tmp.cap) STCHK(ensure(tmp, bytes));
    bytes = tmp.cap;
    HIPCHK(rocprim::exclusive_scan(tmp.p, bytes, d, d, 0, (size_t)n, rocprim::plus<int>(), s));
    return FIPC_OK;
}

// Build the grid over n items without waiting for the host: buffers are
// sized by the capacity, overflow is flagged in flag[0] (see GF_RETRY).
static int grid_build(fipc_ctx* c, int slot, int n, const BoxFn& box, hipStream_t st = nullptr)
{
    const hipStream_t s = st ? st : c->s;
    Grid& g = c->grid[slot];
    // the same boxes at the same x: the grid is already there; or the grid
    // sweeps these boxes from the x of version g.x_version along the current
    // sol, and x lies on that segment
    if (g.reusable && !box.sol && g.kind == box.kind && g.n == n && g.infl == box.infl) {
        if (!g.swept && g.x_version == c->x_version) return FIPC_OK;
        if (g.swept && c->seg_x == c->x_version && c->seg_base == g.x_version && c->seg_sol == g.seg_sol &&
            c->seg_sol == c->sol_version)
            return FIPC_OK;
    }
    g.reusable = false;
    const int np = nblk(n);
    if (np > OFF_PMAX) return FIPC_ERR_CAPACITY;
    STCHK(fipc_tm_begin(c, TM_GRID, s));
    int cap = g.cap;
    // buckets: a power of two at least the entry capacity
    unsigned M = 1024;
    while (M < (unsigned)cap && M < (1u << 30)) M <<= 1;
    g.M = M;
    STCHK(ensure(g.cnt, (size_t)n + 2));
    STCHK(ensure(g.part, (size_t)np + 2));
    STCHK(ensure(g.ekey, cap)); STCHK(ensure(g.eitem, cap)); STCHK(ensure(g.items, cap));
    STCHK(ensure(g.ecell, cap)); STCHK(ensure(g.cells, cap));
    STCHK(ensure(g.bstart, (size_t)M + 1));
    if (np)
        k_count_cells<<<np, TPB, 0, s>>>(n, box, c->inv_dx, g.cnt.p, g.part.p, c->flag.p, g.bstart.p, (long)M + 1);
    else
        HIPCHK(hipMemsetAsync(g.bstart.p, 0, ((size_t)M + 1) * 4, s));
    HIPCHK(hipGetLastError());
    const int* nep = g.part.p + np;  // the entry count, stored by k_grid_fill
    int nb = stride_grid(cap);
    k_grid_fill<<<nb, TPB, (np + 1) * sizeof(int), s>>>(n, np, cap, box, c->inv_dx, M, g.cnt.p, g.part.p, g.ekey.p,
                                                         g.eitem.p, g.ecell.p, g.bstart.p, c->flag.p);
    HIPCHK(hipGetLastError());
    STCHK(scan_inplace(c, g.bstart.p, M + 1, s));
    k_grid_scatter<<<nb, TPB, 0, s>>>(nep, cap, g.ekey.p, g.eitem.p, g.ecell.p, g.bstart.p, g.items.p, g.cells.p);
    HIPCHK(hipGetLastError());
    // reusable once built; an overflow found after the operation's
    // synchronisation clears this again (bp_check).  A swept grid is reusable
    // only when it sweeps along sol itself
    g.swept = box.sol != nullptr;
    g.reusable = !g.swept || box.sol == c->sol.p;
    g.seg_sol = c->sol_version;
    g.x_version = c->x_version;
    g.kind = box.kind;
    g.n = n;
    g.infl = box.infl;
    STCHK(fipc_tm_end(c, TM_GRID, s));
    return FIPC_OK;
}

// query the last grid with nq boxes; no host synchronisation
template <class Visit>
static int run_query(fipc_ctx* c, int slot, int nq, const BoxFn& qbox, Visit visit, int timer,
                     hipStream_t st = nullptr)
{
    if (nq == 0) return FIPC_OK;
    const hipStream_t s = st ? st : c->s;
    Grid& g = c->grid[slot];
    const int np = nblk(nq);
    if (np > OFF_PMAX) return FIPC_ERR_CAPACITY;
    STCHK(ensure(g.qoff, (size_t)nq + 2));
    STCHK(ensure(g.qpart, (size_t)np + 2));
    STCHK(fipc_tm_begin(c, timer, s));
    k_count_cells<<<np, TPB, 0, s>>>(nq, qbox, c->inv_dx, g.qoff.p, g.qpart.p, c->flag.p);
    HIPCHK(hipGetLastError());
    k_grid_query<<<GRID_MAXB, TPB, (np + 1) * sizeof(int), s>>>(nq, np, g.qoff.p, g.qpart.p, qbox, c->inv_dx, g.M,
                                                                 g.bstart.p, g.items.p, g.cells.p,
                                             c->flag.p, visit);
    HIPCHK(hipGetLastError());
    STCHK(fipc_tm_end(c, timer, s));
    return FIPC_OK;
}

// start of a broad-phase operation, one launch: clears the status word and
// needed count (flag[0..2]) and ncnt counters, and sets namin minima to 1.0
__global__ void k_bp_begin(int* flag, int* cnt, int ncnt, unsigned long long* amin, int namin)
{
    const int i = threadIdx.x;
    if (i < 3) flag[i] = 0;
    if (i < ncnt) cnt[i] = 0;
    if (i < namin) amin[i] = 0x3ff0000000000000ull;  // the bits of 1.0
}
static int bp_begin(fipc_ctx* c, int* cnt = nullptr, int ncnt = 0, unsigned long long* amin = nullptr,
                    int namin = 0)
{
    k_bp_begin<<<1, 64, 0, c->s>>>(c->flag.p, cnt, ncnt, amin, namin);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}
// after the operation's synchronisation: error, retry (capacity grown) or done
static int bp_check(fipc_ctx* c, const int* hflag, bool* retry)
{
    *retry = false;
    if (hflag[0] & (GF_HUGE | GF_RETRY)) c->grid[0].reusable = c->grid[1].reusable = false;
    if (hflag[0] & GF_HUGE) return FIPC_ERR_CAPACITY;
    if (hflag[0] & GF_RETRY) {
        long need = (long)hflag[2];
        long cap = need + need / 4 + 4096;
        if (cap > (1L << 30)) return FIPC_ERR_CAPACITY;
        for (Grid& g : c->grid) g.cap = (int)std::max<long>(g.cap, cap);
        *retry = true;
    }
    return FIPC_OK;
}

// ---- constraint search ----------------------------------------------------
// one atomic per wavefront: adds the number of active lanes to *cnt
__device__ __forceinline__ void wave_count(int* cnt)
{
    unsigned long long m = __ballot(1);
    if (__lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(cnt, (int)__popcll(m));
}
struct RowSink {
    int* cnt;                       // 7 raw counters
    int* cand;                      // broad-phase candidates (PT, EE), or null when not counted
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
        if (out.cand) wave_count(out.cand);
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
        if (out.cand) wave_count(out.cand + 1);
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

// decode the unique sorted keys (count read on the device) into rows
__global__ void k_decode_pp(const int* __restrict__ np, const unsigned long long* k, int* rows)
{
    int n = *np;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long v = k[i];
        rows[2 * i] = (int)(v >> 21);
        rows[2 * i + 1] = (int)(v & ((1ull << 21) - 1));
    }
}
__global__ void k_decode_pe(const int* __restrict__ np, const unsigned long long* k, int* rows)
{
    int n = *np;
    unsigned long long m = (1ull << 21) - 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long v = k[i];
        rows[3 * i] = (int)(v >> 42);
        rows[3 * i + 1] = (int)((v >> 21) & m);
        rows[3 * i + 2] = (int)(v & m);
    }
}

// np.unique order: radix sort of the packed keys, unique into uniq; the
// unique count is left on the device (*d_n)
static int dedup(fipc_ctx* c, const unsigned long long* keys, int n, int end_bit, unsigned long long* sorted,
                 unsigned long long* uniq, int* d_n)
{
    size_t b1 = 0, b2 = 0;
    HIPCHK(rocprim::radix_sort_keys(nullptr, b1, keys, sorted, n, 0, end_bit, c->s));
    HIPCHK(rocprim::unique(nullptr, b2, sorted, uniq, d_n, (size_t)n, rocprim::equal_to<unsigned long long>(), c->s));
    size_t need = std::max(b1, b2);
    if (need > c->cub_tmp.cap) STCHK(ensure(c->cub_tmp, need));
    b1 = c->cub_tmp.cap;
    HIPCHK(rocprim::radix_sort_keys(c->cub_tmp.p, b1, keys, sorted, n, 0, end_bit, c->s));
    b2 = c->cub_tmp.cap;
    HIPCHK(rocprim::unique(c->cub_tmp.p, b2, sorted, uniq, d_n, (size_t)n, rocprim::equal_to<unsigned long long>(),
                           c->s));
    return FIPC_OK;
}

// lexicographic order of four-index rows (np.unique order)
struct Row4Less {
    __device__ __forceinline__ bool operator()(const int4& a, const int4& b) const
    {
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        if (a.z != b.z) return a.z < b.z;
        return a.w < b.w;
    }
};
// sorts the n rows of a four-index type in place (through the scratch
// rows_tmp): the row order, and so every later sum over rows, no longer
// depends on the order the search appended them in
static int sort_rows4(fipc_ctx* c, DVec<int>& rows, int n)
{
    if (n < 2) return FIPC_OK;
    STCHK(ensure(c->rows_tmp, (size_t)n * 4));
    const int4* in = reinterpret_cast<const int4*>(rows.p);
    int4* out = reinterpret_cast<int4*>(c->rows_tmp.p);
    size_t b = 0;
    HIPCHK(rocprim::merge_sort(nullptr, b, in, out, (size_t)n, Row4Less(), c->s));
    if (b > c->cub_tmp.cap) STCHK(ensure(c->cub_tmp, b));
    b = c->cub_tmp.cap;
    HIPCHK(rocprim::merge_sort(c->cub_tmp.p, b, in, out, (size_t)n, Row4Less(), c->s));
    std::swap(rows, c->rows_tmp);
    return FIPC_OK;
}

static int bits_for(int n)
{
    int b = 1;
    while ((1 << b) < n) b++;
    return b;
}

// one pass of both searches; raw counts and the status word come back with
// the single synchronisation
static int find_once(fipc_ctx* c, int* counts, int* hflag)
{
    STCHK(bp_begin(c, c->counters.p, 12));
    RowSink sink;
    sink.cnt = c->counters.p;
    // one same-address atomic per wavefront and candidate test would
    // serialise the queries: the candidates are counted only on request
    sink.cand = c->search_stats ? c->counters.p + 10 : nullptr;
    for (int t = 0; t < NTYPES; t++) { sink.cap[t] = c->cap[t]; sink.rows[t] = c->rows[t].p; }
    sink.kpp = c->keys_pp.p;
    sink.kpe = c->keys_pe.p;
    // PT: points inserted with +-dHat, triangles query their box
    BoxFn pts{c->x.p, nullptr, c->bp.p, 0, c->dHat};
    STCHK(grid_build(c, 0, c->nbp, pts));
    BoxFn tris{c->x.p, nullptr, c->bt.p, 2, 0.0};
    VisitPT vpt{c->x.p, c->bp.p, c->bt.p, c->dHat, c->dHat2, sink};
    STCHK(run_query(c, 0, c->nbt, tris, vpt, TM_FIND));
    // EE: edges inserted with their box, queried with box +- dHat
    BoxFn edges{c->x.p, nullptr, c->be.p, 1, 0.0};
    STCHK(grid_build(c, 1, c->nbe, edges));
    BoxFn qedges{c->x.p, nullptr, c->be.p, 1, c->dHat};
    VisitEE vee{c->x.p, c->x0.p, c->be.p, c->dHat, c->dHat2, sink};
    STCHK(run_query(c, 1, c->nbe, qedges, vee, TM_FIND));
    STCHK(d2h_async(c, counts, c->counters.p, 12 * 4));
    STCHK(d2h_async(c, hflag, c->flag.p, 12));
    STCHK(d2h_wait(c));
    return FIPC_OK;
}

extern "C" int fipc_find_constraints(fipc_ctx* c, int* counts_out)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    int raw[12];
    for (int attempt = 0;; attempt++) {
        if (attempt == 6) return FIPC_ERR_CAPACITY;
        STCHK(ensure(c->keys_pp, c->cap[0]));
        STCHK(ensure(c->keys_pe, c->cap[1]));
        for (int t = 2; t < NTYPES; t++) STCHK(ensure(c->rows[t], c->cap[t] * 4));
        int hflag[3];
        STCHK(find_once(c, raw, hflag));
        bool grow = false;
        STCHK(bp_check(c, hflag, &grow));
        if (grow) continue;
        for (int t = 0; t < NTYPES; t++)
            if (raw[t] > c->cap[t]) { c->cap[t] = (int64_t)raw[t] * 2; grow = true; }
        if (!grow) break;
    }
    for (int t = 0; t < NTYPES; t++) c->raw_count[t] = raw[t];
    c->cand_last[0] = c->search_stats ? raw[10] : -1;
    c->cand_last[1] = c->search_stats ? raw[11] : -1;
    int bits = bits_for(c->nv);
    // sorted and unique keys of PP then PE share keys_tmp
    size_t npp_raw = raw[0], npe_raw = raw[1];
    STCHK(ensure(c->keys_tmp, 2 * (npp_raw + npe_raw) + 4));
    unsigned long long* kt = c->keys_tmp.p;
    STCHK(ensure(c->rows[0], std::max<size_t>(npp_raw, 1) * 2));
    STCHK(ensure(c->rows[1], std::max<size_t>(npe_raw, 1) * 3));
    HIPCHK(hipMemsetAsync(c->nuniq.p, 0, 8, c->s));
    if (npp_raw) {
        STCHK(dedup(c, c->keys_pp.p, raw[0], std::min(64, 21 + bits), kt, kt + npp_raw, c->nuniq.p));
        k_decode_pp<<<stride_grid(npp_raw), TPB, 0, c->s>>>(c->nuniq.p, kt + npp_raw, c->rows[0].p);
    }
    if (npe_raw) {
        unsigned long long* k2 = kt + 2 * npp_raw;
        STCHK(dedup(c, c->keys_pe.p, raw[1], std::min(64, 42 + bits), k2, k2 + npe_raw, c->nuniq.p + 1));
        k_decode_pe<<<stride_grid(npe_raw), TPB, 0, c->s>>>(c->nuniq.p + 1, k2 + npe_raw, c->rows[1].p);
    }
    HIPCHK(hipGetLastError());
    for (int t = 2; t < NTYPES; t++) STCHK(sort_rows4(c, c->rows[t], raw[t]));
    int nu[2];
    STCHK(d2h_async(c, nu, c->nuniq.p, 8));
    STCHK(d2h_wait(c));
    c->count[0] = nu[0];
    c->count[1] = nu[1];
    for (int t = 2; t < NTYPES; t++) c->count[t] = raw[t];
    c->sys_version++;   // the contact rows of the SpMV changed
    c->cons_version++;  // and those of the barrier energy
    for (int t = 0; t < NTYPES; t++) {
        int N = 3 * (t == 0 ? 2 : (t == 1 ? 3 : 4));
        STCHK(ensure(c->cH[t], (size_t)std::max(c->count[t], 1) * N * N));
        if (counts_out) counts_out[t] = c->count[t];
    }
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
// Candidate pairs: the grid walk applies the moving-box test of the reference
// and the CCD lower bound (a pair whose bound lb = (1 - eta) d / maxDisp is
// at least 1 keeps toc = 1, wrapper.cpp:254-328), and appends the rest; the
// narrow phase (CTCD) then runs one thread per candidate, so its long,
// data-dependent root finding neither stalls the grid walk nor unbalances
// its waves.
// One atomic per wavefront: the active lanes that append take consecutive slots.
__device__ __forceinline__ int wave_append(int* cnt)
{
    unsigned long long mask = __ballot(1);
    int leader = __ffsll((unsigned long long)mask) - 1;
    int lane = __lane_id();
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt, __popcll(mask));
    base = __shfl(base, leader, 64);
    unsigned long long below = lane == 0 ? 0ull : (mask & (~0ull >> (64 - lane)));
    return base + __popcll(below);
}
struct PairSink {
    int* cnt;
    int cap;
    int2* pairs;
    __device__ __forceinline__ void push(int a, int b) const
    {
        int s = wave_append(cnt);
        if (s < cap) pairs[s] = make_int2(a, b);
    }
};
// compute_filter_3D_PT (NEWTON_IPC.py:377-406); query = triangle, item = point
struct CandCCDPT {
    const double *x, *p;
    const int *bp, *bt;
    double dHat, eta;
    PairSink out;
    // the triangle's vertices and swept box, once per (triangle, cell)
    struct Q {
        int t0, t1, t2;
        v3 mnt, mxt;
    };
    __device__ __forceinline__ Q prep(int tri) const
    {
        Q s;
        s.t0 = bt[3 * tri]; s.t1 = bt[3 * tri + 1]; s.t2 = bt[3 * tri + 2];
        v3 T0 = ldv(x, s.t0), T1 = ldv(x, s.t1), T2 = ldv(x, s.t2);
        v3 d0 = ldv(p, s.t0), d1 = ldv(p, s.t1), d2 = ldv(p, s.t2);
        s.mxt = vmax(vmax(T0, T0 + d0), vmax(vmax(T1, T1 + d1), vmax(T2, T2 + d2)));
        s.mnt = vmin(vmin(T0, T0 + d0), vmin(vmin(T1, T1 + d1), vmin(T2, T2 + d2)));
        return s;
    }
    __device__ __forceinline__ void item(const Q& s, int tri, int pi) const
    {
        const int q = bp[pi], t0 = s.t0, t1 = s.t1, t2 = s.t2;
        if (q == t0 || q == t1 || q == t2) return;
        v3 P = ldv(x, q), dP = ldv(p, q);
        // moving_point_triangle_ccd_broadphase (ipc.py:339-345)
        v3 mxp = vmax(P, P + dP), mnp = vmin(P, P + dP);
        const v3 mxt = s.mxt, mnt = s.mnt;
        if (!(mnp.x < mxt.x + dHat && mnp.y < mxt.y + dHat && mnp.z < mxt.z + dHat)) return;
        if (!(mnt.x - dHat < mxp.x && mnt.y - dHat < mxp.y && mnt.z - dHat < mxp.z)) return;
        v3 T0 = ldv(x, t0), T1 = ldv(x, t1), T2 = ldv(x, t2);
        v3 d0 = ldv(p, t0), d1 = ldv(p, t1), d2 = ldv(p, t2);
        // the lower bound of pt_ccd, same expressions
        double maxDisp = sqrt(n2(dP)) + sqrt(fmax(fmax(n2(d0), n2(d1)), n2(d2)));
        if (maxDisp == 0) return;
        double d2c = PT_dist2(P, T0, T1, T2, PT_type(P, T0, T1, T2));
        if (!((1 - eta) * sqrt(d2c) / maxDisp < 1)) return;
        out.push(tri, pi);
    }
};
// compute_filter_3D_EE (NEWTON_IPC.py:409-440); query edge i, item edge j
struct CandCCDEE {
    const double *x, *p;
    const int* be;
    double dHat, eta;
    PairSink out;
    // the query edge's vertices and swept box, once per (edge, cell)
    struct Q {
        int a0, a1;
        v3 mna, mxa;
    };
    __device__ __forceinline__ Q prep(int i) const
    {
        Q s;
        s.a0 = be[2 * i]; s.a1 = be[2 * i + 1];
        v3 A0 = ldv(x, s.a0), A1 = ldv(x, s.a1), dA0 = ldv(p, s.a0), dA1 = ldv(p, s.a1);
        s.mxa = vmax(vmax(A0, A0 + dA0), vmax(A1, A1 + dA1));
        s.mna = vmin(vmin(A0, A0 + dA0), vmin(A1, A1 + dA1));
        return s;
    }
    __device__ __forceinline__ void item(const Q& s, int i, int j) const
    {
        if (!(i < j)) return;
        const int a0 = s.a0, a1 = s.a1, b0 = be[2 * j], b1 = be[2 * j + 1];
        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
        v3 B0 = ldv(x, b0), B1 = ldv(x, b1), dB0 = ldv(p, b0), dB1 = ldv(p, b1);
        // moving_edge_edge_ccd_broadphase (ipc.py:357-363)
        const v3 mxa = s.mxa, mna = s.mna;
        v3 mxb = vmax(vmax(B0, B0 + dB0), vmax(B1, B1 + dB1)), mnb = vmin(vmin(B0, B0 + dB0), vmin(B1, B1 + dB1));
        if (!(mna.x < mxb.x + dHat && mna.y < mxb.y + dHat && mna.z < mxb.z + dHat)) return;
        if (!(mnb.x - dHat < mxa.x && mnb.y - dHat < mxa.y && mnb.z - dHat < mxa.z)) return;
        v3 A0 = ldv(x, a0), A1 = ldv(x, a1), dA0 = ldv(p, a0), dA1 = ldv(p, a1);
        // the lower bound of ee_ccd, same expressions
        double maxDisp = sqrt(fmax(n2(dA0), n2(dA1))) + sqrt(fmax(n2(dB0), n2(dB1)));
        if (maxDisp == 0) return;
        double d2c = EE_dist2(A0, A1, B0, B1, EE_type(A0, A1, B0, B1));
        if (!((1 - eta) * sqrt(d2c) / maxDisp < 1)) return;
        out.push(i, j);
    }
};
__device__ __forceinline__ void amin_d(unsigned long long* a, double v)
{
    atomicMin(a, (unsigned long long)__double_as_longlong(v));
}
// narrow phase over the candidates (count read on the device, capped at cap:
// an overflowing count makes the host redo the filter with more room).  A
// pair's time of impact is the lesser of its two roundings (strict and
// fused); the two run in different wavefronts (wavefront w takes the pairs
// of group w / 2 of 64 in rounding w % 2), so each wavefront runs one root
// isolation per pair: the minimum over (pair, rounding) is the minimum over
// pairs of the lesser, bit for bit, with twice the wavefronts in flight.
__device__ __forceinline__ bool ccd_slot(long t, int n, int* i, bool* fused)
{
    *i = (int)((t >> 7) * 64 + (t & 63));
    *fused = (t >> 6) & 1;
    return *i < n;
}
__global__ void __launch_bounds__(TPB) k_ccd_pt(const int* __restrict__ cnt, int cap, const int2* __restrict__ pairs,
                                                const double* __restrict__ x, const double* __restrict__ p,
                                                const int* __restrict__ bp, const int* __restrict__ bt, double eta,
                                                unsigned long long* amin)
{
    const int n = min(*cnt, cap);
    const long nt = 2L * ((n + 63) / 64) * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (long)gridDim.x * blockDim.x) {
        int i;
        bool fused;
        if (!ccd_slot(t, n, &i, &fused)) continue;
        int2 pr = pairs[i];
        int q = bp[pr.y], t0 = bt[3 * pr.x], t1 = bt[3 * pr.x + 1], t2 = bt[3 * pr.x + 2];
        v3 P = ldv(x, q), T0 = ldv(x, t0), T1 = ldv(x, t1), T2 = ldv(x, t2);
        v3 dP = ldv(p, q), d0 = ldv(p, t0), d1 = ldv(p, t1), d2 = ldv(p, t2);
        double d2c = PT_dist2(P, T0, T1, T2, PT_type(P, T0, T1, T2));
        double toc = fused ? ccd_fused::pt_ccd(P, T0, T1, T2, dP, d0, d1, d2, eta, d2c)
                           : ccd_strict::pt_ccd(P, T0, T1, T2, dP, d0, d1, d2, eta, d2c);
        if (toc < 1.0) amin_d(amin, toc);
    }
}
__global__ void __launch_bounds__(TPB) k_ccd_ee(const int* __restrict__ cnt, int cap, const int2* __restrict__ pairs,
                                                const double* __restrict__ x, const double* __restrict__ p,
                                                const int* __restrict__ be, double eta, unsigned long long* amin)
{
    const int n = min(*cnt, cap);
    const long nt = 2L * ((n + 63) / 64) * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (long)gridDim.x * blockDim.x) {
        int i;
        bool fused;
        if (!ccd_slot(t, n, &i, &fused)) continue;
        int2 pr = pairs[i];
        int a0 = be[2 * pr.x], a1 = be[2 * pr.x + 1], b0 = be[2 * pr.y], b1 = be[2 * pr.y + 1];
        v3 A0 = ldv(x, a0), A1 = ldv(x, a1), B0 = ldv(x, b0), B1 = ldv(x, b1);
        v3 dA0 = ldv(p, a0), dA1 = ldv(p, a1), dB0 = ldv(p, b0), dB1 = ldv(p, b1);
        double d2c = EE_dist2(A0, A1, B0, B1, EE_type(A0, A1, B0, B1));
        double toc = fused ? ccd_fused::ee_ccd(A0, A1, B0, B1, dA0, dA1, dB0, dB1, eta, d2c)
                           : ccd_strict::ee_ccd(A0, A1, B0, B1, dA0, dA1, dB0, dB1, eta, d2c);
        if (toc < 1.0) amin_d(amin, toc);
    }
}

extern int fipc_inversion_alpha(fipc_ctx* c, unsigned long long* dst);

// compute_intersection_free_step_size (NEWTON_IPC.py:456-468)
extern "C" int fipc_ccd_alpha(fipc_ctx* c, double eta, int inversion, double* alpha, double* parts)
{
    if (!c || !c->nv || !alpha) return FIPC_ERR_STATE;
    double a[3];
    for (int attempt = 0;; attempt++) {
        if (attempt == 6) return FIPC_ERR_CAPACITY;
        int cap = c->ccd_cap;
        STCHK(ensure(c->ccd_pairs[0], cap)); STCHK(ensure(c->ccd_pairs[1], cap));
        int* pc = c->counters.p + 8;
        STCHK(bp_begin(c, pc, 2, c->amin.p, 3));
        // the edge-edge half (the longer one) runs on the second stream while
        // the point-triangle half and the inversion bound run on the first:
        // separate grids, pair lists, counters and minima; the first waits
        // for the second before the readback
        HIPCHK(hipEventRecord(c->ev_rhs, c->s));
        HIPCHK(hipStreamWaitEvent(c->s2, c->ev_rhs, 0));
        BoxFn edges{c->x.p, c->sol.p, c->be.p, 1, 0.0};
        STCHK(grid_build(c, 1, c->nbe, edges, c->s2));
        BoxFn qedges{c->x.p, c->sol.p, c->be.p, 1, c->dHat};
        CandCCDEE vee{c->x.p, c->sol.p, c->be.p, c->dHat, eta, PairSink{pc + 1, cap, c->ccd_pairs[1].p}};
        STCHK(run_query(c, 1, c->nbe, qedges, vee, TM_CCD, c->s2));
        STCHK(fipc_tm_begin(c, TM_NARROW, c->s2));
        k_ccd_ee<<<stride_grid(2L * cap), TPB, 0, c->s2>>>(pc + 1, cap, c->ccd_pairs[1].p, c->x.p, c->sol.p, c->be.p, eta,
                                                      c->amin.p + 1);
        HIPCHK(hipGetLastError());
        STCHK(fipc_tm_end(c, TM_NARROW, c->s2));
        HIPCHK(hipEventRecord(c->ev_contact, c->s2));
        BoxFn pts{c->x.p, c->sol.p, c->bp.p, 0, c->dHat};
        STCHK(grid_build(c, 0, c->nbp, pts));
        BoxFn tris{c->x.p, c->sol.p, c->bt.p, 2, 0.0};
        CandCCDPT vpt{c->x.p, c->sol.p, c->bp.p, c->bt.p, c->dHat, eta, PairSink{pc, cap, c->ccd_pairs[0].p}};
        STCHK(run_query(c, 0, c->nbt, tris, vpt, TM_CCD));
        STCHK(fipc_tm_begin(c, TM_NARROW));
        k_ccd_pt<<<stride_grid(2L * cap), TPB, 0, c->s>>>(pc, cap, c->ccd_pairs[0].p, c->x.p, c->sol.p, c->bp.p, c->bt.p, eta,
                                                     c->amin.p);
        HIPCHK(hipGetLastError());
        STCHK(fipc_tm_end(c, TM_NARROW));
        if (inversion) STCHK(fipc_inversion_alpha(c, c->amin.p + 2));
        HIPCHK(hipStreamWaitEvent(c->s, c->ev_contact, 0));
        unsigned long long out[3];
        int hflag[3], hc[2];
        STCHK(d2h_async(c, out, c->amin.p, 24));
        STCHK(d2h_async(c, hflag, c->flag.p, 12));
        STCHK(d2h_async(c, hc, pc, 8));
        STCHK(d2h_wait(c));
        bool retry = false;
        STCHK(bp_check(c, hflag, &retry));
        int need = std::max(hc[0], hc[1]);
        if (need > cap) {
            c->ccd_cap = need + need / 2 + 1024;
            retry = true;
        }
        if (retry) continue;
        memcpy(a, out, 24);
        c->ccd_last[0] = hc[0];
        c->ccd_last[1] = hc[1];
        break;
    }
    if (parts) { parts[0] = a[0]; parts[1] = a[1]; parts[2] = a[2]; }
    *alpha = std::min(std::min(a[0], a[1]), a[2]);
    return FIPC_OK;
}

extern "C" int fipc_search_stats(fipc_ctx* c, int on)
{
    if (!c) return FIPC_ERR_ARG;
    c->search_stats = on != 0;
    return FIPC_OK;
}

extern "C" int fipc_search_counts(fipc_ctx* c, int* cand)
{
    if (!c || !cand) return FIPC_ERR_ARG;
    cand[0] = c->cand_last[0];
    cand[1] = c->cand_last[1];
    return FIPC_OK;
}

extern "C" int fipc_ccd_counts(fipc_ctx* c, int* pairs)
{
    if (!c || !pairs) return FIPC_ERR_ARG;
    pairs[0] = c->ccd_last[0];
    pairs[1] = c->ccd_last[1];
    return FIPC_OK;
}

// ---- collision check --------------------------------------------------------
struct VisitCollide {
    const double* x;
    const int *be, *bt;
    int* hit;
    // the triangle, once per (triangle, cell)
    struct Q {
        int t0, t1, t2;
        v3 T0, T1, T2;
    };
    __device__ __forceinline__ Q prep(int tri) const
    {
        Q s;
        s.t0 = bt[3 * tri]; s.t1 = bt[3 * tri + 1]; s.t2 = bt[3 * tri + 2];
        s.T0 = ldv(x, s.t0); s.T1 = ldv(x, s.t1); s.T2 = ldv(x, s.t2);
        return s;
    }
    __device__ __forceinline__ void item(const Q& s, int tri, int ei) const
    {
        const int a0 = be[2 * ei], a1 = be[2 * ei + 1], t0 = s.t0, t1 = s.t1, t2 = s.t2;
        if (a0 == t0 || a0 == t1 || a0 == t2 || a1 == t0 || a1 == t1 || a1 == t2) return;
        if (seg_tri(ldv(x, a0), ldv(x, a1), s.T0, s.T1, s.T2)) atomicOr(hit, 1);
    }
};

// check_collision (NEWTON_IPC.py:855-886)
extern "C" int fipc_check_collision(fipc_ctx* c, int* hit)
{
    if (!c || !c->nv || !hit) return FIPC_ERR_STATE;
    for (int attempt = 0;; attempt++) {
        if (attempt == 6) return FIPC_ERR_CAPACITY;
        STCHK(bp_begin(c));
        // the edge grid of the constraint search (boxes without margin), so
        // either reuses the other's at the same x; the margin d-hat moves to
        // the triangle queries, which pairs the same boxes
        BoxFn edges{c->x.p, nullptr, c->be.p, 1, 0.0};
        STCHK(grid_build(c, 1, c->nbe, edges));
        BoxFn tris{c->x.p, nullptr, c->bt.p, 2, c->dHat};
        VisitCollide v{c->x.p, c->be.p, c->bt.p, c->flag.p + 1};
        STCHK(run_query(c, 1, c->nbt, tris, v, TM_FIND));
        int hflag[3];
        STCHK(d2h_async(c, hflag, c->flag.p, 12));
        STCHK(d2h_wait(c));
        bool retry = false;
        STCHK(bp_check(c, hflag, &retry));
        if (retry) continue;
        *hit = hflag[1];
        return FIPC_OK;
    }
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

