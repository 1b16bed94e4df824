// fipc_ctx.h -- device-resident state of one simulation (one context per
// mesh).  Buffers are owned by the context; every entry point is
// stream-ordered on ctx->s and returns a status code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include "fipc_math.h"
#include "../../include/fastipc.h"

#define NTYPES 7

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            fprintf(stderr, "fastipc: HIP error %s at %s:%d\n", hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                              \
            return FIPC_ERR_HIP;                                                      \
        }                                                                             \
    } while (0)

#define STCHK(expr)                       \
    do {                                  \
        int _st = (expr);                 \
        if (_st != FIPC_OK) return _st;   \
    } while (0)

template <class T>
struct DVec {
    T* p = nullptr;
    size_t cap = 0;
};

template <class T>
static inline int ensure(DVec<T>& v, size_t n)
{
    if (n == 0) n = 1;
    if (v.cap >= n) return FIPC_OK;
    if (v.p) HIPCHK(hipFree(v.p));
    v.p = nullptr;
    size_t c = n + n / 4 + 64;
    HIPCHK(hipMalloc(&v.p, c * sizeof(T)));
    v.cap = c;
    return FIPC_OK;
}
template <class T>
static inline void release(DVec<T>& v)
{
    if (v.p) (void)hipFree(v.p);
    v.p = nullptr;
    v.cap = 0;
}

// hashed uniform grid used by every broad phase
// Broad-phase status word (flag[0]) bits: GF_RETRY = a grid had more entries
// than its capacity (the needed count is in flag[2]; the caller grows the
// capacity and redoes the operation), GF_HUGE = one box spans more cells than
// the limit (an error).
#define GF_RETRY 1
#define GF_HUGE 2
struct Grid {
    unsigned M = 0;        // bucket count (power of two, >= cap)
    int cap = 0;           // entry capacity; grows when a build overflows
    DVec<int> cnt;         // per item: first entry within its block of items (two-level offsets)
    DVec<int> part;        // per block of items: its entries; part[blocks] = all entries
    DVec<unsigned> ekey;   // bucket of each entry
    DVec<int> eitem;       // item | lowest-cell flags << 28
    DVec<unsigned long long> ecell, cells;  // exact cell key per entry / ordered by bucket
    DVec<int> bstart;      // M+1 bucket offsets (bucket ends after the scatter)
    DVec<int> items;       // entries ordered by bucket
    DVec<int> qoff, qpart; // query cell offsets within blocks of queries, per-block totals
    // what the grid holds, for reuse by a later operation: the x version,
    // primitive kind, count and inflation of a grid of boxes at x; for a grid
    // of swept boxes (x to x + sol), the x and sol versions it sweeps between
    // (swept): any x = x + alpha sol, 0 <= alpha <= 1, has its boxes inside
    // the swept ones, so a query at that x may use it (the candidates are a
    // superset and the visitors test exactly)
    bool reusable = false, swept = false;
    unsigned long long x_version = 0, seg_sol = 0;
    int kind = -1, n = -1;
    double infl = 0.0;
};

// HIP-event timing of single kernels: pairs are recorded on the context
// stream without synchronising and folded into totals when read
#define FIPC_NTIMERS 11
enum { TM_SPMV = 0, TM_ELASTIC, TM_CCD, TM_FIND, TM_ENERGY, TM_CONTACT, TM_PCGVEC, TM_GRID, TM_GATHER, TM_SPMV_RES,
       TM_NARROW };
struct KernelTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    long launches = 0;
};

struct fipc_ctx {
    hipStream_t s = nullptr;
    int nv = 0, nt = 0, nbp = 0, nbe = 0, nbt = 0, ndir = 0;
    // mesh
    DVec<int> tets, bp, be, bt, dids;
    DVec<double> IB, vol, la, mu, m;
    DVec<unsigned char> dmask;
    // state
    DVec<double> x, x0, xn, xt, xprev, v, sol, rhs, target;
    double dt = 0.01, dHat = 1e-3, dHat2 = 1e-6, kappa = 1.0, inv_dx = 1.0, gravity = -9.8;
    // elastic + inertia Hessian, symmetric: only the upper-triangle 3x3
    // blocks are stored (row i holds columns j >= i, the diagonal block
    // first), 9 values per block row-major (val[9 b + 3 r + c]); for every
    // row i, lent[lptr[i] .. lptr[i + 1]) lists the pairs (block, j) of the
    // blocks (j, i), j < i, stored in row j, which the SpMV applies
    // transposed.  nnzb_full counts both triangles.
    int nnzb = 0, nnzb_full = 0, nlower = 0;
    DVec<int> rowptr, col, diag, brow, lptr, lent;
    DVec<double> val;
    DVec<float> valf;
    // assembly: the per-tet pass writes each of a tet's 10 upper blocks
    // (oriented to its global block) to cpos[10 e + slot] of a scratch sorted
    // by block, so the gather sums contiguous runs bcptr[b] .. bcptr[b + 1];
    // the gradient is gathered through vertex -> (corner, tet) lists
    DVec<double> tetH, tetg;  // FP64 mode
    DVec<int> cpos, bcptr, vcptr, vcent;
    DVec<int> gorder;      // the gather's block order: diagonal blocks first
    // FP32 mode's tiled pass (k_elastic_tile): nchunk chunks of consecutive
    // tets, TILE_META int4 each; per (slot, tet) and (corner, tet) the LDS
    // index | colour << 8 of the entry it adds to; per (chunk, block) and
    // (chunk, vertex) the position of the partial record, whose runs per
    // block / vertex are tile_pptr / tile_vptr
    int nchunk = 0;
    long tile_nbpart = 0, tile_nvpart = 0;
    DVec<int4> tile_meta;
    DVec<unsigned short> tile_tcode, tile_vcode;
    DVec<int> tile_ppos, tile_vpos, tile_pptr, tile_vptr;
    DVec<float> tile_bpart;
    DVec<double> tile_gpart;
    int precision = 0;     // 0 = FP64 Hessian, 1 = FP32 Hessian / SpMV
    // contacts
    int64_t cap[NTYPES] = {0};
    int count[NTYPES] = {0};
    int raw_count[NTYPES] = {0};
    DVec<int> rows[NTYPES];
    DVec<int> rows_tmp;    // the sort of the four-index types writes here, then swaps
    DVec<double> cH[NTYPES];
    // FP32 Hessian: FP32 copies of the contact rows' projected Hessians for
    // the FP32 iterations, and per row a flag (1) for the rows closer than
    // 1e-2 dHat, whose FP64 Hessian those iterations keep using
    DVec<float> cHf[NTYPES];
    DVec<unsigned char> c64[NTYPES];
    // Contact vertices: every (row, vertex) slot of the seven types, slot
    // s = soff[t] + K_t r + a, keyed by its vertex and stably sorted, so the
    // slots of a vertex are contiguous and in (type, row, vertex) order;
    // cv_vert[u] lists the non-Dirichlet vertices with slots, cv_ptr[u] ..
    // cv_ptr[u + 1] their range in cv_slot, cv_count[0] = their number.
    // Sums over a vertex's slots run in this fixed order (no atomics), so
    // the contact parts of the gradient, of the diagonal blocks and of every
    // product are the same bits from run to run.  Rebuilt by the assembly
    // when the rows changed (cv_version = cons_version of the build).
    int soff[NTYPES + 1] = {0};
    DVec<int> cv_keys, cv_vals, cv_keys2, cv_slot, cv_head, cv_vert, cv_ptr, cv_count;
    DVec<double> cgrad;    // per slot the row's gradient at that vertex (3)
    unsigned long long cv_version = 0;
    DVec<unsigned long long> keys_pp, keys_pe, keys_tmp;
    DVec<int> counters;    // NTYPES raw append counters + misc
    DVec<unsigned char> cub_tmp, cub_tmp2;  // rocPRIM scan / sort scratch of the first and second stream
    size_t cub_bytes = 0;
    DVec<int> nuniq;
    // preconditioner + PCG
    DVec<double> cdiag, Minv, r, z, pp, q, qc;   // q = A_bsr p, qc = contact part of A p
    // FP32 Hessian: the refinement rounds iterate on FP32 copies of the
    // preconditioner and of r, z, p, q (qc and the solution stay FP64)
    DVec<float> Minvf, rf, zf, pf, qf;
    // single-reduction PCG: s = A p kept by recurrence (FP64 / FP32 rounds),
    // and the partials of r.u and r.r of two consecutive iterations
    DVec<double> cgs, part_g;
    DVec<float> cgsf;
    DVec<float> dxf;  // an FP32 round's correction to the solution (zero between rounds)
    // prescribed rows of the step (fipc_begin_step_rows)
    DVec<int> trow;
    DVec<double> tval;
    DVec<double> part_a, part_b, part_c, scal;
    DVec<int> x0flag;      // -1 when the PCG start x0 (Dirichlet increments) is zero
    hipEvent_t ws_done = nullptr;  // end of the last warm-start prediction on this stream
    // second stream: the contact assembly, overlapped with the tet pass
    // (ev_rhs: inertia term and cleared diagonal blocks ready; ev_contact:
    // contact rows assembled)
    hipStream_t s2 = nullptr;
    hipEvent_t ev_rhs = nullptr, ev_contact = nullptr;
    // max |sol| read back at the end of a solve; sol_version counts the
    // writes of sol, res_version is the version res_cached belongs to
    double res_cached = 0.0;
    unsigned long long sol_version = 1, res_version = 0;
    // elastic and barrier energy of the last evaluation, valid for its x
    // version and constraint-set version (cons_version: bumped by every
    // constraint search); parameter changes clear e_ok
    unsigned long long cons_version = 1, e_xv = 0, e_cv = 0;
    double e_el = 0.0, e_bar = 0.0;
    bool e_ok = false;
    DVec<int> pcg_state;   // [0] done_at, [1] iterations
    // reductions for energy / residual / ccd
    DVec<double> epart;
    DVec<unsigned long long> amin;
    DVec<int> flag;        // [0] broad-phase status, [1] collision hit, [2] needed grid entries
    // CCD candidate pairs (point-triangle, edge-edge) between the grid walk and the narrow phase
    int ccd_cap = 0;
    DVec<int2> ccd_pairs[2];
    int ccd_last[2] = {0, 0};  // candidate pairs (PT, EE) of the last fipc_ccd_alpha
    int cand_last[2] = {-1, -1}; // broad-phase candidates (PT, EE) of the last fipc_find_constraints
    bool search_stats = false;   // count them (fipc_search_stats)
    // broad phase grids
    Grid grid[2];                         // [0] boundary points, [1] boundary edges
    unsigned long long x_version = 1;     // incremented by every write of x
    // xprev holds x of version xprev_of; the x written by the last
    // fipc_apply_sol with 0 <= alpha <= 1 (version seg_x) lies on the segment
    // from that x along sol (version seg_sol); seg_x = 0 otherwise
    unsigned long long xprev_of = 0, seg_x = 0, seg_base = 0, seg_sol = 0;
    // learned warm start
    DVec<double> dprev;    // previous step's correction x_{n} - xTilde_{n-1} (0 on Dirichlet vertices)
    DVec<float> wsX;       // features, nv x 16
    DVec<unsigned long long> ckey;  // nearest contact row per vertex
    DVec<int> free_ids;    // vertices that are not Dirichlet (built on first capture)
    DVec<int> near_ids, nnear;  // free vertices within d-hat of a contact primitive, their count
    int nfree = -1;
    DVec<int> cap_ids;     // vertices captured at the start of the step
    DVec<float> cap_X, cap_T;
    int cap_n = 0;
    // statistics
    double last_pcg_relres = 0.0;
    int last_pcg_iters = 0;
    int pcg_hint = 0;      // iterations of the last first round, sizes the first chunk
    int refine_hint = 0;   // iterations of the first round of the last continuation
    // fipc_assemble bumps sys_version; fipc_solve records the version it
    // solved, which fipc_solve_refine requires
    unsigned long long sys_version = 1, solve_version = 0;
    // the last solve ended on an FP32-mode restart (r, z, p and the scalars
    // at its solution, FP64 residual) that a continuation can iterate from
    bool restart_ready = false;
    // certified Newton test (fipc_solve_certify): while cert_bound > 0 a
    // continuation stops once max|sol| + |b - A sol|_2 / cert_lam < cert_bound;
    // cert_lam is half the least lumped mass of a free vertex (A >= M)
    double cert_bound = 0.0, cert_lam = 0.0, cert_x = 0.0;
    bool cert_hit = false;
    int cert_fail = 0;
    int cert_hint = 0;  // iterations the last certificate's round took
    KernelTimer tm[FIPC_NTIMERS];
    std::vector<hipEvent_t> event_pool;
    bool timing = false;
    // pinned host staging of the small device-to-host readbacks (counts,
    // flags, PCG state and scalars, reduction partials): a copy into pageable
    // memory is staged by the runtime and returns only once the stream has
    // reached it, so two readbacks cost two round trips; into pinned memory
    // they queue behind the kernels and one synchronisation collects them
    char* pin = nullptr;
    size_t pin_used = 0;
    struct PendingCopy {
        void* dst;
        size_t off, n;
    };
    std::vector<PendingCopy> pend;
};

#define FIPC_PIN_BYTES (256u << 10)
// queues the copy of n bytes from device memory src to host memory dst;
// dst is written by the next d2h_wait (copies that do not fit the staging
// area go straight to dst and complete by then too)
static inline int d2h_async(fipc_ctx* c, void* dst, const void* src, size_t n)
{
    size_t off = (c->pin_used + 63) & ~(size_t)63;
    if (!c->pin || off + n > FIPC_PIN_BYTES) {
        HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->s));
        return FIPC_OK;
    }
    hipError_t e = hipMemcpyAsync(c->pin + off, src, n, hipMemcpyDeviceToHost, c->s);
    if (e != hipSuccess) {  // no stale destinations are left queued
        c->pend.clear();
        c->pin_used = 0;
        HIPCHK(e);
    }
    c->pend.push_back({dst, off, n});
    c->pin_used = off + n;
    return FIPC_OK;
}
// waits for the stream and delivers the queued readbacks
static inline int d2h_wait(fipc_ctx* c)
{
    hipError_t e = hipStreamSynchronize(c->s);
    if (e == hipSuccess)
        for (const auto& q : c->pend) memcpy(q.dst, c->pin + q.off, q.n);
    c->pend.clear();
    c->pin_used = 0;
    HIPCHK(e);
    return FIPC_OK;
}

int fipc_tm_begin(fipc_ctx* c, int which, hipStream_t st = nullptr);
int fipc_tm_end(fipc_ctx* c, int which, hipStream_t st = nullptr);
int fipc_tm_fold(fipc_ctx* c);
