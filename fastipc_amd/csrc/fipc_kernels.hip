// fipc_kernels.hip -- assembly, solve, energy and state kernels of the IPC
// Newton step for gfx950, and the context entry points that drive them.
//
// Data layout in HBM (per context):
//   x, x0, xn, xTilde, xPrev, v, p, rhs : float64 [nv][3]
//   elastic + inertia Hessian          : symmetric BSR, upper-triangle 3x3
//                                        blocks only, val[9 b + 3 r + c], so a
//                                        block read in either orientation is
//                                        one contiguous 36 / 72 byte run
//   contact Hessians                   : per row, the projected (3k x 3k)
//                                        matrix row-major, k = 2, 3, 4
//   block-Jacobi preconditioner        : float64 [nv][9] inverse blocks
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include "fipc_ctx.h"
#include "fipc_rows.h"

using namespace fipc;

#define TPB 256

static inline int nblk(long n, int t = TPB) { return (int)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double warp_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double warp_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
// deterministic block sum (fixed tree); result valid in thread 0
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double sh[TPB / 64];
    v = warp_sum(v);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
    __syncthreads();
    return r;
}
__device__ __forceinline__ double block_max(double v)
{
    __shared__ double sh[TPB / 64];
    v = warp_max(v);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = fmax(r, sh[i]);
    __syncthreads();
    return r;
}
// every block reduces the same partial array in the same order
__device__ __forceinline__ double reduce_partials(const double* part, int n)
{
    __shared__ double bc;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    s = block_sum(s);
    if (threadIdx.x == 0) bc = s;
    __syncthreads();
    double r = bc;
    __syncthreads();
    return r;
}
// two partial arrays reduced in one pass (their loads overlap and the block
// synchronises once); every thread receives both sums, identical in every
// block.  Called at most once per kernel.
__device__ __forceinline__ void reduce_partials2(const double* a, const double* b, int n, double& ra, double& rb)
{
    __shared__ double sh[2][TPB / 64];
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { s0 += a[i]; s1 += b[i]; }
    s0 = warp_sum(s0);
    s1 = warp_sum(s1);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = s0; sh[1][w] = s1; }
    __syncthreads();
    ra = 0.0; rb = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) { ra += sh[0][i]; rb += sh[1][i]; }
}

// Cross-block reductions never synchronise blocks inside a kernel.  On gfx950
// the L2 is private to each XCD, so an agent-scope release / acquire pair
// (a __threadfence before a ticket atomic) writes back and invalidates the
// XCD's whole L2 and destroys the cache for every other block.  Instead each
// block stores its partial with a plain store; the next kernel on the stream
// (the kernel boundary makes the stores visible) reduces the partials with
// reduce_partials, redundantly in every block and in index order, so every
// block sees the same bits.

// Workgroups are dispatched round-robin over the 8 XCDs.  For a grid whose
// size is a multiple of 8 this maps the hardware block index to a logical one
// such that each XCD runs a contiguous range of logical blocks, which keeps
// neighbouring rows (and the vertices they gather) in one XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) { return (b & 7) * (nb >> 3) + (b >> 3); }

// ---------------------------------------------------------------------------
// step state
// ---------------------------------------------------------------------------
// compute_xn_and_xTilde (NEWTON_IPC.py:497-502) + move_nodes (:505-511)
__global__ void k_begin_step(int nv, const double* __restrict__ x, const double* __restrict__ v, double* xn,
                             double* xt, const unsigned char* dmask, const double* target, double dt, double g)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    for (int d = 0; d < 3; d++) {
        double xi = x[3 * i + d];
        xn[3 * i + d] = xi;
        double t = xi + dt * v[3 * i + d];
        if (d == 1) t += dt * dt * g;
        xt[3 * i + d] = dmask[i] ? target[3 * i + d] : t;
    }
}
// compute_v (NEWTON_IPC.py:799-802)
// (also keeps x - xTilde of free vertices, a feature of the learned warm start)
__global__ void k_end_step(int n3, const double* x, const double* xn, const double* xt, const unsigned char* dmask,
                           double* v, double* dprev, double inv_dt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    v[i] = (x[i] - xn[i]) * inv_dt;
    dprev[i] = dmask[i / 3] ? 0.0 : x[i] - xt[i];
}
// apply_sol (NEWTON_IPC.py:792-796)
__global__ void k_apply(int n3, const double* xp, const double* p, double* x, double a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) x[i] = xp[i] + p[i] * a;
}
// output_residual (NEWTON_IPC.py:805-812): block maxima of |p|
__global__ void k_maxabs(int n, const double* p, double* part)
{
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, fabs(p[i]));
    m = block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// ---------------------------------------------------------------------------
// assembly: inertia + elasticity into the BSR, contacts into per-row blocks
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tet_load(const int* tets, const double* IB, const double* x, int e, int (&t)[4],
                                         double (&F)[9], double (&Bm)[4][3])
{
    for (int a = 0; a < 4; a++) t[a] = tets[4 * e + a];
    double ib[9];
    for (int k = 0; k < 9; k++) ib[k] = IB[9 * e + k];
    v3 x0 = ldv(x, t[0]);
    v3 d1 = ldv(x, t[1]) - x0, d2 = ldv(x, t[2]) - x0, d3 = ldv(x, t[3]) - x0;
    double Ds[9] = {d1.x, d2.x, d3.x, d1.y, d2.y, d3.y, d1.z, d2.z, d3.z};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) F[i * 3 + j] = Ds[i * 3] * ib[j] + Ds[i * 3 + 1] * ib[3 + j] + Ds[i * 3 + 2] * ib[6 + j];
    for (int j = 0; j < 3; j++) {
        Bm[1][j] = ib[j];
        Bm[2][j] = ib[3 + j];
        Bm[3][j] = ib[6 + j];
        Bm[0][j] = -ib[j] - ib[3 + j] - ib[6 + j];
    }
}

// index of the upper 3x3 block (a <= c) of a tet's 12x12 Hessian, 0..9
__host__ __device__ constexpr int uslot(int a, int c) { return a * 4 - a * (a - 1) / 2 + (c - a); }

// compute_elasticity (NEWTON_IPC.py:568-691), first pass in FP64 mode: one
// thread per tet writes each of its ten upper 3x3 blocks, oriented as the
// stored global block, to tetH[cpos[10 e + slot] * 9 ..] (positions sorted by
// block, so the gather reads contiguous runs and sums every block's tets in
// index order) and its gradient to tetg[(corner * nt + e) * 3 + i]
// (corner-major: a wave's stores are contiguous)
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(3)))
k_elastic(int nt, const int* __restrict__ tets, const double* __restrict__ IB, const double* __restrict__ vol,
          const double* __restrict__ la, const double* __restrict__ mu, const double* __restrict__ x, double dt2,
          const int* __restrict__ cpos, double* __restrict__ tetH, double* __restrict__ tetg)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    int t[4];
    double F[9], Bm[4][3];
    tet_load(tets, IB, x, e, t, F, Bm);
    double sc = dt2 * vol[e];
    double lae = la[e], mue = mu[e];
    double P[9];
    nh_P(F, lae, mue, P);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < 3; i++)
            tetg[((size_t)a * nt + e) * 3 + i] = (P[i * 3] * Bm[a][0] + P[i * 3 + 1] * Bm[a][1] + P[i * 3 + 2] * Bm[a][2]) * sc;
    NHFactor f;
    nh_factor(F, lae, mue, f);
    // block (a, c), a <= c, belongs to the upper block (min, max) of the global
    // vertices: transposed when the global order is reversed
    tet_hessian_blocks(f, Bm, sc, [&](int a, int c, const double* blk) {
        double* o = tetH + (size_t)cpos[(size_t)e * 10 + uslot(a, c)] * 9;
        if (t[a] <= t[c]) {
#pragma unroll
            for (int k = 0; k < 9; k++) o[k] = blk[k];
        } else {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) o[i * 3 + j] = blk[j * 3 + i];
        }
    });
}

// FP32 mode, tiled: one wavefront per chunk of up to TILE_T consecutive tets
// whose distinct blocks and vertices fit the wave's LDS (the chunks are cut
// at set-up).  The wave sums its tets' oriented FP32 blocks per distinct
// block, and their FP64 corner gradients per distinct vertex, in LDS, then
// writes one partial record per (chunk, block) and (chunk, vertex) to a
// scratch sorted by block / vertex (chunks in order within a run), about a
// third / a fifth of the per-tet records on lattice-ordered meshes.  The sums
// are in a fixed order: slot by slot (corner by corner), and within a slot
// the tets that hit the same entry add one after another in index order
// ("colours" from the set-up tables), so no atomics are needed.
constexpr int TILE_T = 64;       // tets per chunk: one per lane
constexpr int TILE_MAXB = 256;   // distinct blocks per chunk (9 floats of LDS each)
constexpr int TILE_MAXV = 128;   // distinct vertices per chunk (3 doubles of LDS each)
constexpr int TILE_META = 3;     // int4 per chunk: {e0, n, nb, nv}, {pbase, vbase, -, -}, colour counts
static_assert(TPB % 64 == 0, "whole wavefronts per workgroup");

// orders the LDS accesses of the wave's lanes around a colour step (LDS
// instructions of one wavefront execute in issue order; this keeps the
// compiler from moving them across the step)
__device__ __forceinline__ void wave_step()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
template <int N, class A>
__device__ __forceinline__ void tile_add(A* lds, unsigned code, int ncol, const A (&v)[N], bool act)
{
    const int li = (int)(code & 0xffu), col = (int)(code >> 8);
    for (int cc = 0; cc < ncol; cc++) {
        if (act && col == cc) {
#pragma unroll
            for (int q = 0; q < N; q++) lds[N * li + q] += v[q];
        }
        wave_step();
    }
}

__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(3)))
k_elastic_tile(int nchunk, int nt, const int4* __restrict__ meta, const unsigned short* __restrict__ tcode,
               const unsigned short* __restrict__ vcode, const int* __restrict__ ppos, const int* __restrict__ vpos,
               const int* __restrict__ tets, const double* __restrict__ IB, const double* __restrict__ vol,
               const double* __restrict__ la, const double* __restrict__ mu, const double* __restrict__ x,
               double dt2, float* __restrict__ bpart, double* __restrict__ gpart)
{
    __shared__ float lb[TPB / 64][TILE_MAXB * 9];
    __shared__ double lg[TPB / 64][TILE_MAXV * 3];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)blockIdx.x * (TPB / 64) + wv);
    if (w >= nchunk) return;
    const int4 m0 = meta[TILE_META * w], m1 = meta[TILE_META * w + 1], m2 = meta[TILE_META * w + 2];
    const int e0 = __builtin_amdgcn_readfirstlane(m0.x), n = __builtin_amdgcn_readfirstlane(m0.y);
    const int nb = __builtin_amdgcn_readfirstlane(m0.z), nvx = __builtin_amdgcn_readfirstlane(m0.w);
    const int pbase = __builtin_amdgcn_readfirstlane(m1.x), vbase = __builtin_amdgcn_readfirstlane(m1.y);
    // colour counts: bytes 0..9 the block slots, 10..13 the corners
    const unsigned cw[4] = {(unsigned)__builtin_amdgcn_readfirstlane(m2.x), (unsigned)__builtin_amdgcn_readfirstlane(m2.y),
                            (unsigned)__builtin_amdgcn_readfirstlane(m2.z), (unsigned)__builtin_amdgcn_readfirstlane(m2.w)};
    float* B = lb[wv];
    double* G = lg[wv];
    for (int i = lane; i < 9 * nb; i += 64) B[i] = 0.0f;
    for (int i = lane; i < 3 * nvx; i += 64) G[i] = 0.0;
    wave_step();
    const bool act = lane < n;
    const int e = act ? e0 + lane : e0;  // idle lanes recompute the first tet and add nothing
    int t[4];
    double F[9], Bm[4][3];
    tet_load(tets, IB, x, e, t, F, Bm);
    const double sc = dt2 * vol[e];
    const double lae = la[e], mue = mu[e];
    double P[9];
    nh_P(F, lae, mue, P);
#pragma unroll
    for (int a = 0; a < 4; a++) {
        double g[3];
#pragma unroll
        for (int i = 0; i < 3; i++) g[i] = (P[i * 3] * Bm[a][0] + P[i * 3 + 1] * Bm[a][1] + P[i * 3 + 2] * Bm[a][2]) * sc;
        const int k = 10 + a;
        tile_add<3, double>(G, vcode[(size_t)a * nt + e], (int)((cw[k >> 2] >> (8 * (k & 3))) & 0xffu), g, act);
    }
    NHFactor f;
    nh_factor(F, lae, mue, f);
    tet_hessian_blocks_t<float>(f, Bm, sc, [&](int a, int c, const float* blk) {
        const int k = uslot(a, c);
        float o[9];
        if (t[a] <= t[c]) {
#pragma unroll
            for (int q = 0; q < 9; q++) o[q] = blk[q];
        } else {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) o[i * 3 + j] = blk[j * 3 + i];
        }
        tile_add<9, float>(B, tcode[(size_t)k * nt + e], (int)((cw[k >> 2] >> (8 * (k & 3))) & 0xffu), o, act);
    });
    // nine (three) consecutive lanes write a record's consecutive entries
    for (int i = lane; i < 9 * nb; i += 64) bpart[(size_t)ppos[pbase + i / 9] * 9 + i % 9] = B[i];
    for (int i = lane; i < 3 * nvx; i += 64) gpart[(size_t)vpos[vbase + i / 3] * 3 + i % 3] = G[i];
}

// second pass: sum each stored (upper) block's contiguous run of oriented tet
// blocks in the scratch (tets in index order; in FP32 mode the chunk partials
// of the tiled pass, chunks in order), add the lumped mass on diagonal
// blocks (compute_inertia, NEWTON_IPC.py:558-567), store the FP64 values and,
// when valf != nullptr, their FP32 copy for the FP32 SpMV.  Nine lanes per
// block, one per entry, seven blocks per wavefront (lane 63 idles): the lanes
// of a block read the 72 contiguous bytes of each record together, so a
// load instruction touches a few cache lines instead of one per lane.  Each
// entry is summed in record order, as by a single thread.
constexpr int GATHER_BPW = 7;                       // blocks per wavefront
constexpr int GATHER_BPG = GATHER_BPW * (TPB / 64); // blocks per workgroup
static inline int gather_grid(long nnzb) { return (int)((nnzb + GATHER_BPG - 1) / GATHER_BPG); }
// FP64 mode takes the blocks in the order gorder (diagonal blocks first), so
// the blocks of a wavefront loop about equally long (gorder == nullptr: index
// order).
template <class R>
__global__ void __launch_bounds__(TPB) k_gather_blocks(int nnzb, const int* __restrict__ gorder,
                                                       const int* __restrict__ bcptr, const R* __restrict__ tetH,
                                                       const int* __restrict__ brow, const int* __restrict__ col,
                                                       const double* __restrict__ m, double* __restrict__ val,
                                                       float* __restrict__ valf)
{
    const int lane = threadIdx.x & 63;
    if (lane >= 9 * GATHER_BPW) return;
    const long slot = ((long)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6)) * GATHER_BPW + lane / 9;
    const int e = lane % 9;
    if (slot >= nnzb) return;
    const int b = gorder ? gorder[slot] : (int)slot;
    double acc = 0.0;
    const int t1 = bcptr[b + 1];
    // unrolled so that several records are in flight (same summation order)
#pragma unroll 4
    for (int t = bcptr[b]; t < t1; t++) acc += (double)tetH[(size_t)t * 9 + e];
    if ((e == 0 || e == 4 || e == 8) && col[b] == brow[b]) acc += m[brow[b]];
    if (val) val[(size_t)b * 9 + e] = acc;
    if (valf) valf[(size_t)b * 9 + e] = (float)acc;
}

// the FP32 tiled pass's gather: as k_gather_blocks in index order, but each
// group of nine lanes sums two blocks (wavefront w takes blocks 14 w .. +13,
// group g the blocks 14 w + g and 14 w + 7 + g) so that twice the records are
// in flight; each block's partials are still summed in chunk order
__global__ void __launch_bounds__(TPB) k_gather_parts(int nnzb, const int* __restrict__ pptr,
                                                      const float* __restrict__ part, const int* __restrict__ brow,
                                                      const int* __restrict__ col, const double* __restrict__ m,
                                                      float* __restrict__ valf)
{
    const int lane = threadIdx.x & 63;
    if (lane >= 9 * GATHER_BPW) return;
    const long w = (long)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    const long b0 = w * 2 * GATHER_BPW + lane / 9, b1 = b0 + GATHER_BPW;
    const int e = lane % 9;
    if (b0 >= nnzb) return;
    const bool two = b1 < nnzb;
    const int s0 = pptr[b0], t0 = pptr[b0 + 1];
    const int s1 = two ? pptr[b1] : 0, t1 = two ? pptr[b1 + 1] : 0;
    double a0 = 0.0, a1 = 0.0;
    const int n0 = t0 - s0, n1 = t1 - s1, n = n0 > n1 ? n0 : n1;
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        if (k < n0) a0 += (double)part[(size_t)(s0 + k) * 9 + e];
        if (k < n1) a1 += (double)part[(size_t)(s1 + k) * 9 + e];
    }
    if ((e == 0 || e == 4 || e == 8) && col[b0] == brow[b0]) a0 += m[brow[b0]];
    valf[(size_t)b0 * 9 + e] = (float)a0;
    if (two) {
        if ((e == 0 || e == 4 || e == 8) && col[b1] == brow[b1]) a1 += m[brow[b1]];
        valf[(size_t)b1 * 9 + e] = (float)a1;
    }
}

// the gradient in two parts around the contact assembly, which runs on a
// second stream meanwhile: k_rhs_init writes the inertia term
// (compute_inertia, NEWTON_IPC.py:558-567; rhs = -g, rhs[D] = 0,
// NEWTON_IPC.py:772) and clears the contact diagonal blocks before
// k_contact adds its rows; k_gather_tetg then subtracts the vertex's tet
// contributions in order (after the tiled pass: its chunk partials in chunk
// order).  Three lanes per vertex, one per component, 21
// vertices per wavefront (lane 63 idles): the lanes of a vertex read each
// 24-byte contribution together.
constexpr int RHS_VPW = 21;                     // vertices per wavefront
constexpr int RHS_VPG = RHS_VPW * (TPB / 64);   // vertices per workgroup
static inline int rhs_grid(long nv) { return (int)((nv + RHS_VPG - 1) / RHS_VPG); }
__global__ void __launch_bounds__(TPB) k_rhs_init(int nv, const double* __restrict__ m, const double* __restrict__ x,
                                                  const double* __restrict__ xt,
                                                  const unsigned char* __restrict__ dmask, double* rhs, double* cdiag)
{
    const int lane = threadIdx.x & 63;
    if (lane >= 3 * RHS_VPW) return;
    const long i = ((long)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6)) * RHS_VPW + lane / 3;
    const int d = lane % 3;
    if (i >= nv) return;
    for (int k = 0; k < 3; k++) cdiag[9 * i + 3 * d + k] = 0.0;
    rhs[3 * i + d] = dmask[i] ? 0.0 : -m[i] * (x[3 * i + d] - xt[3 * i + d]);
}
__global__ void __launch_bounds__(TPB) k_gather_tetg(int nv, const int* __restrict__ vcptr,
                                                     const int* __restrict__ vcent, const double* __restrict__ tetg,
                                                     const unsigned char* __restrict__ dmask, double* rhs)
{
    const int lane = threadIdx.x & 63;
    if (lane >= 3 * RHS_VPW) return;
    const long i = ((long)blockIdx.x * (TPB / 64) + (threadIdx.x >> 6)) * RHS_VPW + lane / 3;
    const int d = lane % 3;
    if (i >= nv || dmask[i]) return;
    double r = rhs[3 * i + d];
    const int t1 = vcptr[i + 1];
    if (vcent) {
#pragma unroll 4
        for (int t = vcptr[i]; t < t1; t++) r -= tetg[(size_t)vcent[t] * 3 + d];
    } else {  // the tiled pass's partials: a contiguous run per vertex
#pragma unroll 4
        for (int t = vcptr[i]; t < t1; t++) r -= tetg[(size_t)t * 3 + d];
    }
    rhs[3 * i + d] = r;
}

// new position of position k after a round of the round-robin ("circle")
// schedule on M (even) positions, which always pairs positions (i, M-1-i):
// position 0 stays, the others rotate by one, so the M - 1 rounds of a sweep
// pair every two indices exactly once
template <int M>
__host__ __device__ constexpr int rr_next(int k)
{
    return k == 0 ? 0 : (k == M - 1 ? 1 : k + 1);
}
// orders LDS accesses between the lanes of one wavefront
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// contact rows: a group of G lanes per row (G = 8 for PP, 16 otherwise).
// The projection onto the positive semi-definite cone (makePD semantics,
// UTILS.h:7-25: V max(L, 0) V^T of the eigen-decomposition) runs a parallel
// cyclic Jacobi on the matrix padded to M = N rounded up to even (a zero row
// and column, never rotated): lane l < M owns row l, a round applies the M / 2
// disjoint rotations of the pairs (i, M-1-i) at once, A <- J^T A J and
// V <- V J, then relabels the indices by the round-robin rotation rr_next, so
// every round is the same code with static register indices (the Brent-Luk
// arrangement); rows move between lanes, and the columns of J are exchanged,
// through LDS.  A pair rotates when |a_pq| > 1e-17 |A|_F; sweeps (M - 1
// rounds) stop once a whole sweep rotates no pair, after at most max_sweeps.
// LDS of one contact-assembly block: per lane group the M x M matrix and the
// two coefficient columns, sized for the largest row (M = 12, G = 16)
#define CONTACT_LDS ((TPB / 16) * (144 + 2 * 12))
template <int KIND>
__device__ __forceinline__ void contact_row(int n, const int* __restrict__ rows, const double* __restrict__ x,
                                            const double* __restrict__ x0, double dHat2, double kappa,
                                            double* __restrict__ Hout, double* cgrad, double* Eout,
                                            double* gout, int project, int accumulate, int max_sweeps, int gid,
                                            double* lds, float* __restrict__ Hf, unsigned char* __restrict__ c64)
{
    using KI = KindInfo<KIND>;
    constexpr int N = KI::N, G = KI::G;
    constexpr int M = (N + 1) & ~1;
    int r = gid / G, l = gid % G, grp = (int)threadIdx.x / G;
    if (r >= n) return;  // uniform over the lanes of a row
    const bool act = l < N;
    double* A = lds + grp * (M * M + 2 * M);
    double* jd = A + M * M;
    double* jo = jd + M;
    int id[4];
    v3 px[4], p0[4];
    load_row<KIND>(rows, r, x, x0, id, px, p0);
    double g[N], H[N * N];
    double E = constraint_gH<KIND>(px, p0, dHat2, kappa, g, H);
    // own row, from the upper triangle
    double o[N];
#pragma unroll
    for (int k = 0; k < N; k++) o[k] = 0.0;
#pragma unroll
    for (int L = 0; L < N; L++)
        if (L == l) {
#pragma unroll
            for (int k = 0; k < N; k++) o[k] = k >= L ? H[L * N + k] : H[k * N + L];
        }
    double out[N];
    if (project) {
        // derivative-only arithmetic: contraction into FMA is allowed here
#pragma clang fp contract(fast)
        const bool pos = l < M;
        double a[M], v[M];
#pragma unroll
        for (int k = 0; k < M; k++) {
            a[k] = k < N ? o[k] : 0.0;
            v[k] = k == l ? 1.0 : 0.0;
        }
        double f = 0.0;
#pragma unroll
        for (int k = 0; k < M; k++) f += a[k] * a[k];
#pragma unroll
        for (int mm = G / 2; mm > 0; mm >>= 1) f += __shfl_xor(f, mm, G);
        const double tiny = 1e-17 * sqrt(f);
        if (pos) {
#pragma unroll
            for (int k = 0; k < M; k++) A[l * M + k] = a[k];
        }
        wave_sync();
        const unsigned long long gmask = ((1ull << G) - 1) << (((int)threadIdx.x % 64) / G * G);
        const int m = M - 1 - l;
        const int dest = l == 0 ? 0 : (l == M - 1 ? 1 : l + 1);
        for (int sweep = 0; sweep < max_sweeps; sweep++) {
            bool rot = false;
#pragma unroll 1
            for (int rd = 0; rd < M - 1; rd++) {
                double c = 1.0, s = 0.0, t = 0.0, app = 0.0, aqq = 0.0, apq = 0.0;
                double w[M];
#pragma unroll
                for (int k = 0; k < M; k++) w[k] = 0.0;
                if (pos) {
                    int p = min(l, m), q = max(l, m);
                    app = A[p * M + p];
                    aqq = A[q * M + q];
                    apq = A[p * M + q];
#pragma unroll
                    for (int k = 0; k < M; k++) w[k] = A[m * M + k];
                }
                const bool did = pos && fabs(apq) > tiny;
                if (did) {
                    double theta = (aqq - app) / (2.0 * apq);
                    t = fabs(theta) > 1e150 ? 0.5 / theta
                      : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                }
                // column l of J: J[l][l] = c, J[m][l] = -s (l < m) or s (l > m)
                const double cl = c, sl = l < m ? -s : s;
                if (pos) { jd[l] = cl; jo[l] = sl; }
                wave_sync();
                double u[M];
#pragma unroll
                for (int k = 0; k < M; k++) u[k] = cl * a[k] + sl * w[k];
                double an[M], vn[M];
#pragma unroll
                for (int k = 0; k < M; k++) {
                    double ck = jd[k], sk = jo[k];
                    an[k] = ck * u[k] + sk * u[M - 1 - k];
                    vn[k] = ck * v[k] + sk * v[M - 1 - k];
                }
                if (did) {
                    // the rotated pair's entries, exactly
                    double dnew = l < m ? app - t * apq : aqq + t * apq;
#pragma unroll
                    for (int k = 0; k < M; k++) {
                        if (k == m) an[k] = 0.0;
                        if (k == l) an[k] = dnew;
                    }
                }
                rot = rot || did;
                // relabel the columns (static) and move the row to its new position
#pragma unroll
                for (int k = 0; k < M; k++) {
                    a[rr_next<M>(k)] = an[k];
                    v[rr_next<M>(k)] = vn[k];
                }
                if (pos) {
#pragma unroll
                    for (int k = 0; k < M; k++) A[dest * M + k] = a[k];
                }
                wave_sync();
                if (pos) {
#pragma unroll
                    for (int k = 0; k < M; k++) a[k] = A[l * M + k];
                }
            }
            if ((__ballot(rot) & gmask) == 0) break;
        }
        // eigenvalues clamped at zero; rows of V replace A in LDS
        double lam = pos ? fmax(A[l * M + l], 0.0) : 0.0;
        wave_sync();
        if (pos) {
#pragma unroll
            for (int k = 0; k < M; k++) A[l * M + k] = v[k];
            jd[l] = lam;
        }
        wave_sync();
        double lv[M];
#pragma unroll
        for (int k = 0; k < M; k++) lv[k] = v[k] * jd[k];
#pragma unroll
        for (int cc = 0; cc < N; cc++) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < M; k++) sum += lv[k] * A[cc * M + k];
            out[cc] = sum;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) out[k] = o[k];
    }
    if (!act) return;
    double* Hr = Hout + (size_t)r * N * N + (size_t)l * N;
#pragma unroll
    for (int k = 0; k < N; k++) Hr[k] = out[k];
    if (Hf) {
        // FP32 copy for the FP32 iterations; rows closer than 1e-2 dHat
        // (squared distance below 1e-4 dHat^2) are flagged to stay FP64
        float* Hfr = Hf + (size_t)r * N * N + (size_t)l * N;
#pragma unroll
        for (int k = 0; k < N; k++) Hfr[k] = (float)out[k];
        if (l == 0) c64[r] = constraint_dist(KIND, px) < 1e-4 * dHat2 ? 1 : 0;
    }
    double gl = g[0];
#pragma unroll
    for (int k = 1; k < N; k++)
        if (k == l) gl = g[k];
    if (accumulate) {
        // the row's gradient at its vertex l / 3 (slot K r + l / 3 of the
        // type), summed per vertex by k_cv_gather (load_hessian_and_gradient,
        // NEWTON_IPC.py:546-555)
        cgrad[3 * (size_t)r * KI::K + l] = gl;
    } else {
        if (l == 0 && Eout) Eout[r] = E;
        if (gout) gout[(size_t)r * N + l] = gl;
    }
}

// every constraint type in one launch: blocks [boff[t], boff[t + 1]) run the
// rows of type t, so the (latency-bound) types overlap instead of queueing
struct ContactBatch {
    const int* rows[NTYPES];
    double* H[NTYPES];
    float* Hf[NTYPES];            // null unless FP32 copies are wanted
    unsigned char* c64[NTYPES];
    double* cgrad[NTYPES];        // slot gradients of the type (accumulate mode)
    int n[NTYPES];
    int boff[NTYPES + 1];
};
__global__ void __launch_bounds__(TPB) k_contact(ContactBatch cb, const double* __restrict__ x,
                                                 const double* __restrict__ x0, double dHat2, double kappa,
                                                 double* Eout, double* gout, int project, int accumulate,
                                                 int max_sweeps)
{
    __shared__ double lds[CONTACT_LDS];
    int b = (int)blockIdx.x, t = 0;
    while (t < NTYPES - 1 && b >= cb.boff[t + 1]) t++;
    int gid = (b - cb.boff[t]) * TPB + (int)threadIdx.x;
#define CONTACT_CASE(K)                                                                                            \
    case K:                                                                                                         \
        contact_row<K>(cb.n[K], cb.rows[K], x, x0, dHat2, kappa, cb.H[K], cb.cgrad[K], Eout, gout, project,     \
                       accumulate, max_sweeps, gid, lds, cb.Hf[K], cb.c64[K]);                                      \
        break;
    switch (t) {
        CONTACT_CASE(0)
        CONTACT_CASE(1)
        CONTACT_CASE(2)
        CONTACT_CASE(3)
        CONTACT_CASE(4)
        CONTACT_CASE(5)
        CONTACT_CASE(6)
    }
#undef CONTACT_CASE
}
// block ranges of the types with rows (G lanes per row)
static int contact_blocks(ContactBatch& cb)
{
    int nb = 0;
    for (int t = 0; t < NTYPES; t++) {
        int G = kind_dofs(t) <= 8 ? 8 : 16;
        cb.boff[t] = nb;
        nb += cb.n[t] ? nblk((long)cb.n[t] * G) : 0;
    }
    cb.boff[NTYPES] = nb;
    return nb;
}

// ---- contact vertices: the slots of every row keyed by vertex ----
// the per-type tables a slot decodes through
struct ContactSlots {
    const int* rows[NTYPES];
    const double* H[NTYPES];
    const float* Hf[NTYPES];            // FP32 copies (FP32 iterations), or null
    const unsigned char* c64[NTYPES];   // rows that keep their FP64 Hessian
    int soff[NTYPES + 1];               // first slot of each type
};
__device__ __forceinline__ void slot_decode(const ContactSlots& cs, int s, int& t, int& r, int& a)
{
    t = 0;
    while (t < NTYPES - 1 && s >= cs.soff[t + 1]) t++;
    const int K = kind_points(t), o = s - cs.soff[t];
    r = o / K;
    a = o - r * K;
}
// key of a slot: its vertex, or nv for a Dirichlet vertex (sorted last, never summed)
__global__ void k_cv_keys(int S, ContactSlots cs, const unsigned char* __restrict__ dmask, int nv, int* keys,
                          int* vals)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    int t, r, a;
    slot_decode(cs, s, t, r, a);
    int v = cs.rows[t][kind_points(t) * r + a];
    keys[s] = dmask[v] ? nv : v;
    vals[s] = s;
}
// head[i] = 1 where a vertex's run starts (head[S] = 0 for the scan's total)
__global__ void k_cv_heads(int S, int nv, const int* __restrict__ keys, int* head)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > S) return;
    head[i] = i < S && keys[i] != nv && (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}
// pos = exclusive scan of head: run u starts at slot index i with pos[i] = u
__global__ void k_cv_compact(int S, int nv, const int* __restrict__ keys, const int* __restrict__ pos, int* vert,
                             int* ptr, int* count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const int k = keys[i];
    if (k == nv) return;
    const bool head = i == 0 || keys[i - 1] != k;
    if (head) { vert[pos[i]] = k; ptr[pos[i]] = i; }
    if (i == S - 1 || keys[i + 1] == nv) {  // the last run ends here
        const int n = pos[i] + (head ? 1 : 0);
        ptr[n] = i + 1;
        count[0] = n;
    }
}
// lanes per contact vertex: lane l of a vertex's group takes its slots
// l, l + CV_G, ... in order, and the group combines the lanes' sums with a
// fixed butterfly, so a vertex with many slots costs a few dependent loads,
// not one chain per slot, and the result is the same bits on every run
#define CV_G 8
// gradient and diagonal block of every contact vertex, summed over its slots
// (CV_G lanes per vertex): rhs -= g (after the inertia term), the diagonal
// block written whole (k_rhs_init cleared every vertex's)
__global__ void __launch_bounds__(TPB) k_cv_gather(const int* __restrict__ count, const int* __restrict__ vert,
                                                   const int* __restrict__ ptr, const int* __restrict__ slot,
                                                   ContactSlots cs, const double* __restrict__ cgrad, double* rhs,
                                                   double* cdiag)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int u = gid / CV_G, l = gid % CV_G;
    const bool live = u < count[0];
    double acc[12];  // g (3), then the diagonal block row-major (9)
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = 0.0;
    if (live) {
        for (int e = ptr[u] + l; e < ptr[u + 1]; e += CV_G) {
            int sl = slot[e], t, r, a;
            slot_decode(cs, sl, t, r, a);
            const int N = kind_dofs(t);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                acc[i] += cgrad[3 * (size_t)sl + i];
                const double* h = cs.H[t] + (size_t)r * N * N + (size_t)(3 * a + i) * N + 3 * a;
                acc[3 + 3 * i] += h[0]; acc[4 + 3 * i] += h[1]; acc[5 + 3 * i] += h[2];
            }
        }
    }
    // every lane of the wavefront takes part (groups never straddle one)
#pragma unroll
    for (int o = CV_G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 12; k++) acc[k] += __shfl_xor(acc[k], o, 64);
    if (!live || l >= 3) return;
    const int v = vert[u], i = l;
    double g = i == 0 ? acc[0] : (i == 1 ? acc[1] : acc[2]);
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = i == 0 ? acc[3 + k] : (i == 1 ? acc[6 + k] : acc[9 + k]);
    rhs[3 * v + i] -= g;
    cdiag[9 * v + 3 * i] = d[0]; cdiag[9 * v + 3 * i + 1] = d[1]; cdiag[9 * v + 3 * i + 2] = d[2];
}

// block-Jacobi preconditioner: inverse of (elastic diag block + contact diag
// block); Dirichlet vertices get the identity (their residual is zero)
template <class T>
__global__ void k_precond(int nv, const T* __restrict__ val, const int* __restrict__ diag,
                          const double* __restrict__ cdiag, const unsigned char* __restrict__ dmask, double* Minv,
                          float* Minvf)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    double A[9];
    int b = diag[i];
    for (int k = 0; k < 9; k++) A[k] = (double)val[(size_t)b * 9 + k] + cdiag[9 * i + k];
    double* o = Minv + 9 * i;
    double d = det3(A);
    double M[9];
    if (dmask[i] || !(fabs(d) > 0.0)) {
        for (int k = 0; k < 9; k++) M[k] = (k % 4 == 0) ? (dmask[i] ? 1.0 : 1.0 / fmax(fabs(A[k]), 1e-300)) : 0.0;
    } else {
        double id = 1.0 / d;
        M[0] = (A[4] * A[8] - A[5] * A[7]) * id; M[1] = (A[2] * A[7] - A[1] * A[8]) * id;
        M[2] = (A[1] * A[5] - A[2] * A[4]) * id; M[3] = (A[5] * A[6] - A[3] * A[8]) * id;
        M[4] = (A[0] * A[8] - A[2] * A[6]) * id; M[5] = (A[2] * A[3] - A[0] * A[5]) * id;
        M[6] = (A[3] * A[7] - A[4] * A[6]) * id; M[7] = (A[1] * A[6] - A[0] * A[7]) * id;
        M[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    }
    for (int k = 0; k < 9; k++) o[k] = M[k];
    if (Minvf)
        for (int k = 0; k < 9; k++) Minvf[9 * i + k] = (float)M[k];
}

// ---------------------------------------------------------------------------
// SpMV: q += A p  (BSR part) and q += sum_rows H_r p_r (contact part)
// ---------------------------------------------------------------------------
// lanes per block row: with 4, a wavefront covers 16 consecutive rows, whose
// gathers share cache lines of p and of the value stream (faster than 2, 8 or
// 16 lanes on the 1M-tet slab)
#ifndef SPMV_G
#define SPMV_G 4
#endif
// One launch computes A p: blocks [0, nb_c) run the contact rows of every
// type (first, so that they overlap the BSR rows instead of trailing them),
// the nb_bsr blocks after them the BSR rows (nb_c is a multiple of 8, so the
// XCD of a BSR block is that of its own index).
//  - BSR rows: SPMV_G lanes per block row stride over the row's entries
//    (stored blocks, then transposed lower blocks) and reduce with shuffles;
//    a block handles TPB / SPMV_G consecutive rows per pass.  q[row] =
//    (A_bsr p)[row] is stored (zero on Dirichlet rows), so q needs no
//    clearing.
//  - contact part: one lane per (contact vertex, component) sums the rows of
//    H_r p_r at that vertex over its slots in their fixed order and stores
//    the sum into qc (Dirichlet vertices have no slots).
// With part != 0 every block stores its partial of p.(A p): p_row . q[row] for
// BSR rows, p_v . (H_r p_r)_v for contact lanes (Dirichlet entries of the
// PCG directions are zero, so this is the masked p.q).
#define SPMV_ROWS (TPB / SPMV_G)
// entries of a row per lane whose loads are in flight together
#ifndef SPMV_UNROLL
#define SPMV_UNROLL _Pragma("unroll 2")
#endif
#ifndef SPMV_MAXB
#define SPMV_MAXB 1024
#endif
static inline int spmv_grid(int nv)
{
    int groups = (nv + SPMV_ROWS - 1) / SPMV_ROWS;
    int nb = std::min(groups, SPMV_MAXB);
    return std::max(8, (nb + 7) & ~7);
}
// the contact part of a product: the contact vertices with their slot lists
struct ContactSet {
    ContactSlots sl;
    const int* count;   // number of contact vertices (device)
    const int* vert;
    const int* ptr;
    const int* slot;
};
// lane l's share of (sum_r H_r p_r) at contact vertex u: the vertex's slots
// l, l + CV_G, ..., each the product of the slot's three Hessian rows 3 a + i
// with the row's points, added into s[3].  With Hf set (FP32 iterations) a
// row uses its FP32 copy unless flagged in c64.
template <class V>
__device__ __forceinline__ void contact_vertex_spmv(const ContactSet& cs, const V* __restrict__ p, int u, int l,
                                                    double* s)
{
    for (int e = cs.ptr[u] + l; e < cs.ptr[u + 1]; e += CV_G) {
        int t, r, a;
        slot_decode(cs.sl, cs.slot[e], t, r, a);
        const int K = kind_points(t), N = 3 * K;
        const int* rw = cs.sl.rows[t] + K * r;
        const size_t off = (size_t)r * N * N + (size_t)(3 * a) * N;
        // point by point (static indices, guarded up to the widest row), so
        // only one point's p and the three partial rows are live
        const bool f32 = cs.sl.Hf[t] && !cs.sl.c64[t][r];
        const float* hf = f32 ? cs.sl.Hf[t] + off : nullptr;
        const double* hd = cs.sl.H[t] + off;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            if (b < K) {
                const int w = rw[b];
                const double p0 = (double)p[3 * w], p1 = (double)p[3 * w + 1], p2 = (double)p[3 * w + 2];
                if (f32) {
                    const float* h = hf + 3 * b;
                    a0 += (double)h[0] * p0 + (double)h[1] * p1 + (double)h[2] * p2;
                    a1 += (double)h[N] * p0 + (double)h[N + 1] * p1 + (double)h[N + 2] * p2;
                    a2 += (double)h[2 * N] * p0 + (double)h[2 * N + 1] * p1 + (double)h[2 * N + 2] * p2;
                } else {
                    const double* h = hd + 3 * b;
                    a0 += h[0] * p0 + h[1] * p1 + h[2] * p2;
                    a1 += h[N] * p0 + h[N + 1] * p1 + h[N + 2] * p2;
                    a2 += h[2 * N] * p0 + h[2 * N + 1] * p1 + h[2 * N + 2] * p2;
                }
            }
        }
        s[0] += a0; s[1] += a1; s[2] += a2;
    }
}
template <class T, class V>
__global__ void __launch_bounds__(TPB) k_spmv(int nv, const int* __restrict__ rowptr, const int* __restrict__ col,
                                              const int* __restrict__ lptr, const int2* __restrict__ lent,
                                              const T* __restrict__ val, const V* __restrict__ p,
                                              V* __restrict__ q, double* qc,
                                              const unsigned char* __restrict__ dmask, ContactSet cs, int nb_c,
                                              int nb_bsr, double* part, const int* __restrict__ state, int it)
{
    if (state && it > state[0]) return;
    double pq = 0.0;
    if ((int)blockIdx.x >= nb_c) {
        // the row groups are split into 8 contiguous ranges, one per XCD
        // (blocks are dispatched to XCDs round-robin); the XCD's blocks
        // interleave over its range, so at any time they cover consecutive
        // rows and the transposed blocks a row needs (stored up to a few
        // thousand rows above it) were read into the XCD's L2 by the sweep
        // just before
        int bb = (int)blockIdx.x - nb_c;
        int nbx = nb_bsr >> 3;
        int groups = (nv + SPMV_ROWS - 1) / SPMV_ROWS;
        int gpx = (groups + 7) >> 3;
        int g1 = min((bb & 7) * gpx + gpx, groups);
        int g0 = (bb & 7) * gpx + (bb >> 3);
        int lane = threadIdx.x % SPMV_G;
        int rsub = threadIdx.x / SPMV_G;
        // row bounds of the next group are loaded while the current one runs
        int row = g0 * SPMV_ROWS + rsub;
        bool live = row < nv && g0 < g1;
        int nu0 = live ? rowptr[row] : 0, nu1 = live ? rowptr[row + 1] : 0;
        int nl0 = live ? lptr[row] : 0, nl1 = live ? lptr[row + 1] : 0;
        for (int g = g0; g < g1; g += nbx) {
            row = g * SPMV_ROWS + rsub;
            int u0 = nu0, l0 = nl0;
            int nu = nu1 - nu0, n = nu + (nl1 - nl0);
            int nrow = row + nbx * SPMV_ROWS;
            if (g + nbx < g1 && nrow < nv) {
                nu0 = rowptr[nrow]; nu1 = rowptr[nrow + 1];
                nl0 = lptr[nrow]; nl1 = lptr[nrow + 1];
            }
            // the row's stored blocks (i, j), j >= i, then the transposes of
            // the blocks (j, i), j < i, stored in the rows above
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            SPMV_UNROLL
            for (int k = lane; k < n; k += SPMV_G) {
                bool tr = k >= nu;
                int b, c;
                if (!tr) {
                    b = u0 + k;
                    c = col[b];
                } else {
                    int2 e = lent[l0 + k - nu];
                    b = e.x;
                    c = e.y;
                }
                const T* A = val + (size_t)b * 9;
                double v0 = A[0], v1 = A[1], v2 = A[2], v3 = A[3], v4 = A[4], v5 = A[5], v6 = A[6], v7 = A[7],
                       v8 = A[8];
                if (tr) {
                    double t;
                    t = v1; v1 = v3; v3 = t;
                    t = v2; v2 = v6; v6 = t;
                    t = v5; v5 = v7; v7 = t;
                }
                double x0 = (double)p[3 * c], x1 = (double)p[3 * c + 1], x2 = (double)p[3 * c + 2];
                a0 += v0 * x0 + v1 * x1 + v2 * x2;
                a1 += v3 * x0 + v4 * x1 + v5 * x2;
                a2 += v6 * x0 + v7 * x1 + v8 * x2;
            }
#pragma unroll
            for (int o = SPMV_G / 2; o > 0; o >>= 1) {
                a0 += __shfl_xor(a0, o, 64);
                a1 += __shfl_xor(a1, o, 64);
                a2 += __shfl_xor(a2, o, 64);
            }
            if (row < nv && lane == 0) {
                if (dmask[row]) a0 = a1 = a2 = 0.0;
                q[3 * row] = (V)a0; q[3 * row + 1] = (V)a1; q[3 * row + 2] = (V)a2;
                pq += (double)p[3 * row] * a0 + (double)p[3 * row + 1] * a1 + (double)p[3 * row + 2] * a2;
            }
        }
    } else {
        // contact part: CV_G lanes per contact vertex, combined by a fixed
        // butterfly; the sum is stored (each entry of qc has one writer)
        const int gid = (int)blockIdx.x * TPB + (int)threadIdx.x;
        const int u = gid / CV_G, l = gid % CV_G;
        const bool live = u < cs.count[0];
        double sv[3] = {0.0, 0.0, 0.0};
        if (live) contact_vertex_spmv<V>(cs, p, u, l, sv);
#pragma unroll
        for (int o = CV_G / 2; o > 0; o >>= 1) {
            sv[0] += __shfl_xor(sv[0], o, 64);
            sv[1] += __shfl_xor(sv[1], o, 64);
            sv[2] += __shfl_xor(sv[2], o, 64);
        }
        if (live && l < 3) {
            const int v = cs.vert[u];
            const double si = l == 0 ? sv[0] : (l == 1 ? sv[1] : sv[2]);
            qc[3 * v + l] = si;
            pq = (double)p[3 * v + l] * si;
        }
    }
    if (part) {
        pq = block_sum(pq);
        if (threadIdx.x == 0) part[blockIdx.x] = pq;
    }
}

// ---------------------------------------------------------------------------
// PCG with block-Jacobi preconditioner (replaces the CHOLMOD solve of
// NEWTON_IPC.py:773-775).  Dirichlet dofs carry a prescribed increment in the
// initial guess and a zero residual afterwards.
// ---------------------------------------------------------------------------
// x = p0 (prescribed on Dirichlet dofs, 0 elsewhere)
// x0 = prescribed Dirichlet increments (0 elsewhere); *nonzero (set to -1
// beforehand) becomes 0 if any increment is non-zero, so that the product
// A x0 and its use can be skipped when x0 = 0 (then b - A x0 = b exactly)
__global__ void k_pcg_x0(int nv, const double* xt, const double* xcur, const unsigned char* dmask, double* sol,
                         double* qc, int* nonzero)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool nz = false;
    if (i < nv) {
        for (int d = 0; d < 3; d++) {
            double v = dmask[i] ? xt[3 * i + d] - xcur[3 * i + d] : 0.0;
            sol[3 * i + d] = v;
            qc[3 * i + d] = 0.0;
            nz |= v != 0.0;
        }
    }
    if (__ballot(nz) != 0 && (threadIdx.x & 63) == 0) nonzero[0] = 0;
}
// r = b - (q + qc) (= b - A x, from FP64 products); z = Minv r; p = z; qc
// cleared for the iterations; partials of r.z and r.r.  V is the type the
// iterations keep r, z, p in (float for the FP32 Hessian's refinement rounds).
// x0zero (may be null): -1 when x = 0, whose product was skipped (q, qc
// not written: r = b)
template <class V>
__global__ void k_pcg_init(int nv, const double* b, const double* q, double* qc, const unsigned char* dmask,
                           const double* Minv, V* r, V* z, V* p, double* part_rz, double* part_bb,
                           const int* x0zero)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0.0, bb = 0.0;
    const bool skip = x0zero && x0zero[0] == -1;
    if (i < nv) {
        double ri[3], zi[3];
        bool fixed = dmask[i];
        for (int d = 0; d < 3; d++)
            ri[d] = fixed ? 0.0 : (skip ? b[3 * i + d] : b[3 * i + d] - (q[3 * i + d] + qc[3 * i + d]));
        const double* M = Minv + 9 * i;
        for (int d = 0; d < 3; d++) zi[d] = fixed ? 0.0 : M[3 * d] * ri[0] + M[3 * d + 1] * ri[1] + M[3 * d + 2] * ri[2];
        for (int d = 0; d < 3; d++) {
            V rv = (V)ri[d], zv = (V)zi[d];
            r[3 * i + d] = rv; z[3 * i + d] = zv; p[3 * i + d] = zv; qc[3 * i + d] = 0.0;
            rz += (double)rv * (double)zv;
            bb += ri[d] * ri[d];
        }
    }
    rz = block_sum(rz);
    bb = block_sum(bb);
    if (threadIdx.x == 0) { part_rz[blockIdx.x] = rz; part_bb[blockIdx.x] = bb; }
}
// scal[0|1] = rz from partials.  First start: scal[2|3] = |r0|^2, the
// reference norm of the relative tolerance.  Restart (refinement): the
// partials hold |r|^2 of the recomputed FP64 residual, kept in scal[4];
// state[0] = -1 when that residual already meets the tolerance.
__global__ void k_pcg_scalars(const double* part_rz, const double* part_bb, int n, double* scal, int* state,
                              int restart, double tol2)
{
    double rz = reduce_partials(part_rz, n);
    double bb = reduce_partials(part_bb, n);
    if (threadIdx.x == 0) {
        scal[0] = rz; scal[1] = rz;
        if (!restart) {
            scal[2] = bb; scal[3] = bb; scal[4] = bb;
            state[0] = bb == 0.0 ? -1 : 0x7fffffff;  // done_at
        } else {
            scal[4] = bb;
            state[0] = (bb <= tol2 * scal[2] || !(bb == bb)) ? -1 : 0x7fffffff;
        }
        state[1] = 0;
    }
}
// the convergence test of a kept restart (scal[4] = |r|^2 of its FP64
// residual) against a new tolerance
__global__ void k_pcg_retest(const double* scal, int* state, double tol2)
{
    if (threadIdx.x == 0) {
        double bb = scal[4];
        state[0] = (bb <= tol2 * scal[2] || !(bb == bb)) ? -1 : 0x7fffffff;
        state[1] = 0;
    }
}
// scal[0] = r.z of a round's start, scal[2] = |r0|^2, scal[4] = |r|^2 of
// the last tested residual, scal[5] = u.A u of the last iteration,
// scal[6|7] and scal[8|9] = gamma and alpha of an iteration (by parity);
// state[0] = iteration at which the test passed (done_at), state[1] =
// iterations run.  The update kernel runs a grid-stride loop over vertices on
// at most PCG_MAXB blocks, so the partials that the next kernel reduces stay
// few.
// the FP32 iterations stop at this fraction of the tolerance, so that the
// FP64 residual that follows usually meets it
#ifndef PCG_FP32_INNER
#define PCG_FP32_INNER 1.0
#endif
// iterations per host convergence check after the first chunk, and how many
// iterations short of the hint the first chunk stops
#ifndef PCG_CHUNK
#define PCG_CHUNK 8
#endif
#ifndef PCG_FIRST_SHORT
#define PCG_FIRST_SHORT 4
#endif
#ifndef PCG_MAXB
#define PCG_MAXB 512
#endif
static inline int pcg_grid(int nv) { return std::min(nblk(nv), PCG_MAXB); }
// One iteration of PCG in the single-reduction form of Chronopoulos and Gear:
// the product of this iteration is w = A u (u = Minv r; w = q + qc) with the
// partials of u.w (delta); the partials of gamma = r.u and of r.r were
// written by the previous iteration into part_g[(it + 1) & 1] (it = 0: gamma
// is scal[0] from the round's start, whose residual its own test covered).
//   beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev)
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = Minv r
// and the partials of the new r.u and r.r go to part_g[it & 1].  x is the
// solution (FP64 rounds) or, in the FP32 rounds, the round's correction in
// FP32, which k_pcg_flush adds to the FP64 solution when the round ends.  The test of
// an iteration's residual runs at the start of the next one (every block
// decides it from the same partials; block 0 records it): on success
// done_at = it - 1 and nothing more is written.  gamma and alpha of an
// iteration are kept by parity in scal[6 + (it & 1)] and scal[8 + (it & 1)],
// so no block reads a value another block of the same launch writes.  One
// launch replaces the two update kernels of the textbook form (whose second
// waits for the r.z reduction of the first).
template <class V, class M>
__global__ void __launch_bounds__(TPB) k_pcg_cg(int nv, double* scal, const unsigned char* __restrict__ dmask,
                                                const M* __restrict__ Minv, V* __restrict__ u,
                                                const V* __restrict__ w, double* __restrict__ qc,
                                                V* __restrict__ p, V* __restrict__ s, V* __restrict__ x,
                                                V* __restrict__ r, const double* part_uw, int nuw, double* part_g,
                                                int npg, int* state, int it, double tol2)
{
    if (it > state[0]) return;
    const int cur = it & 1, prv = cur ^ 1;
    double gamma, rr;
    if (it == 0) {
        gamma = scal[0];
        rr = scal[4];
    } else {
        reduce_partials2(part_g + 2 * PCG_MAXB * prv, part_g + 2 * PCG_MAXB * prv + PCG_MAXB, npg, gamma, rr);
        if (rr <= tol2 * scal[2] || !(rr == rr)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                state[0] = it - 1;
                state[1] = it;
                scal[4] = rr;
            }
            return;
        }
    }
    const double delta = reduce_partials(part_uw, nuw);
    const double beta = it == 0 ? 0.0 : gamma / scal[6 + prv];
    const double alpha = it == 0 ? gamma / delta : gamma / (delta - beta * gamma / scal[8 + prv]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[6 + cur] = gamma;
        scal[8 + cur] = alpha;
        scal[4] = rr;
        scal[5] = delta;
        state[1] = it + 1;
    }
    double ru = 0.0, rn2 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
        if (dmask[i]) continue;
        // every load of the vertex is issued before the first store
        const int k = 3 * i;
        double uv[3], wv[3], qcv[3], pv[3], sv[3], xv[3], rv[3], Mi[9];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            uv[d] = (double)u[k + d];
            wv[d] = (double)w[k + d];
            qcv[d] = qc[k + d];
            pv[d] = it == 0 ? 0.0 : (double)p[k + d];
            sv[d] = it == 0 ? 0.0 : (double)s[k + d];
            xv[d] = (double)x[k + d];
            rv[d] = (double)r[k + d];
        }
#pragma unroll
        for (int j = 0; j < 9; j++) Mi[j] = (double)Minv[9 * i + j];
        double ri[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            // qc is non-zero only at contact vertices: cleared here, where
            // it is consumed, for the next product
            if (qcv[d] != 0.0) qc[k + d] = 0.0;
            const V pn = (V)(uv[d] + beta * pv[d]);
            const V sn = (V)((wv[d] + qcv[d]) + beta * sv[d]);
            p[k + d] = pn;
            s[k + d] = sn;
            x[k + d] = (V)(xv[d] + alpha * (double)pn);
            const V rnv = (V)(rv[d] - alpha * (double)sn);
            r[k + d] = rnv;
            ri[d] = (double)rnv;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const V un = (V)(Mi[3 * d] * ri[0] + Mi[3 * d + 1] * ri[1] + Mi[3 * d + 2] * ri[2]);
            u[k + d] = un;
            ru += ri[d] * (double)un;
            rn2 += ri[d] * ri[d];
        }
    }
    ru = block_sum(ru);
    rn2 = block_sum(rn2);
    if (threadIdx.x == 0) {
        part_g[2 * PCG_MAXB * cur + blockIdx.x] = ru;
        part_g[2 * PCG_MAXB * cur + PCG_MAXB + blockIdx.x] = rn2;
    }
}
// the test of the last iteration's residual when the iteration budget ended
// before the next iteration could run it
__global__ void k_pcg_cg_last(double* scal, const double* part_g, int npg, int* state, int it, double tol2)
{
    if (it > state[0]) return;
    const int prv = (it & 1) ^ 1;
    double gamma, rr;
    reduce_partials2(part_g + 2 * PCG_MAXB * prv, part_g + 2 * PCG_MAXB * prv + PCG_MAXB, npg, gamma, rr);
    if (threadIdx.x == 0) {
        scal[4] = rr;
        if (rr <= tol2 * scal[2] || !(rr == rr)) state[0] = it - 1;
    }
}

// end of an FP32 round: the solution takes the round's correction, which is
// cleared for the next round
__global__ void k_pcg_flush(int n, float* __restrict__ dx, double* __restrict__ x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dx[i];
    if (d != 0.0f) {
        x[i] += (double)d;
        dx[i] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// energy (NEWTON_IPC.py:514-542): per-block partial sums, summed in order
// ---------------------------------------------------------------------------
__global__ void k_energy_inertia(int nv, const double* m, const double* x, const double* xt, double* part)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < nv) {
        v3 d = ldv(x, i) - ldv(xt, i);
        e = 0.5 * m[i] * n2(d);
    }
    e = block_sum(e);
    if (threadIdx.x == 0) part[blockIdx.x] = e;
}
__global__ void k_energy_elastic(int nt, const int* tets, const double* IB, const double* vol, const double* la,
                                 const double* mu, const double* x, double dt2, double* part)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    double en = 0.0;
    if (e < nt) {
        int t[4];
        double F[9], Bm[4][3];
        tet_load(tets, IB, x, e, t, F, Bm);
        en = nh_energy_F(F, la[e], mu[e]) * dt2 * vol[e];
    }
    en = block_sum(en);
    if (threadIdx.x == 0) part[blockIdx.x] = en;
}
template <int KIND>
__global__ void k_energy_contact(int n, const int* rows, const double* x, const double* x0, double dHat2, double kappa,
                                 double* part)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    double en = 0.0;
    if (r < n) {
        int id[4];
        v3 px[4], p0[4];
        load_row<KIND>(rows, r, x, x0, id, px, p0);
        en = constraint_E(KIND, px, p0, dHat2, kappa);
    }
    en = block_sum(en);
    if (threadIdx.x == 0) part[blockIdx.x] = en;
}
template <int KIND>
__global__ void k_min_dist(int n, const int* rows, const double* x, unsigned long long* out)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int id[4];
    v3 px[4], p0[4];
    load_row<KIND>(rows, r, x, x, id, px, p0);
    double d = constraint_dist(KIND, px);
    atomicMin(out, (unsigned long long)__double_as_longlong(d));
}

// inversion filter (NEWTON_IPC.py:443-453): the minimum over the tets, folded
// into *amin (a non-negative double's bit pattern, 1 on entry) with one
// atomic per wavefront whose minimum is below 1 (same-address atomics
// serialise, so the waves that cannot lower the minimum issue none)
__global__ void k_inversion(int nt, const int* tets, const double* x, const double* p, unsigned long long* amin)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    double al = 1.0;
    if (e < nt) {
        int a = tets[4 * e], b = tets[4 * e + 1], c = tets[4 * e + 2], d = tets[4 * e + 3];
        v3 x1 = ldv(x, a), x2 = ldv(x, b), x3 = ldv(x, c), x4 = ldv(x, d);
        v3 p1 = ldv(p, a), p2 = ldv(p, b), p3 = ldv(p, c), p4 = ldv(p, d);
        double ca, cb, cc, cd;
        tet_det_cubic(x1, x2, x3, x4, p1, p2, p3, p4, ca, cb, cc, cd);
        // the root finder only for the tets whose bound may be below 1
        if (!tet_inversion_clear(ca, cb, cc, cd, 0.2)) al = tet_inversion_alpha(x1, x2, x3, x4, p1, p2, p3, p4, 0.2);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) al = fmin(al, __shfl_xor(al, o, 64));
    if ((threadIdx.x & 63) == 0 && al < 1.0) atomicMin(amin, (unsigned long long)__double_as_longlong(al));
}

// batched tet evaluation for tests
__global__ void k_batch_tet(int n, const double* Fin, const double* IB, const double* scale, const double* la,
                            const double* mu, double* psi, double* P, double* Hout)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double F[9], ib[9];
    for (int k = 0; k < 9; k++) { F[k] = Fin[9 * e + k]; ib[k] = IB[9 * e + k]; }
    psi[e] = nh_energy_F(F, la[e], mu[e]);
    double Pl[9];
    nh_P(F, la[e], mu[e], Pl);
    for (int k = 0; k < 9; k++) P[9 * e + k] = Pl[k];
    if (!Hout) return;
    double Bm[4][3];
    for (int j = 0; j < 3; j++) {
        Bm[1][j] = ib[j]; Bm[2][j] = ib[3 + j]; Bm[3][j] = ib[6 + j];
        Bm[0][j] = -ib[j] - ib[3 + j] - ib[6 + j];
    }
    NHFactor f;
    nh_factor(F, la[e], mu[e], f);
    double* H = Hout + 144 * (size_t)e;
    tet_hessian_blocks(f, Bm, scale[e], [&](int a, int c, const double* blk) {
        for (int i = 0; i < 3; i++)
            for (int l = 0; l < 3; l++) {
                H[(3 * a + i) * 12 + 3 * c + l] = blk[i * 3 + l];
                H[(3 * c + l) * 12 + 3 * a + i] = blk[i * 3 + l];
            }
    });
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// timing events: no system-scope fence (they only time the kernels between
// them; host reads of results go through stream synchronisation), which
// keeps their own cost out of the bracketed launch
static hipEvent_t pool_get(fipc_ctx* c)
{
    if (c->event_pool.empty()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
}
// the events of a timed class are recorded on the stream its launches use
// (c->s unless given); fipc_tm_fold waits for both streams
int fipc_tm_begin(fipc_ctx* c, int which, hipStream_t st)
{
    if (!c->timing) return FIPC_OK;
    hipEvent_t a = pool_get(c), b = pool_get(c);
    if (!a || !b) return FIPC_ERR_HIP;
    HIPCHK(hipEventRecord(a, st ? st : c->s));
    c->tm[which].pending.push_back({a, b});
    return FIPC_OK;
}
int fipc_tm_end(fipc_ctx* c, int which, hipStream_t st)
{
    if (!c->timing) return FIPC_OK;
    HIPCHK(hipEventRecord(c->tm[which].pending.back().second, st ? st : c->s));
    if (c->tm[which].pending.size() > 4096) STCHK(fipc_tm_fold(c));
    return FIPC_OK;
}
int fipc_tm_fold(fipc_ctx* c)
{
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->s2) HIPCHK(hipStreamSynchronize(c->s2));  // its timed launches too
    for (auto& t : c->tm) {
        for (auto& pr : t.pending) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
            t.total_ms += ms;
            t.launches++;
            c->event_pool.push_back(pr.first);
            c->event_pool.push_back(pr.second);
        }
        t.pending.clear();
    }
    return FIPC_OK;
}

template <class T>
static int upload(DVec<T>& d, const T* h, size_t n, hipStream_t s)
{
    STCHK(ensure(d, n));
    if (n) HIPCHK(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return FIPC_OK;
}

extern "C" int fipc_create(fipc_ctx** out)
{
    if (!out) return FIPC_ERR_ARG;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return FIPC_ERR_HIP;
    fipc_ctx* c = new fipc_ctx();
    hipError_t e = hipStreamCreate(&c->s);
    if (e == hipSuccess) e = hipStreamCreate(&c->s2);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_rhs, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_contact, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->pin, FIPC_PIN_BYTES, hipHostMallocDefault);
    if (e != hipSuccess) {  // release what was created
        (void)fipc_destroy(c);
        HIPCHK(e);
    }
    *out = c;
    return FIPC_OK;
}

extern "C" int fipc_destroy(fipc_ctx* c)
{
    if (!c) return FIPC_OK;
    if (c->s) (void)hipStreamSynchronize(c->s);
    if (c->s2) (void)hipStreamSynchronize(c->s2);
    // pending timing events are read while both streams still exist
    if (c->s) (void)fipc_tm_fold(c);
    for (auto& t : c->tm) {  // events left by a failed fold go back to the pool
        for (auto& pr : t.pending) { c->event_pool.push_back(pr.first); c->event_pool.push_back(pr.second); }
        t.pending.clear();
    }
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->ev_rhs) (void)hipEventDestroy(c->ev_rhs);
    if (c->ev_contact) (void)hipEventDestroy(c->ev_contact);
    if (c->s2) (void)hipStreamDestroy(c->s2);
    c->s2 = nullptr;
    if (c->ws_done) (void)hipEventDestroy(c->ws_done);
    release(c->tets); release(c->bp); release(c->be); release(c->bt); release(c->dids);
    release(c->IB); release(c->vol); release(c->la); release(c->mu); release(c->m); release(c->dmask);
    release(c->x); release(c->x0); release(c->xn); release(c->xt); release(c->xprev); release(c->v);
    release(c->sol); release(c->rhs); release(c->target);
    release(c->rowptr); release(c->col); release(c->diag); release(c->val); release(c->valf);
    release(c->brow); release(c->tetH); release(c->tetg); release(c->bcptr); release(c->gorder); release(c->cpos); release(c->vcptr);
    release(c->lptr); release(c->lent);
    release(c->vcent);
    release(c->tile_meta); release(c->tile_tcode); release(c->tile_vcode); release(c->tile_ppos); release(c->tile_vpos);
    release(c->tile_pptr); release(c->tile_vptr); release(c->tile_bpart); release(c->tile_gpart);
    for (int t = 0; t < NTYPES; t++) { release(c->rows[t]); release(c->cH[t]); release(c->cHf[t]); release(c->c64[t]); }
    release(c->cv_keys); release(c->cv_vals); release(c->cv_keys2); release(c->cv_slot); release(c->cv_head);
    release(c->cv_vert); release(c->cv_ptr); release(c->cv_count); release(c->cgrad);
    release(c->keys_pp); release(c->keys_pe); release(c->keys_tmp); release(c->counters); release(c->cub_tmp); release(c->cub_tmp2); release(c->rows_tmp);
    release(c->nuniq); release(c->cdiag); release(c->Minv); release(c->r); release(c->z); release(c->pp); release(c->q); release(c->qc);
    release(c->Minvf); release(c->rf); release(c->zf); release(c->pf); release(c->qf);
    release(c->cgs); release(c->cgsf); release(c->part_g); release(c->dxf);
    release(c->trow); release(c->tval);
    release(c->part_a); release(c->part_b); release(c->part_c); release(c->scal); release(c->pcg_state);
    release(c->x0flag);
    release(c->epart); release(c->amin); release(c->flag);
    for (Grid& g : c->grid) {
        release(g.cnt); release(g.ekey); release(g.eitem); release(g.bstart);
        release(g.items); release(g.ecell); release(g.cells); release(g.qoff); release(g.part); release(g.qpart);
    }
    release(c->ccd_pairs[0]); release(c->ccd_pairs[1]);
    release(c->dprev); release(c->wsX); release(c->ckey); release(c->free_ids); release(c->cap_ids);
    release(c->cap_X); release(c->cap_T); release(c->near_ids); release(c->nnear);
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return FIPC_OK;
}

// symmetric BSR pattern of the elastic + inertia Hessian: one upper block per
// vertex pair sharing a tet (and every diagonal block), the lower lists of the
// SpMV, and the block-sorted positions of the per-tet contributions
// chunks of the tiled FP32 pass: greedy over the tets in index order, a
// chunk closes at TILE_T tets or when the next tet's new blocks or vertices
// would overflow the wave's LDS (TILE_MAXB, TILE_MAXV); a single tet always
// fits.  Colour of (slot k, tet e): how many earlier tets of the chunk add to
// the same block in slot k (for vertices: in corner a).
static int build_tiles(fipc_ctx* c, int nv, int nt, int nnzb, const int* tets, const std::vector<int>& cblock)
{
    std::vector<int4> meta;
    std::vector<unsigned short> tcode((size_t)std::max(nt, 1) * 10), vcode((size_t)std::max(nt, 1) * 4);
    std::vector<int> cblk, cvert;  // per (chunk, local index): the block / vertex
    std::vector<int> bstamp(std::max(nnzb, 1), -1), blid(std::max(nnzb, 1)), vstamp(nv, -1), vlid(nv);
    std::vector<int> ccnt(14 * (size_t)TILE_MAXB);
    int e = 0, ch = 0;
    while (e < nt) {
        const int e0 = e;
        int nb = 0, nvx = 0;
        unsigned char ncol[16] = {0};
        std::fill(ccnt.begin(), ccnt.end(), 0);
        const int pbase = (int)cblk.size(), vbase = (int)cvert.size();
        while (e < nt && e - e0 < TILE_T) {
            int newb = 0, newv = 0;
            for (int k = 0; k < 10; k++) newb += bstamp[cblock[(size_t)e * 10 + k]] != ch;
            for (int a = 0; a < 4; a++) newv += vstamp[tets[4 * e + a]] != ch;
            if (nb + newb > TILE_MAXB || nvx + newv > TILE_MAXV) break;
            for (int k = 0; k < 10; k++) {
                int b = cblock[(size_t)e * 10 + k];
                if (bstamp[b] != ch) { bstamp[b] = ch; blid[b] = nb++; cblk.push_back(b); }
                int col = ccnt[(size_t)k * TILE_MAXB + blid[b]]++;
                tcode[(size_t)k * nt + e] = (unsigned short)(blid[b] | col << 8);
                ncol[k] = (unsigned char)std::max<int>(ncol[k], col + 1);
            }
            for (int a = 0; a < 4; a++) {
                int v = tets[4 * e + a];
                if (vstamp[v] != ch) { vstamp[v] = ch; vlid[v] = nvx++; cvert.push_back(v); }
                int col = ccnt[(size_t)(10 + a) * TILE_MAXB + vlid[v]]++;
                vcode[(size_t)a * nt + e] = (unsigned short)(vlid[v] | col << 8);
                ncol[10 + a] = (unsigned char)std::max<int>(ncol[10 + a], col + 1);
            }
            e++;
        }
        int cw[4];
        memcpy(cw, ncol, 16);
        meta.push_back(make_int4(e0, e - e0, nb, nvx));
        meta.push_back(make_int4(pbase, vbase, 0, 0));
        meta.push_back(make_int4(cw[0], cw[1], cw[2], cw[3]));
        ch++;
    }
    // partial positions: sorted by block (vertex), chunks in order within a run
    auto place = [](const std::vector<int>& ids, int n, std::vector<int>& ptr, std::vector<int>& pos) {
        ptr.assign(n + 1, 0);
        for (int id : ids) ptr[id + 1]++;
        for (int i = 0; i < n; i++) ptr[i + 1] += ptr[i];
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        pos.resize(std::max<size_t>(ids.size(), 1));
        for (size_t k = 0; k < ids.size(); k++) pos[k] = fill[ids[k]]++;
    };
    std::vector<int> pptr, ppos, vptr, vpos;
    place(cblk, nnzb, pptr, ppos);
    place(cvert, nv, vptr, vpos);
    c->nchunk = ch;
    if (meta.empty()) meta.push_back(make_int4(0, 0, 0, 0));
    STCHK(upload(c->tile_meta, meta.data(), meta.size(), c->s));
    STCHK(upload(c->tile_tcode, tcode.data(), tcode.size(), c->s));
    STCHK(upload(c->tile_vcode, vcode.data(), vcode.size(), c->s));
    STCHK(upload(c->tile_ppos, ppos.data(), ppos.size(), c->s));
    STCHK(upload(c->tile_vpos, vpos.data(), vpos.size(), c->s));
    STCHK(upload(c->tile_pptr, pptr.data(), pptr.size(), c->s));
    STCHK(upload(c->tile_vptr, vptr.data(), vptr.size(), c->s));
    // the partial scratch itself is allocated by the first FP32 assembly
    c->tile_nbpart = (long)cblk.size();
    c->tile_nvpart = (long)cvert.size();
    // upload() is asynchronous from the host vectors: wait before they go
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

static int build_pattern(fipc_ctx* c, int nv, int nt, const int* tets)
{
    // upper pattern: row i holds i and its neighbours j > i, sorted
    std::vector<std::vector<int>> adj(nv);
    for (int i = 0; i < nv; i++) adj[i].push_back(i);
    for (int e = 0; e < nt; e++)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                int i = tets[4 * e + a], j = tets[4 * e + b];
                if (j > i) adj[i].push_back(j);
            }
    std::vector<int> rowptr(nv + 1, 0), diag(nv);
    for (int i = 0; i < nv; i++) {
        auto& v = adj[i];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        rowptr[i + 1] = rowptr[i] + (int)v.size();
        diag[i] = rowptr[i];
    }
    int nnzb = rowptr[nv];
    std::vector<int> col(nnzb), brow(nnzb);
    for (int i = 0; i < nv; i++) {
        std::copy(adj[i].begin(), adj[i].end(), col.begin() + rowptr[i]);
        for (int b = rowptr[i]; b < rowptr[i + 1]; b++) brow[b] = i;
    }
    auto block_of = [&](int i, int j) {  // i <= j
        return rowptr[i] + (int)(std::lower_bound(adj[i].begin(), adj[i].end(), j) - adj[i].begin());
    };
    // lower lists: for row j, the blocks (i, j), i < j, in order of i
    std::vector<int> lptr(nv + 1, 0);
    for (int b = 0; b < nnzb; b++)
        if (col[b] > brow[b]) lptr[col[b] + 1]++;
    for (int i = 0; i < nv; i++) lptr[i + 1] += lptr[i];
    int nlower = lptr[nv];
    std::vector<int> lent(2 * (size_t)std::max(nlower, 1)), lfill(lptr.begin(), lptr.end() - 1);
    for (int b = 0; b < nnzb; b++)
        if (col[b] > brow[b]) {
            int k = lfill[col[b]]++;
            lent[2 * k] = b;
            lent[2 * k + 1] = brow[b];
        }
    // contributions: the 10 upper slots of every tet, sorted by their block
    // (counting sort; tets in index order within a block)
    std::vector<int> cblock((size_t)nt * 10);
    std::vector<int> bcnt(nnzb + 1, 0);
    for (int e = 0; e < nt; e++)
        for (int a = 0; a < 4; a++)
            for (int cc = a; cc < 4; cc++) {
                int i = tets[4 * e + a], j = tets[4 * e + cc];
                int bu = block_of(std::min(i, j), std::max(i, j));
                cblock[(size_t)e * 10 + uslot(a, cc)] = bu;
                bcnt[bu + 1]++;
            }
    for (int b = 0; b < nnzb; b++) bcnt[b + 1] += bcnt[b];
    std::vector<int> cpos((size_t)std::max(nt, 1) * 10), fill(bcnt.begin(), bcnt.end() - 1);
    for (size_t k = 0; k < (size_t)nt * 10; k++) cpos[k] = fill[cblock[k]]++;
    // vertex -> (corner, tet) entries for the gradient
    std::vector<int> vcnt(nv + 1, 0);
    for (int e = 0; e < nt; e++)
        for (int a = 0; a < 4; a++) vcnt[tets[4 * e + a] + 1]++;
    for (int i = 0; i < nv; i++) vcnt[i + 1] += vcnt[i];
    std::vector<int> vcent((size_t)std::max(nt, 1) * 4), vfill(vcnt.begin(), vcnt.end() - 1);
    for (int e = 0; e < nt; e++)
        for (int a = 0; a < 4; a++) vcent[vfill[tets[4 * e + a]]++] = a * nt + e;
    // gather order: the rows' diagonal blocks first (each gathers about 20
    // records), then the off-diagonal blocks (1 to 6 records) in index order,
    // so that the seven blocks of a wavefront loop about equally long and the
    // off-diagonal runs are still read in address order
    std::vector<int> gorder(std::max(nnzb, 1));
    {
        int k = 0;
        for (int i = 0; i < nv; i++) gorder[k++] = rowptr[i];  // the diagonal block leads its row
        for (int i = 0; i < nv; i++)
            for (int b = rowptr[i] + 1; b < rowptr[i + 1]; b++) gorder[k++] = b;
    }
    STCHK(build_tiles(c, nv, nt, nnzb, tets, cblock));
    STCHK(upload(c->gorder, gorder.data(), gorder.size(), c->s));
    STCHK(upload(c->bcptr, bcnt.data(), bcnt.size(), c->s));
    STCHK(upload(c->cpos, cpos.data(), cpos.size(), c->s));
    STCHK(upload(c->brow, brow.data(), brow.size(), c->s));
    STCHK(upload(c->lptr, lptr.data(), lptr.size(), c->s));
    STCHK(upload(c->lent, lent.data(), lent.size(), c->s));
    STCHK(upload(c->vcptr, vcnt.data(), vcnt.size(), c->s));
    STCHK(upload(c->vcent, vcent.data(), vcent.size(), c->s));
    c->nnzb = nnzb;
    c->nlower = nlower;
    c->nnzb_full = nnzb + nlower;
    STCHK(upload(c->rowptr, rowptr.data(), rowptr.size(), c->s));
    STCHK(upload(c->col, col.data(), col.size(), c->s));
    STCHK(upload(c->diag, diag.data(), diag.size(), c->s));
    STCHK(ensure(c->val, (size_t)nnzb * 9));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

extern "C" int fipc_upload_mesh(fipc_ctx* c, int nv, const double* x0, const double* m, int nt, const int* tets,
                                const double* IB, const double* vol, const double* la, const double* mu, int nbp,
                                const int* bp, int nbe, const int* be, int nbt, const int* bt,
                                const unsigned char* dirichlet)
{
    if (!c || nv <= 0 || nt < 0) return FIPC_ERR_ARG;
    if (nv >= (1 << 21)) return FIPC_ERR_CAPACITY;
    c->nv = nv; c->nt = nt; c->nbp = nbp; c->nbe = nbe; c->nbt = nbt;
    STCHK(upload(c->x0, x0, (size_t)nv * 3, c->s));
    STCHK(upload(c->m, m, nv, c->s));
    STCHK(upload(c->tets, tets, (size_t)nt * 4, c->s));
    STCHK(upload(c->IB, IB, (size_t)nt * 9, c->s));
    STCHK(upload(c->vol, vol, nt, c->s));
    STCHK(upload(c->la, la, nt, c->s));
    STCHK(upload(c->mu, mu, nt, c->s));
    STCHK(upload(c->bp, bp, nbp, c->s));
    STCHK(upload(c->be, be, (size_t)nbe * 2, c->s));
    STCHK(upload(c->bt, bt, (size_t)nbt * 3, c->s));
    STCHK(upload(c->dmask, dirichlet, nv, c->s));
    size_t n3 = (size_t)nv * 3;
    STCHK(ensure(c->x, n3)); STCHK(ensure(c->xn, n3)); STCHK(ensure(c->xt, n3)); STCHK(ensure(c->xprev, n3));
    STCHK(ensure(c->v, n3)); STCHK(ensure(c->sol, n3)); STCHK(ensure(c->rhs, n3)); STCHK(ensure(c->target, n3));
    STCHK(ensure(c->r, n3)); STCHK(ensure(c->z, n3)); STCHK(ensure(c->pp, n3)); STCHK(ensure(c->q, n3)); STCHK(ensure(c->qc, n3));
    STCHK(ensure(c->cdiag, (size_t)nv * 9)); STCHK(ensure(c->Minv, (size_t)nv * 9));
    HIPCHK(hipMemcpyAsync(c->x.p, x0, n3 * 8, hipMemcpyHostToDevice, c->s));
    c->x_version++;
    c->sys_version++;
    HIPCHK(hipMemcpyAsync(c->xt.p, x0, n3 * 8, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemsetAsync(c->v.p, 0, n3 * 8, c->s));
    STCHK(ensure(c->dprev, n3));
    HIPCHK(hipMemsetAsync(c->dprev.p, 0, n3 * 8, c->s));
    c->nfree = -1;
    HIPCHK(hipMemsetAsync(c->sol.p, 0, n3 * 8, c->s));
    c->sol_version++;
    HIPCHK(hipMemsetAsync(c->rhs.p, 0, n3 * 8, c->s));
    int nmax = std::max(std::max(nv, nt), 4096);
    // block partials: k_pcg_init writes one per block of nblk(nv), k_maxabs
    // (into part_a) at most 1024; the SpMV grid has at most SPMV_MAXB blocks
    STCHK(ensure(c->part_a, std::max(nblk(nmax), 1024) + 64)); STCHK(ensure(c->part_b, nblk(nmax) + 64));
    STCHK(ensure(c->part_c, SPMV_MAXB + 64));
    STCHK(ensure(c->scal, 16)); STCHK(ensure(c->pcg_state, 4));
    STCHK(ensure(c->epart, nblk(nmax) * 4 + 64));
    STCHK(ensure(c->amin, 4)); STCHK(ensure(c->flag, 4)); STCHK(ensure(c->counters, 16)); STCHK(ensure(c->nuniq, 4));
    // broad-phase capacities start from the primitive counts and grow on demand
    for (Grid& g : c->grid)
        if (g.cap == 0) g.cap = 4 * std::max(nbp, std::max(nbe, nbt)) + 4096;
    if (c->ccd_cap == 0) c->ccd_cap = 65536;
    for (int t = 0; t < NTYPES; t++) {
        if (c->cap[t] == 0) c->cap[t] = 1024;
        int w = t == 0 ? 2 : (t == 1 ? 3 : 4);
        int N = 3 * (t == 0 ? 2 : (t == 1 ? 3 : 4));
        STCHK(ensure(c->rows[t], c->cap[t] * w));
        STCHK(ensure(c->cH[t], c->cap[t] * N * N));
        c->count[t] = 0;
    }
    c->cons_version++;  // no rows (and a new Dirichlet mask) from here
    {
        double mmin = 0.0;
        for (int i = 0; i < nv; i++)
            if (!dirichlet[i] && (mmin == 0.0 || m[i] < mmin)) mmin = m[i];
        c->cert_lam = 0.5 * mmin;
    }
    return build_pattern(c, nv, nt, tets);
}

extern "C" int fipc_set_params(fipc_ctx* c, double dt, double dHat, double kappa, double inv_dx, double gravity)
{
    if (!c || !(dt > 0) || !(dHat > 0) || !(inv_dx > 0)) return FIPC_ERR_ARG;
    c->dt = dt; c->dHat = dHat; c->dHat2 = dHat * dHat; c->kappa = kappa; c->inv_dx = inv_dx; c->gravity = gravity;
    c->e_ok = false;
    return FIPC_OK;
}

extern "C" int fipc_set_state(fipc_ctx* c, const double* x, const double* v)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    size_t n = (size_t)c->nv * 3 * 8;
    if (x) {
        HIPCHK(hipMemcpyAsync(c->x.p, x, n, hipMemcpyHostToDevice, c->s));
        c->x_version++;
    }
    if (v) HIPCHK(hipMemcpyAsync(c->v.p, v, n, hipMemcpyHostToDevice, c->s));
    // a new (x, v) state has no previous step
    if (v) HIPCHK(hipMemsetAsync(c->dprev.p, 0, n, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}
extern "C" int fipc_get_state(fipc_ctx* c, double* x, double* v)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    size_t n = (size_t)c->nv * 3 * 8;
    if (x) HIPCHK(hipMemcpyAsync(x, c->x.p, n, hipMemcpyDeviceToHost, c->s));
    if (v) HIPCHK(hipMemcpyAsync(v, c->v.p, n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}
static DVec<double>* vec_of(fipc_ctx* c, int which)
{
    switch (which) {
    case 0: return &c->x;
    case 1: return &c->sol;
    case 2: return &c->rhs;
    case 3: return &c->xt;
    case 4: return &c->v;
    case 5: return &c->xprev;
    default: return nullptr;
    }
}
extern "C" int fipc_get_vector(fipc_ctx* c, int which, double* out)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    DVec<double>* d = vec_of(c, which);
    if (!d || !out) return FIPC_ERR_ARG;
    HIPCHK(hipMemcpyAsync(out, d->p, (size_t)c->nv * 24, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}
extern "C" int fipc_set_vector(fipc_ctx* c, int which, const double* in)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    DVec<double>* d = vec_of(c, which);
    if (!d || !in) return FIPC_ERR_ARG;
    HIPCHK(hipMemcpyAsync(d->p, in, (size_t)c->nv * 24, hipMemcpyHostToDevice, c->s));
    if (d == &c->x) c->x_version++;
    if (d != &c->sol) c->sys_version++;   // a continued solve needs the system it solved
    c->restart_ready = false;
    if (d == &c->sol) c->sol_version++;
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

extern "C" int fipc_begin_step(fipc_ctx* c, const double* target)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    if (target) HIPCHK(hipMemcpyAsync(c->target.p, target, (size_t)c->nv * 24, hipMemcpyHostToDevice, c->s));
    else HIPCHK(hipMemcpyAsync(c->target.p, c->x.p, (size_t)c->nv * 24, hipMemcpyDeviceToDevice, c->s));
    k_begin_step<<<nblk(c->nv), TPB, 0, c->s>>>(c->nv, c->x.p, c->v.p, c->xn.p, c->xt.p, c->dmask.p, c->target.p, c->dt, c->gravity);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}
// prescribed positions of the listed rows (Dirichlet vertices) into target
__global__ void k_scatter_rows(int n, const int* __restrict__ rows, const double* __restrict__ val, double* target)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = rows[i];
    target[3 * r] = val[3 * i]; target[3 * r + 1] = val[3 * i + 1]; target[3 * r + 2] = val[3 * i + 2];
}
extern "C" int fipc_begin_step_rows(fipc_ctx* c, int n, const int* rows, const double* target)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    if (n < 0 || (n && (!rows || !target))) return FIPC_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (rows[i] < 0 || rows[i] >= c->nv) return FIPC_ERR_ARG;
    if (n) {
        STCHK(upload(c->trow, rows, n, c->s));
        STCHK(upload(c->tval, target, (size_t)n * 3, c->s));
        k_scatter_rows<<<nblk(n), TPB, 0, c->s>>>(n, c->trow.p, c->tval.p, c->target.p);
    }
    k_begin_step<<<nblk(c->nv), TPB, 0, c->s>>>(c->nv, c->x.p, c->v.p, c->xn.p, c->xt.p, c->dmask.p, c->target.p, c->dt, c->gravity);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}
extern "C" int fipc_end_step(fipc_ctx* c)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    k_end_step<<<nblk(c->nv * 3), TPB, 0, c->s>>>(c->nv * 3, c->x.p, c->xn.p, c->xt.p, c->dmask.p, c->v.p, c->dprev.p,
                                                   1.0 / c->dt);
    HIPCHK(hipGetLastError());
    return FIPC_OK;  // ordered before the next step's work; host reads of v synchronise
}
extern "C" int fipc_save_prev(fipc_ctx* c)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    HIPCHK(hipMemcpyAsync(c->xprev.p, c->x.p, (size_t)c->nv * 24, hipMemcpyDeviceToDevice, c->s));
    c->xprev_of = c->x_version;
    return FIPC_OK;
}
extern "C" int fipc_apply_sol(fipc_ctx* c, double alpha)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    k_apply<<<nblk(c->nv * 3), TPB, 0, c->s>>>(c->nv * 3, c->xprev.p, c->sol.p, c->x.p, alpha);
    HIPCHK(hipGetLastError());
    c->x_version++;
    // x = xprev + alpha sol with 0 <= alpha <= 1 lies between xprev and
    // xprev + sol coordinate by coordinate (fl(alpha s) is no larger than s
    // and rounding is monotone), inside any box swept along sol from xprev
    if (alpha >= 0.0 && alpha <= 1.0) {
        c->seg_x = c->x_version;
        c->seg_base = c->xprev_of;
        c->seg_sol = c->sol_version;
    } else {
        c->seg_x = 0;
    }
    return FIPC_OK;
}
extern "C" int fipc_residual(fipc_ctx* c, double* res)
{
    if (!c || !c->nv || !res) return FIPC_ERR_STATE;
    if (c->res_version == c->sol_version) {  // read back at the end of the solve that wrote sol
        *res = c->res_cached;
        return FIPC_OK;
    }
    int nb = std::min(nblk(c->nv * 3), 1024);
    k_maxabs<<<nb, TPB, 0, c->s>>>(c->nv * 3, c->sol.p, c->part_a.p);
    HIPCHK(hipGetLastError());
    std::vector<double> h(nb);
    STCHK(d2h_async(c, h.data(), c->part_a.p, nb * 8));
    STCHK(d2h_wait(c));
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    *res = m;
    return FIPC_OK;
}

// ---- assembly ----
static ContactSlots contact_slots(fipc_ctx* c, bool fp32)
{
    ContactSlots sl;
    for (int t = 0; t < NTYPES; t++) {
        sl.rows[t] = c->rows[t].p;
        sl.H[t] = c->cH[t].p;
        sl.Hf[t] = fp32 ? c->cHf[t].p : nullptr;
        sl.c64[t] = fp32 ? c->c64[t].p : nullptr;
        sl.soff[t] = c->soff[t];
    }
    sl.soff[NTYPES] = c->soff[NTYPES];
    return sl;
}
// the contact vertices of the current rows (see fipc_ctx): slot keys, a
// stable radix sort by vertex, run heads, their scan and the compaction; no
// host synchronisation (the vertex count stays on the device)
static int build_contact_vertices(fipc_ctx* c, hipStream_t st)
{
    int S = 0;
    for (int t = 0; t < NTYPES; t++) { c->soff[t] = S; S += c->count[t] * kind_points(t); }
    c->soff[NTYPES] = S;
    STCHK(ensure(c->cv_count, 1));
    STCHK(ensure(c->cgrad, (size_t)std::max(S, 1) * 3));
    c->cv_version = c->cons_version;
    if (S == 0) {
        HIPCHK(hipMemsetAsync(c->cv_count.p, 0, 4, st));
        return FIPC_OK;
    }
    STCHK(ensure(c->cv_keys, S)); STCHK(ensure(c->cv_vals, S)); STCHK(ensure(c->cv_keys2, S));
    STCHK(ensure(c->cv_slot, S)); STCHK(ensure(c->cv_head, (size_t)S + 1));
    STCHK(ensure(c->cv_vert, S)); STCHK(ensure(c->cv_ptr, (size_t)S + 1));
    HIPCHK(hipMemsetAsync(c->cv_count.p, 0, 4, st));
    k_cv_keys<<<nblk(S), TPB, 0, st>>>(S, contact_slots(c, false), c->dmask.p, c->nv, c->cv_keys.p, c->cv_vals.p);
    HIPCHK(hipGetLastError());
    unsigned bits = 1;
    while ((1u << bits) <= (unsigned)c->nv) bits++;  // keys up to nv
    size_t b1 = 0, b2 = 0;
    HIPCHK(rocprim::radix_sort_pairs(nullptr, b1, c->cv_keys.p, c->cv_keys2.p, c->cv_vals.p, c->cv_slot.p, S, 0, bits,
                                     st));
    HIPCHK(rocprim::exclusive_scan(nullptr, b2, c->cv_head.p, c->cv_head.p, 0, (size_t)S + 1, rocprim::plus<int>(),
                                   st));
    size_t need = std::max(b1, b2);
    if (need > c->cub_tmp2.cap) STCHK(ensure(c->cub_tmp2, need));
    b1 = c->cub_tmp2.cap;
    HIPCHK(rocprim::radix_sort_pairs(c->cub_tmp2.p, b1, c->cv_keys.p, c->cv_keys2.p, c->cv_vals.p, c->cv_slot.p, S, 0,
                                     bits, st));
    k_cv_heads<<<nblk((long)S + 1), TPB, 0, st>>>(S, c->nv, c->cv_keys2.p, c->cv_head.p);
    HIPCHK(hipGetLastError());
    b2 = c->cub_tmp2.cap;
    HIPCHK(rocprim::exclusive_scan(c->cub_tmp2.p, b2, c->cv_head.p, c->cv_head.p, 0, (size_t)S + 1,
                                   rocprim::plus<int>(), st));
    k_cv_compact<<<nblk(S), TPB, 0, st>>>(S, c->nv, c->cv_keys2.p, c->cv_head.p, c->cv_vert.p, c->cv_ptr.p,
                                         c->cv_count.p);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}
static int launch_contacts(fipc_ctx* c, int accumulate, hipStream_t st)
{
    if (c->cv_version != c->cons_version) STCHK(build_contact_vertices(c, st));
    ContactBatch cb;
    for (int t = 0; t < NTYPES; t++) {
        cb.rows[t] = c->rows[t].p;
        cb.H[t] = c->cH[t].p;
        cb.n[t] = c->count[t];
        const int N = kind_dofs(t);
        if (c->precision == 1) {
            STCHK(ensure(c->cHf[t], (size_t)std::max(c->count[t], 1) * N * N));
            STCHK(ensure(c->c64[t], (size_t)std::max(c->count[t], 1)));
        }
        cb.Hf[t] = c->precision == 1 ? c->cHf[t].p : nullptr;
        cb.c64[t] = c->precision == 1 ? c->c64[t].p : nullptr;
        cb.cgrad[t] = c->cgrad.p + 3 * (size_t)c->soff[t];
    }
    int nb = contact_blocks(cb);
    if (!nb) return FIPC_OK;
    k_contact<<<nb, TPB, 0, st>>>(cb, c->x.p, c->x0.p, c->dHat2, c->kappa, nullptr, nullptr, 1, accumulate, 20);
    HIPCHK(hipGetLastError());
    const int S = c->soff[NTYPES];
    k_cv_gather<<<nblk((long)CV_G * S), TPB, 0, st>>>(c->cv_count.p, c->cv_vert.p, c->cv_ptr.p, c->cv_slot.p,
                                              contact_slots(c, false), c->cgrad.p, c->rhs.p, c->cdiag.p);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}

extern "C" int fipc_assemble(fipc_ctx* c, int precision)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    c->precision = precision;
    c->sys_version++;
    if (precision == 1) {
        STCHK(ensure(c->valf, (size_t)c->nnzb * 9));
        STCHK(ensure(c->Minvf, (size_t)c->nv * 9));
        STCHK(ensure(c->tile_bpart, (size_t)std::max(c->tile_nbpart, 1L) * 9));
        STCHK(ensure(c->tile_gpart, (size_t)std::max(c->tile_nvpart, 1L) * 3));
    } else {
        STCHK(ensure(c->val, (size_t)c->nnzb * 9));
        STCHK(ensure(c->tetH, (size_t)std::max(c->nt, 1) * 90));
        STCHK(ensure(c->tetg, (size_t)std::max(c->nt, 1) * 12));
    }
    // the contact rows (few, latency-bound) run on the second stream while
    // the tet pass and the block gather fill the GPU: the second stream
    // starts after the inertia term and the cleared diagonal blocks, and the
    // first waits for it before the tet gradients and the preconditioner
    k_rhs_init<<<rhs_grid(c->nv), TPB, 0, c->s>>>(c->nv, c->m.p, c->x.p, c->xt.p, c->dmask.p, c->rhs.p, c->cdiag.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev_rhs, c->s));
    HIPCHK(hipStreamWaitEvent(c->s2, c->ev_rhs, 0));
    STCHK(fipc_tm_begin(c, TM_CONTACT, c->s2));
    STCHK(launch_contacts(c, 1, c->s2));
    STCHK(fipc_tm_end(c, TM_CONTACT, c->s2));
    HIPCHK(hipEventRecord(c->ev_contact, c->s2));
    // FP64: per-tet FP64 records, each block sums its tets in index order.
    // FP32: the tiled pass sums FP32 blocks per chunk in LDS, each block sums
    // its chunk partials in FP64 and stores the FP32 value only (the
    // refinement's residual products read the same FP32 values in FP64
    // arithmetic)
    STCHK(fipc_tm_begin(c, TM_ELASTIC));
    if (c->nt) {
        if (precision == 1)
            k_elastic_tile<<<(c->nchunk + TPB / 64 - 1) / (TPB / 64), TPB, 0, c->s>>>(
                c->nchunk, c->nt, c->tile_meta.p, c->tile_tcode.p, c->tile_vcode.p, c->tile_ppos.p, c->tile_vpos.p,
                c->tets.p, c->IB.p, c->vol.p, c->la.p, c->mu.p, c->x.p, c->dt * c->dt, c->tile_bpart.p,
                c->tile_gpart.p);
        else
            k_elastic<<<nblk(c->nt), TPB, 0, c->s>>>(c->nt, c->tets.p, c->IB.p, c->vol.p, c->la.p, c->mu.p,
                                                              c->x.p, c->dt * c->dt, c->cpos.p, c->tetH.p, c->tetg.p);
    }
    HIPCHK(hipGetLastError());
    STCHK(fipc_tm_end(c, TM_ELASTIC));
    STCHK(fipc_tm_begin(c, TM_GATHER));
    if (precision == 1)
        k_gather_parts<<<(c->nnzb + 2 * GATHER_BPG - 1) / (2 * GATHER_BPG), TPB, 0, c->s>>>(
            c->nnzb, c->tile_pptr.p, c->tile_bpart.p, c->brow.p, c->col.p, c->m.p, c->valf.p);
    else
        k_gather_blocks<double><<<gather_grid(c->nnzb), TPB, 0, c->s>>>(c->nnzb, c->gorder.p, c->bcptr.p, c->tetH.p, c->brow.p,
                                                                         c->col.p, c->m.p, c->val.p, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamWaitEvent(c->s, c->ev_contact, 0));
    if (precision == 1)
        k_gather_tetg<<<rhs_grid(c->nv), TPB, 0, c->s>>>(c->nv, c->tile_vptr.p, nullptr, c->tile_gpart.p, c->dmask.p,
                                                         c->rhs.p);
    else
        k_gather_tetg<<<rhs_grid(c->nv), TPB, 0, c->s>>>(c->nv, c->vcptr.p, c->vcent.p, c->tetg.p, c->dmask.p,
                                                         c->rhs.p);
    HIPCHK(hipGetLastError());
    STCHK(fipc_tm_end(c, TM_GATHER));
    if (precision == 1)
        k_precond<float><<<nblk(c->nv), TPB, 0, c->s>>>(c->nv, c->valf.p, c->diag.p, c->cdiag.p, c->dmask.p,
                                                         c->Minv.p, c->Minvf.p);
    else
        k_precond<double><<<nblk(c->nv), TPB, 0, c->s>>>(c->nv, c->val.p, c->diag.p, c->cdiag.p, c->dmask.p,
                                                          c->Minv.p, nullptr);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}

// ---- SpMV ----
// contact blocks follow the BSR blocks; returns the total grid size
// the contact vertices of the last assembly (an upper bound of their number,
// the slot count, sizes the grid; the lanes past the device count exit)
static int spmv_contacts(fipc_ctx* c, ContactSet& cs, bool fp32 = false)
{
    cs.sl = contact_slots(c, fp32);
    cs.count = c->cv_count.p;
    cs.vert = c->cv_vert.p;
    cs.ptr = c->cv_ptr.p;
    cs.slot = c->cv_slot.p;
    int nb = c->soff[NTYPES] ? nblk((long)CV_G * c->soff[NTYPES]) : 0;
    return (nb + 7) & ~7;
}
// q + qc = A p (q = BSR part, qc = contact part accumulated: qc must be zero
// on entry); prec selects the FP32 copy of the BSR values, V the type of p, q
template <class V>
static int spmv(fipc_ctx* c, int prec, const V* p, V* q, double* qc, double* part, const int* state, int it)
{
    // the contact vertices (and Hessians) are those of the last assembly:
    // rows found since then have none yet
    if (c->cv_version != c->cons_version) return FIPC_ERR_STATE;
    ContactSet cs;
    // the FP32 iterations (FP32 vectors) read the FP32 contact copies
    int nbc = spmv_contacts(c, cs, prec == 1 && sizeof(V) == 4);
    int nbb = spmv_grid(c->nv);
    // products inside the PCG iterations and the residual products (initial
    // guess, FP64 refinement) are timed as separate classes
    const int timer = part ? TM_SPMV : TM_SPMV_RES;
    STCHK(fipc_tm_begin(c, timer));
    if (prec == 1)
        k_spmv<float, V><<<nbc + nbb, TPB, 0, c->s>>>(c->nv, c->rowptr.p, c->col.p, c->lptr.p, (const int2*)c->lent.p,
                                                      c->valf.p, p, q, qc, c->dmask.p, cs, nbc, nbb, part, state, it);
    else
        k_spmv<double, V><<<nbc + nbb, TPB, 0, c->s>>>(c->nv, c->rowptr.p, c->col.p, c->lptr.p, (const int2*)c->lent.p,
                                                       c->val.p, p, q, qc, c->dmask.p, cs, nbc, nbb, part, state, it);
    HIPCHK(hipGetLastError());
    STCHK(fipc_tm_end(c, timer));
    return FIPC_OK;
}
// partials of one product (contact blocks, padded to a multiple of 8, then BSR blocks)
static int spmv_nparts(fipc_ctx* c)
{
    ContactSet cs;
    return spmv_grid(c->nv) + spmv_contacts(c, cs);
}
__global__ void k_add(int n, const double* a, double* b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] += a[i];
}
extern "C" int fipc_spmv(fipc_ctx* c, const double* x, double* y)
{
    if (!c || !c->nv || c->cv_version != c->cons_version) return FIPC_ERR_STATE;
    size_t n = (size_t)c->nv * 24;
    HIPCHK(hipMemcpyAsync(c->pp.p, x, n, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemsetAsync(c->qc.p, 0, n, c->s));
    STCHK(spmv(c, c->precision, c->pp.p, c->q.p, c->qc.p, nullptr, nullptr, 0));
    k_add<<<nblk((long)c->nv * 3), TPB, 0, c->s>>>(c->nv * 3, c->qc.p, c->q.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, c->q.p, n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

// the iterations of one PCG round on vectors of type V (FP64 with the FP64
// Hessian; FP32 in the FP32 Hessian's refinement rounds)
// the certificate of the Newton test at the current solution: max|sol| and
// |b - A sol|_2 from scal[4] (the FP64 residual of the restart just made);
// true when max|sol| + |r|_2 / cert_lam < cert_bound
static int cert_check(fipc_ctx* c, bool* ok)
{
    const int nbm = std::min(nblk(c->nv * 3), 1024);
    k_maxabs<<<nbm, TPB, 0, c->s>>>(c->nv * 3, c->sol.p, c->part_a.p);
    HIPCHK(hipGetLastError());
    std::vector<double> hm(nbm);
    double rr = 0.0;
    STCHK(d2h_async(c, hm.data(), c->part_a.p, (size_t)nbm * 8));
    STCHK(d2h_async(c, &rr, c->scal.p + 4, 8));
    STCHK(d2h_wait(c));
    double mx = 0.0;
    for (double v : hm) mx = std::max(mx, v);
    c->cert_x = mx;
    *ok = rr == rr && mx + sqrt(rr) / c->cert_lam < c->cert_bound;
    return FIPC_OK;
}

template <class V, class M>
static int pcg_round(fipc_ctx* c, int prec, const M* Minv, V* r, V* z, V* p, V* q, int round, int budget,
                     double tol2, int hint, int state[2], bool reuse, bool* init_only, bool* cert_stop)
{
    const int nv = c->nv, nbv = nblk(nv), nsp = spmv_nparts(c), nu = pcg_grid(nv), chunk = PCG_CHUNK;
    const double tol_it = prec == 1 ? tol2 * (PCG_FP32_INNER * PCG_FP32_INNER) : tol2;
    V* sv = sizeof(V) == 4 ? (V*)(void*)c->cgsf.p : (V*)(void*)c->cgs.p;
    // the iterations' x: the FP64 solution, or the FP32 round's correction
    V* xs = sizeof(V) == 4 ? (V*)(void*)c->dxf.p : (V*)(void*)c->sol.p;
    int it = 0;
    state[0] = 0x7fffffff;
    state[1] = 0;
    *init_only = false;
    *cert_stop = false;
    const bool cert = c->cert_bound > 0.0 && sizeof(V) == 4 && round > 0;
    if (reuse) {
        // r, z, p and the scalars still hold the restart at the current
        // solution that ended the last solve; only the test changes
        k_pcg_retest<<<1, 64, 0, c->s>>>(c->scal.p, c->pcg_state.p, tol2);
        HIPCHK(hipGetLastError());
    } else {
        // r = b - A x with the FP64 blocks (q + qc), z = Minv r, p = z; in
        // round 0 the product exits at once when x0 = 0 (the flag is -1)
        const int* x0zero = round == 0 ? c->x0flag.p : nullptr;
        STCHK(spmv<double>(c, c->precision, c->sol.p, c->q.p, c->qc.p, nullptr, x0zero, 0));
        k_pcg_init<V><<<nbv, TPB, 0, c->s>>>(nv, c->rhs.p, c->q.p, c->qc.p, c->dmask.p, c->Minv.p, r, z, p,
                                             c->part_a.p, c->part_b.p, x0zero);
        HIPCHK(hipGetLastError());
        k_pcg_scalars<<<1, TPB, 0, c->s>>>(c->part_a.p, c->part_b.p, nbv, c->scal.p, c->pcg_state.p, round > 0, tol2);
        HIPCHK(hipGetLastError());
        if (round > 0) {
            // the recomputed residual may already meet the tolerance
            STCHK(d2h_async(c, state, c->pcg_state.p, 8));
            STCHK(d2h_wait(c));
            if (state[0] == -1) {
                *init_only = true;
                return FIPC_OK;
            }
        }
    }
    if (cert) {
        // the FP64 residual of the current solution is in scal[4]: the
        // certificate either holds (the round ends, the solve is done) or
        // sets max|sol| for the estimates between chunks
        bool ok = false;
        STCHK(cert_check(c, &ok));
        if (ok) {
            c->cert_hit = true;
            state[0] = -1;
            return FIPC_OK;
        }
        if (it == 0) c->cert_fail++;
    }
    // the host checks convergence between chunks of iterations; the first
    // chunk is sized from a previous solve of the same kind (hint), whose
    // count is usually close
    // (towards a certificate: from the iterations the last one took)
    int first = cert ? std::max(chunk / 2, c->cert_hint - PCG_FIRST_SHORT) : std::max(chunk, hint - PCG_FIRST_SHORT);
    // with a certificate to reach, the next chunk is sized from the
    // residual's decay over the last one, to end about where the estimate
    // crosses the bound (at most chunk iterations)
    int next = chunk;
    double rr_last = -1.0;
    while (it < budget) {
        const int it0 = it;
        int end = std::min(budget, it + (it == 0 ? first : next));
        // single-reduction iterations: the product of u = z, then one update
        for (; it < end; it++) {
            STCHK(spmv<V>(c, prec, z, q, c->qc.p, c->part_c.p, c->pcg_state.p, it));
            STCHK(fipc_tm_begin(c, TM_PCGVEC));
            k_pcg_cg<V, M><<<nu, TPB, 0, c->s>>>(nv, c->scal.p, c->dmask.p, Minv, z, q, c->qc.p, p, sv, xs, r,
                                                 c->part_c.p, nsp, c->part_g.p, nu, c->pcg_state.p, it, tol_it);
            STCHK(fipc_tm_end(c, TM_PCGVEC));
        }
        HIPCHK(hipGetLastError());
        double rr = 0.0;
        STCHK(d2h_async(c, state, c->pcg_state.p, 8));
        if (cert) STCHK(d2h_async(c, &rr, c->scal.p + 4, 8));
        STCHK(d2h_wait(c));
        if (state[0] != 0x7fffffff) break;
        // the recursive residual suggests the certificate holds: end the
        // round, whose successor's restart checks it with the FP64 residual
        // (after a failed check, only at half the bound)
        const double bound = c->cert_bound * (c->cert_fail > 1 ? 0.5 : 1.0);
        if (cert && c->cert_x + sqrt(rr) / c->cert_lam < bound) {
            *cert_stop = true;
            c->cert_hint = it;
            break;
        }
        if (cert) {
            const double target = (bound - c->cert_x) * c->cert_lam;  // the |r|_2 that passes
            const int len = it - it0;
            next = chunk;
            if (target > 0.0 && rr_last > 0.0 && rr < rr_last && len > 0) {
                const double n = ceil(log(target / sqrt(rr)) / (0.5 * log(rr / rr_last) / len));
                next = n >= chunk ? chunk : (n < 1.0 ? 1 : (int)n);
            }
            rr_last = rr;
        }
    }
    // the solution takes an FP32 round's correction
    if (sizeof(V) == 4 && it > 0) {
        k_pcg_flush<<<nblk(3 * nv), TPB, 0, c->s>>>(3 * nv, c->dxf.p, c->sol.p);
        HIPCHK(hipGetLastError());
    }
    // the budget ran out: the last iteration's residual is tested here
    if (state[0] == 0x7fffffff && it > 0 && !*cert_stop) {
        k_pcg_cg_last<<<1, TPB, 0, c->s>>>(c->scal.p, c->part_g.p, nu, c->pcg_state.p, it, tol_it);
        HIPCHK(hipGetLastError());
        STCHK(d2h_async(c, state, c->pcg_state.p, 8));
        STCHK(d2h_wait(c));
    }
    if (budget <= 0) {
        STCHK(d2h_async(c, state, c->pcg_state.p, 8));
        STCHK(d2h_wait(c));
    }
    return FIPC_OK;
}

// PCG rounds from first_round on.  Round 0 starts from the prescribed
// Dirichlet increments; a round > 0 restarts from the current solution with
// the residual b - A x recomputed with the FP64 blocks, against the reference
// norm |r0| of the last round 0.  With the FP32 Hessian the iterations use the
// FP32 copy of the elastic blocks, of the preconditioner and of the iteration
// vectors, and rounds repeat until the recomputed FP64 residual meets the
// tolerance (iterative refinement; the solution accumulates in FP64); with the
// FP64 Hessian a call runs a single round.
static int solve_rounds(fipc_ctx* c, int first_round, double tol, int maxit, int* iters, double* relres)
{
    int nv = c->nv;
    int nbv = nblk(nv);
    const int nsp = spmv_nparts(c);
    const double tol2 = tol * tol;
    const int prec = c->precision;
    STCHK(ensure(c->part_c, (size_t)nsp + 64));
    if (prec == 1) {
        STCHK(ensure(c->rf, (size_t)nv * 3)); STCHK(ensure(c->zf, (size_t)nv * 3));
        STCHK(ensure(c->pf, (size_t)nv * 3)); STCHK(ensure(c->qf, (size_t)nv * 3));
        STCHK(ensure(c->cgsf, (size_t)nv * 3));
        if (c->dxf.cap < (size_t)nv * 3) {
            STCHK(ensure(c->dxf, (size_t)nv * 3));
            HIPCHK(hipMemsetAsync(c->dxf.p, 0, c->dxf.cap * sizeof(float), c->s));
        }
    } else {
        STCHK(ensure(c->cgs, (size_t)nv * 3));
    }
    STCHK(ensure(c->part_g, 4 * PCG_MAXB));
    // a certified continuation may end rounds early: two more rounds
    const int rounds = prec == 1 ? (c->cert_bound > 0.0 ? 6 : 4) : 1;
    int total = 0, state[2] = {0, 0};
    bool converged = false, init_only = false, cert_stop = false;
    // a continuation in FP32 mode starts from the restart that ended the
    // solve it continues, when there is one (its FP64 residual is current)
    bool reuse = first_round > 0 && prec == 1 && c->restart_ready;
    c->restart_ready = false;
    c->sol_version++;  // this solve writes sol
    for (int round = first_round; round < first_round + rounds; round++) {
        if (round == 0) {
            // x0 = prescribed Dirichlet increments
            STCHK(ensure(c->x0flag, 1));
            HIPCHK(hipMemsetAsync(c->x0flag.p, 0xff, 4, c->s));
            k_pcg_x0<<<nbv, TPB, 0, c->s>>>(nv, c->xt.p, c->x.p, c->dmask.p, c->sol.p, c->qc.p, c->x0flag.p);
            HIPCHK(hipGetLastError());
        } else {
            HIPCHK(hipMemsetAsync(c->qc.p, 0, (size_t)nv * 24, c->s));
        }
        int budget = maxit - total;
        // first chunk: the last first round's count, or that of the last
        // continuation (fipc_solve_refine) for its first round
        int hint = round == 0 ? c->pcg_hint : (round == first_round ? c->refine_hint : 0);
        const bool ru = reuse && round == first_round;
        if (prec == 1)
            STCHK((pcg_round<float, float>(c, prec, c->Minvf.p, c->rf.p, c->zf.p, c->pf.p, c->qf.p, round, budget,
                                           tol2, hint, state, ru, &init_only, &cert_stop)));
        else
            STCHK((pcg_round<double, double>(c, prec, c->Minv.p, c->r.p, c->z.p, c->pp.p, c->q.p, round, budget,
                                             tol2, hint, state, ru, &init_only, &cert_stop)));
        if (state[0] == -1 && round > 0) { converged = true; break; }
        int ran = state[0] == -1 ? 0 : state[1];
        total += ran;
        if (round == 0 && ran > 0) c->pcg_hint = ran;
        if (round == first_round && first_round > 0 && ran > 0 && !cert_stop) c->refine_hint = ran;
        if (cert_stop) continue;                  // the next round's restart checks the certificate
        if (state[0] == 0x7fffffff) break;        // iteration budget spent
        if (prec != 1 || state[0] == -1) { converged = true; break; }
        // FP32 iterations converged: the next round checks the FP64 residual
    }
    // refinement rounds exhausted after a converged FP32 round: the FP64
    // residual of its result decides (one more product, no iterations); the
    // solve is reported as not converged when it misses the tolerance
    if (!converged && state[0] != 0x7fffffff) {
        int chk[2] = {0, 0};
        STCHK((pcg_round<float, float>(c, prec, c->Minvf.p, c->rf.p, c->zf.p, c->pf.p, c->qf.p,
                                       first_round + rounds, 0, tol2, 0, chk, false, &init_only, &cert_stop)));
        converged = chk[0] == -1;
    }
    // a solve that ended on a restart whose FP64 residual met the tolerance
    // leaves that restart for a continuation
    c->restart_ready = prec == 1 && init_only;
    // the scalars and max |sol| (the Newton residual, output_residual) are
    // read back together: fipc_residual returns the kept value while sol is
    // unchanged
    double sc[5];
    const int nbm = std::min(nblk(nv * 3), 1024);
    k_maxabs<<<nbm, TPB, 0, c->s>>>(nv * 3, c->sol.p, c->part_a.p);
    HIPCHK(hipGetLastError());
    std::vector<double> hm(nbm);
    STCHK(d2h_async(c, sc, c->scal.p, 40));
    STCHK(d2h_async(c, hm.data(), c->part_a.p, (size_t)nbm * 8));
    STCHK(d2h_wait(c));
    double mx = 0.0;
    for (double v : hm) mx = std::max(mx, v);
    c->res_cached = mx;
    c->res_version = c->sol_version;
    double rel = sc[2] > 0 ? sqrt(sc[4] / sc[2]) : 0.0;
    c->last_pcg_iters = total;
    c->last_pcg_relres = rel;
    if (iters) *iters = total;
    if (relres) *relres = rel;
    if (!(rel == rel)) return FIPC_ERR_NONFINITE;
    if (!converged) return FIPC_ERR_NOT_CONVERGED;
    return FIPC_OK;
}
// solve_system (NEWTON_IPC.py:750-777): block-Jacobi PCG from the prescribed
// Dirichlet increments
extern "C" int fipc_solve(fipc_ctx* c, double tol, int maxit, int* iters, double* relres)
{
    if (!c || !c->nv || c->cv_version != c->cons_version) return FIPC_ERR_STATE;
    int rc = solve_rounds(c, 0, tol, maxit, iters, relres);
    c->solve_version = c->sys_version;
    return rc;
}
// continues the last fipc_solve of the current system from its solution to
// the tighter tolerance tol (relative to the same |r0|)
extern "C" int fipc_solve_refine(fipc_ctx* c, double tol, int maxit, int* iters, double* relres)
{
    if (!c || !c->nv || c->solve_version != c->sys_version || c->cv_version != c->cons_version)
        return FIPC_ERR_STATE;
    return solve_rounds(c, 1, tol, maxit, iters, relres);
}

// fipc_solve_refine that stops as soon as the exact direction provably passes
// the Newton test max|p| < bound: A >= M (the lumped mass; the elastic and
// barrier Hessians are projected to PSD), so |p* - p|_inf <= |p* - p|_2 <=
// |b - A p|_2 / min m, halved for the FP32 storage of A.  Checked with the
// FP64 residual at round restarts, which the recursive residual triggers.
extern "C" int fipc_solve_certify(fipc_ctx* c, double tol, int maxit, double bound, int* iters, double* relres,
                                  int* certified)
{
    if (!c || !c->nv || c->solve_version != c->sys_version || c->cv_version != c->cons_version)
        return FIPC_ERR_STATE;
    if (!(bound > 0.0)) return FIPC_ERR_ARG;
    c->cert_hit = false;
    c->cert_fail = 0;
    c->cert_bound = c->precision == 1 && c->cert_lam > 0.0 ? bound : 0.0;
    const int rc = solve_rounds(c, 1, tol, maxit, iters, relres);
    c->cert_bound = 0.0;
    if (certified) *certified = c->cert_hit ? 1 : 0;
    return rc;
}

// ---- energy ----
extern "C" int fipc_energy(fipc_ctx* c, double* E, double* parts)
{
    if (!c || !c->nv || !E) return FIPC_ERR_STATE;
    int nbv = nblk(c->nv), nbt = c->nt ? nblk(c->nt) : 0;
    if (c->e_ok && c->e_xv == c->x_version && c->e_cv == c->cons_version) {
        // same x, constraints and parameters as the last evaluation: only
        // the inertia term (x tilde may have moved) is evaluated again; the
        // sum is formed in the same order, so the bits are those of a full
        // evaluation
        double* ep = c->epart.p;
        STCHK(fipc_tm_begin(c, TM_ENERGY));
        k_energy_inertia<<<nbv, TPB, 0, c->s>>>(c->nv, c->m.p, c->x.p, c->xt.p, ep);
        HIPCHK(hipGetLastError());
        STCHK(fipc_tm_end(c, TM_ENERGY));
        std::vector<double> h(nbv);
        STCHK(d2h_async(c, h.data(), ep, (size_t)nbv * 8));
        STCHK(d2h_wait(c));
        double ei = 0;
        for (int i = 0; i < nbv; i++) ei += h[i];
        *E = ei + c->e_el + c->e_bar;
        if (parts) { parts[0] = ei; parts[1] = c->e_el; parts[2] = c->e_bar; }
        if (!(*E == *E)) return FIPC_ERR_NONFINITE;
        return FIPC_OK;
    }
    std::vector<int> nbc(NTYPES, 0);
    size_t off = nbv + nbt;
    size_t total = off;
    for (int t = 0; t < NTYPES; t++) { nbc[t] = c->count[t] ? nblk(c->count[t]) : 0; total += nbc[t]; }
    // the buffer may grow with the contact counts: size it before any launch writes into it
    STCHK(ensure(c->epart, total + 8));
    double* ep = c->epart.p;
    STCHK(fipc_tm_begin(c, TM_ENERGY));
    k_energy_inertia<<<nbv, TPB, 0, c->s>>>(c->nv, c->m.p, c->x.p, c->xt.p, ep);
    if (nbt) k_energy_elastic<<<nbt, TPB, 0, c->s>>>(c->nt, c->tets.p, c->IB.p, c->vol.p, c->la.p, c->mu.p, c->x.p, c->dt * c->dt, ep + nbv);
    HIPCHK(hipGetLastError());
    size_t o = off;
    for (int t = 0; t < NTYPES; t++) {
        if (!nbc[t]) continue;
#define LAUNCH_E(K) k_energy_contact<K><<<nbc[t], TPB, 0, c->s>>>(c->count[t], c->rows[t].p, c->x.p, c->x0.p, c->dHat2, c->kappa, ep + o)
        switch (t) {
        case 0: LAUNCH_E(0); break;
        case 1: LAUNCH_E(1); break;
        case 2: LAUNCH_E(2); break;
        case 3: LAUNCH_E(3); break;
        case 4: LAUNCH_E(4); break;
        case 5: LAUNCH_E(5); break;
        default: LAUNCH_E(6); break;
        }
#undef LAUNCH_E
        o += nbc[t];
    }
    HIPCHK(hipGetLastError());
    STCHK(fipc_tm_end(c, TM_ENERGY));
    std::vector<double> h(total);
    STCHK(d2h_async(c, h.data(), ep, total * 8));
    STCHK(d2h_wait(c));
    double ei = 0, ee = 0, eb = 0;
    for (int i = 0; i < nbv; i++) ei += h[i];
    for (int i = 0; i < nbt; i++) ee += h[nbv + i];
    for (size_t i = off; i < total; i++) eb += h[i];
    *E = ei + ee + eb;
    if (parts) { parts[0] = ei; parts[1] = ee; parts[2] = eb; }
    c->e_el = ee;
    c->e_bar = eb;
    c->e_xv = c->x_version;
    c->e_cv = c->cons_version;
    c->e_ok = true;
    if (!(*E == *E)) return FIPC_ERR_NONFINITE;
    return FIPC_OK;
}

extern "C" int fipc_min_distance(fipc_ctx* c, double* d2min)
{
    if (!c || !c->nv || !d2min) return FIPC_ERR_STATE;
    double big = 1999999999.0;
    unsigned long long init;
    memcpy(&init, &big, 8);
    HIPCHK(hipMemcpyAsync(c->amin.p, &init, 8, hipMemcpyHostToDevice, c->s));
    for (int t = 0; t < NTYPES; t++) {
        int n = c->count[t];
        if (!n) continue;
#define LAUNCH_M(K) k_min_dist<K><<<nblk(n), TPB, 0, c->s>>>(n, c->rows[t].p, c->x.p, c->amin.p)
        switch (t) {
        case 0: LAUNCH_M(0); break;
        case 1: LAUNCH_M(1); break;
        case 2: LAUNCH_M(2); break;
        case 3: LAUNCH_M(3); break;
        case 4: LAUNCH_M(4); break;
        case 5: LAUNCH_M(5); break;
        default: LAUNCH_M(6); break;
        }
#undef LAUNCH_M
    }
    HIPCHK(hipGetLastError());
    unsigned long long r;
    STCHK(d2h_async(c, &r, c->amin.p, 8));
    STCHK(d2h_wait(c));
    double d;
    memcpy(&d, &r, 8);
    *d2min = d;
    return FIPC_OK;
}

int fipc_inversion_alpha(fipc_ctx* c, unsigned long long* dst)
{
    if (!c->nt) return FIPC_OK;
    k_inversion<<<nblk(c->nt), TPB, 0, c->s>>>(c->nt, c->tets.p, c->x.p, c->sol.p, dst);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}

// plain streaming copy between two device buffers (bandwidth reference and
// hardware-counter calibration: its algorithmic traffic is exactly 2 n bytes)
__global__ void k_stream_copy(long n2, const double2* __restrict__ a, double2* __restrict__ b)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}
extern "C" int fipc_stream_copy(fipc_ctx* c, long nbytes, double* ms)
{
    if (!c || nbytes <= 0 || nbytes % 16) return FIPC_ERR_ARG;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, nbytes));
    HIPCHK(hipMalloc(&b, nbytes));
    HIPCHK(hipMemsetAsync(a, 0, nbytes, c->s));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    long n2 = nbytes / 16;
    int nb = (int)std::min<long>((n2 + TPB - 1) / TPB, 8192);
    HIPCHK(hipEventRecord(e0, c->s));
    k_stream_copy<<<nb, TPB, 0, c->s>>>(n2, (const double2*)a, (double2*)b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, c->s));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return FIPC_OK;
}

extern "C" int fipc_get_bsr(fipc_ctx* c, int* nnzb, int* rowptr, int* col, double* val)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    if (nnzb) *nnzb = c->nnzb_full;
    if (!rowptr && !col && !val) return FIPC_OK;
    // the stored upper triangle expanded to the full matrix: row i lists the
    // transposes of (j, i), j < i, then its own blocks, columns ascending
    int nv = c->nv, nu = c->nnzb, nl = c->nlower;
    std::vector<int> up(nv + 1), ucol(nu), lp(nv + 1), le(2 * (size_t)std::max(nl, 1));
    std::vector<double> v((size_t)nu * 9);
    HIPCHK(hipMemcpyAsync(up.data(), c->rowptr.p, up.size() * 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(ucol.data(), c->col.p, ucol.size() * 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(lp.data(), c->lptr.p, lp.size() * 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(le.data(), c->lent.p, le.size() * 4, hipMemcpyDeviceToHost, c->s));
    if (val) {
        if (c->precision == 1) {
            std::vector<float> vf(v.size());
            HIPCHK(hipMemcpyAsync(vf.data(), c->valf.p, vf.size() * 4, hipMemcpyDeviceToHost, c->s));
            HIPCHK(hipStreamSynchronize(c->s));
            for (size_t i = 0; i < vf.size(); i++) v[i] = vf[i];
        } else {
            HIPCHK(hipMemcpyAsync(v.data(), c->val.p, v.size() * 8, hipMemcpyDeviceToHost, c->s));
        }
    }
    HIPCHK(hipStreamSynchronize(c->s));
    size_t o = 0;
    for (int i = 0; i < nv; i++) {
        if (rowptr) rowptr[i] = (int)o;
        for (int k = lp[i]; k < lp[i + 1]; k++, o++) {
            int b = le[2 * k];
            if (col) col[o] = le[2 * k + 1];
            if (val)
                for (int r = 0; r < 3; r++)
                    for (int cc = 0; cc < 3; cc++) val[o * 9 + 3 * r + cc] = v[(size_t)b * 9 + 3 * cc + r];
        }
        for (int b = up[i]; b < up[i + 1]; b++, o++) {
            if (col) col[o] = ucol[b];
            if (val)
                for (int k = 0; k < 9; k++) val[o * 9 + k] = v[(size_t)b * 9 + k];
        }
    }
    if (rowptr) rowptr[nv] = (int)o;
    return FIPC_OK;
}

extern "C" int fipc_bsr_storage(fipc_ctx* c, int* nupper, int* nlower)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    if (nupper) *nupper = c->nnzb;
    if (nlower) *nlower = c->nlower;
    return FIPC_OK;
}

extern "C" int fipc_enable_timing(fipc_ctx* c, int on)
{
    if (!c) return FIPC_ERR_ARG;
    STCHK(fipc_tm_fold(c));
    c->timing = on != 0;
    for (auto& t : c->tm) { t.total_ms = 0; t.launches = 0; }
    return FIPC_OK;
}
extern "C" int fipc_kernel_times(fipc_ctx* c, int n, double* ms, long* launches)
{
    if (!c || n < 0) return FIPC_ERR_ARG;
    STCHK(fipc_tm_fold(c));
    for (int i = 0; i < n && i < FIPC_NTIMERS; i++) {
        if (ms) ms[i] = c->tm[i].total_ms;
        if (launches) launches[i] = c->tm[i].launches;
    }
    return FIPC_OK;
}
extern "C" int fipc_sync(fipc_ctx* c)
{
    if (!c) return FIPC_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

// ---- batched per-constraint / per-tet evaluation (tests, drivers) ----
extern "C" int fipc_batch_constraint(int type, int n, const double* x, const double* x0, double dHat2, double kappa,
                                     int project, double* E, double* g, double* H)
{
    if (type < 0 || type >= NTYPES || n < 0) return FIPC_ERR_ARG;
    if (n == 0) return FIPC_OK;
    int K = type == 0 ? 2 : (type == 1 ? 3 : 4);
    int W = K, N = 3 * K;
    // the kernel reads rows of vertex ids: make identity rows over the packed points
    std::vector<int> rows((size_t)n * W);
    for (int r = 0; r < n; r++)
        for (int a = 0; a < W; a++) rows[(size_t)r * W + a] = r * K + a;
    double *dx, *dx0, *dE, *dg, *dH;
    int* drows;
    size_t npts = (size_t)n * K * 3;
    HIPCHK(hipMalloc(&dx, npts * 8));
    HIPCHK(hipMalloc(&dx0, npts * 8));
    HIPCHK(hipMalloc(&dE, (size_t)n * 8));
    HIPCHK(hipMalloc(&dg, (size_t)n * N * 8));
    HIPCHK(hipMalloc(&dH, (size_t)n * N * N * 8));
    HIPCHK(hipMalloc(&drows, rows.size() * 4));
    HIPCHK(hipMemcpy(dx, x, npts * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dx0, x0 ? x0 : x, npts * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(drows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
    ContactBatch cb;
    for (int t = 0; t < NTYPES; t++) {
        cb.rows[t] = drows; cb.H[t] = dH; cb.n[t] = t == type ? n : 0;
        cb.Hf[t] = nullptr; cb.c64[t] = nullptr; cb.cgrad[t] = nullptr;
    }
    int nb = contact_blocks(cb);
    k_contact<<<nb, TPB>>>(cb, dx, dx0, dHat2, kappa, dE, dg, project, 0, 20);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    if (E) HIPCHK(hipMemcpy(E, dE, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (g) HIPCHK(hipMemcpy(g, dg, (size_t)n * N * 8, hipMemcpyDeviceToHost));
    if (H) HIPCHK(hipMemcpy(H, dH, (size_t)n * N * N * 8, hipMemcpyDeviceToHost));
    (void)hipFree(dx); (void)hipFree(dx0); (void)hipFree(dE); (void)hipFree(dg); (void)hipFree(dH); (void)hipFree(drows);
    return FIPC_OK;
}

extern "C" int fipc_batch_tet(int n, const double* F, const double* IB, const double* scale, const double* la,
                              const double* mu, double* psi, double* P, double* Hout)
{
    if (n < 0 || !F || !IB || !scale || !la || !mu || !psi || !P) return FIPC_ERR_ARG;
    if (n == 0) return FIPC_OK;
    double *dF, *dIB, *ds, *dla, *dmu, *dpsi, *dP, *dH = nullptr;
    HIPCHK(hipMalloc(&dF, (size_t)n * 72)); HIPCHK(hipMalloc(&dIB, (size_t)n * 72)); HIPCHK(hipMalloc(&ds, (size_t)n * 8));
    HIPCHK(hipMalloc(&dla, (size_t)n * 8)); HIPCHK(hipMalloc(&dmu, (size_t)n * 8)); HIPCHK(hipMalloc(&dpsi, (size_t)n * 8));
    HIPCHK(hipMalloc(&dP, (size_t)n * 72));
    if (Hout) HIPCHK(hipMalloc(&dH, (size_t)n * 144 * 8));
    HIPCHK(hipMemcpy(dF, F, (size_t)n * 72, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dIB, IB, (size_t)n * 72, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ds, scale, (size_t)n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dla, la, (size_t)n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dmu, mu, (size_t)n * 8, hipMemcpyHostToDevice));
    k_batch_tet<<<nblk(n), TPB>>>(n, dF, dIB, ds, dla, dmu, dpsi, dP, dH);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(psi, dpsi, (size_t)n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(P, dP, (size_t)n * 72, hipMemcpyDeviceToHost));
    if (Hout) HIPCHK(hipMemcpy(Hout, dH, (size_t)n * 144 * 8, hipMemcpyDeviceToHost));
    (void)hipFree(dF); (void)hipFree(dIB); (void)hipFree(ds); (void)hipFree(dla); (void)hipFree(dmu); (void)hipFree(dpsi);
    (void)hipFree(dP);
    if (dH) (void)hipFree(dH);
    return FIPC_OK;
}
