// fipc_mlp.hip -- per-vertex MLP of the learned Newton warm start on gfx950:
// products on the BF16 matrix cores (v_mfma_f32_32x32x16_bf16) with FP32
// accumulation, FP32 master weights trained with Adam, and a device-resident
// dataset of (feature, target) rows captured from solver trajectories.
//
// Network: n_in -> depth hidden layers of `width` (SiLU) -> n_out (linear).
// Parameters are packed per layer as W_l [fan_in][fan_out] (row-major, so a
// layer is H_out = act(H_in W_l + b_l)) followed by b_l [fan_out].
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "fipc_ctx.h"
#include "fipc_mlp.h"

#define TPB 256
static inline int nblk(long n, int t = TPB) { return (int)((n + t - 1) / t); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

enum { EPI_NONE = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_DSILU = 3 };

// ---------------------------------------------------------------------------
// GEMM: C[M][N] = epilogue(A[M][K] B[K][N]).  A and B are FP32 in memory with
// arbitrary strides (A(m,k) = A[m sam + k sak], B(k,n) = B[k sbk + n sbn]), so
// transposed operands cost nothing; tiles are rounded to BF16 when staged in
// LDS.  A block of 4 waves computes a 64 x 64 tile, each wave one 32 x 32
// MFMA tile.  Operand layout of v_mfma_f32_32x32x16_bf16 (wave64): lane l
// supplies A[l % 32][8 (l / 32) + i] and B[8 (l / 32) + i][l % 32], i < 8;
// accumulator element j of lane l is D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].
// gridDim.z > 1 splits K: slice z covers [z kslice, (z + 1) kslice) and
// writes its partial product to C + z c_slice (EPI_NONE only).
// ---------------------------------------------------------------------------
#define GBM 64
#define GBN 64
#define GBK 32
#define LDSK (GBK + 8)

struct GemmArgs {
    int M, N, K;
    const float* A;
    long sam, sak;
    const float* B;
    long sbk, sbn;
    float* C;
    long ldc;
    long c_slice;
    int kslice;
    int epi;
    const float* bias;   // EPI_BIAS, EPI_BIAS_SILU
    const float* Z;      // EPI_DSILU: pre-activation, leading dimension ldc
    float* Zout;         // EPI_BIAS_SILU: pre-activation output (may be null)
};

// logistic sigmoid with the hardware reciprocal (about 1 ulp): the products
// that consume it round to BF16
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[GBM][LDSK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[GBN][LDSK];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;
    const int kbeg = blockIdx.z * g.kslice;
    const int kend = min(g.K, kbeg + g.kslice);
    floatx16 acc;
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = 0.f;
    // staging: thread -> tile row sr (0..63), k offsets sk..sk+7
    const int sr = tid >> 2, sk = (tid & 3) * 8;
    const int am = m0 + sr, bn = n0 + sr;
    for (int k0 = kbeg; k0 < kend; k0 += GBK) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            int k = k0 + sk + j;
            bool kin = k < kend;
            float a = (am < g.M && kin) ? g.A[(long)am * g.sam + (long)k * g.sak] : 0.f;
            float b = (bn < g.N && kin) ? g.B[(long)k * g.sbk + (long)bn * g.sbn] : 0.f;
            As[sr][sk + j] = (__bf16)a;
            Bs[sr][sk + j] = (__bf16)b;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 16) {
            bf16x8 a = *(const bf16x8*)&As[wm * 32 + (lane & 31)][kk + (lane >> 5) * 8];
            bf16x8 b = *(const bf16x8*)&Bs[wn * 32 + (lane & 31)][kk + (lane >> 5) * 8];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    float* C = g.C + (long)blockIdx.z * g.c_slice;
    const int col = n0 + wn * 32 + (lane & 31);
    if (col >= g.N) return;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        int row = m0 + wm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        if (row >= g.M) continue;
        long o = (long)row * g.ldc + col;
        float v = acc[j];
        if (g.epi == EPI_BIAS_SILU) {
            v += g.bias[col];
            if (g.Zout) g.Zout[o] = v;
            v = v * sigm(v);
        } else if (g.epi == EPI_BIAS) {
            v += g.bias[col];
        } else if (g.epi == EPI_DSILU) {
            float z = g.Z[o], s = sigm(z);
            v *= s * (1.f + z * (1.f - s));
        }
        C[o] = v;
    }
}

static int gemm(hipStream_t s, int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk,
                long sbn, float* C, long ldc, int epi, const float* bias, const float* Z, float* Zout, int splits = 1,
                long c_slice = 0)
{
    if (M <= 0 || N <= 0) return FIPC_OK;
    GemmArgs g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.sam = sam; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbn = sbn;
    g.C = C; g.ldc = ldc; g.c_slice = c_slice;
    int per = (K + splits - 1) / splits;
    g.kslice = ((per + GBK - 1) / GBK) * GBK;
    g.epi = epi; g.bias = bias; g.Z = Z; g.Zout = Zout;
    dim3 grid((M + GBM - 1) / GBM, (N + GBN - 1) / GBN, splits);
    k_gemm_bf16<<<grid, 256, 0, s>>>(g);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}

// ---------------------------------------------------------------------------
// elementwise and reduction kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float unif(unsigned long long seed, unsigned long long i)
{
    return (float)((mix64(seed * 0x2545f4914f6cdd1dull + i) >> 40) * (1.0 / 16777216.0));  // [0, 1)
}
// W uniform in (-a, a), b = 0
__global__ void k_init_layer(long nw, int nb, float* W, float* b, float a, unsigned long long seed)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < nw) W[i] = (2.f * unif(seed, i) - 1.f) * a;
    if (i < nb) b[i] = 0.f;
}
// out[i] = sum over slices of P[s n + i], in slice order
__global__ void k_sum_slices(long n, int S, const float* P, float* out)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = 0.f;
    for (int s = 0; s < S; s++) v += P[s * n + i];
    out[i] = v;
}
// column sums of A[M][N] (N <= 64 per block column group): out[c] = sum_m A[m][c]
__global__ void k_colsum(int M, int N, const float* A, float* out)
{
    __shared__ float sh[4][64];
    int c = blockIdx.x * 64 + (threadIdx.x & 63), r0 = threadIdx.x >> 6;
    float v = 0.f;
    if (c < N)
        for (int m = r0; m < M; m += 4) v += A[(long)m * N + c];
    sh[r0][threadIdx.x & 63] = v;
    __syncthreads();
    if (r0 == 0 && c < N) out[c] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}
// dY = 2 (Y - T) / (M N); per-block partial sums of (Y - T)^2 / (M N)
__global__ void k_mse(long n, float inv, const float* Y, const float* T, float* dY, float* part)
{
    __shared__ float sh[TPB];
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        float d = Y[i] - T[i];
        dY[i] = 2.f * d * inv;
        e = d * d * inv;
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_adam(long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// batch of B rows drawn uniformly (with replacement) from the dataset
__global__ void k_gather_batch(int B, long nds, unsigned long long seed, int nin, int nout, const float* dsX,
                               const float* dsT, float* X, float* T)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    long r = (long)(mix64(seed * 0x9e3779b97f4a7c15ull + (unsigned long long)i) % (unsigned long long)nds);
    for (int k = 0; k < nin; k++) X[(long)i * nin + k] = dsX[r * nin + k];
    for (int k = 0; k < nout; k++) T[(long)i * nout + k] = dsT[r * nout + k];
}

#ifndef FI_RT
#define FI_RT 4
#endif
#ifndef FI_ROWS
#define FI_ROWS (32 * FI_RT)
#endif
// a wave takes 64 output features (two 32-feature tiles) of FI_RT 32-row
// tiles: wave w takes features 64 (w % 4) .. +64 of rows 32 FI_RT (w / 4) ..
constexpr int FI_WROWS = 32 * FI_RT;
static_assert(FI_ROWS % FI_WROWS == 0, "rows per block");
constexpr int FI_THREADS = 4 * (FI_ROWS / FI_WROWS) * 64;
#define FI_W 256
#define FI_LDS (FI_W + 8)
// D[feature][row] += W[feature][k..k+15] . H[row][k..k+15] over the layer's
// fan-in KIN for NF 32-feature tiles from f0 and RT 32-row tiles from rbase.
// The A operand is a weight fragment (8 consecutive k of one output feature,
// one 16-byte load from the transposed BF16 weights), the B operand an
// activation fragment (8 consecutive k of one row, one 16-byte LDS read);
// the next k step's weight fragments are loaded before this step's products.
template <int NF, int RT>
__device__ __forceinline__ void fi_step(const bf16x8 (&a)[NF], const __bf16 (*H)[FI_LDS], int rbase, int lane, int k0,
                                        floatx16 (&acc)[NF][RT])
{
    bf16x8 b[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) b[rt] = *(const bf16x8*)&H[rbase + rt * 32 + (lane & 31)][k0 + (lane >> 5) * 8];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
            acc[f][rt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[f], b[rt], acc[f][rt], 0, 0, 0);
}
template <int KIN, int NF, int RT>
__device__ __forceinline__ void fi_layer(const __bf16* __restrict__ W, const __bf16 (*H)[FI_LDS], int f0, int rbase,
                                         int lane, floatx16 (&acc)[NF][RT])
{
    static_assert(KIN == 16 || KIN % 32 == 0, "fan-in");
    const __bf16* wp[NF];
    bf16x8 a[NF], an[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        wp[f] = W + (long)(f0 + f * 32 + (lane & 31)) * KIN + (lane >> 5) * 8;
        a[f] = *(const bf16x8*)wp[f];
    }
    if (KIN == 16) {
        fi_step<NF, RT>(a, H, rbase, lane, 0, acc);
        return;
    }
    // two k steps per pass in alternating registers: the fragments of the
    // next step are in flight during this step's products (the last pass
    // reloads its own last step rather than read past the row)
    for (int k0 = 0; k0 < KIN; k0 += 32) {
#pragma unroll
        for (int f = 0; f < NF; f++) an[f] = *(const bf16x8*)(wp[f] + k0 + 16);
        fi_step<NF, RT>(a, H, rbase, lane, k0, acc);
        const int kn = k0 + 32 < KIN ? k0 + 32 : KIN - 16;
#pragma unroll
        for (int f = 0; f < NF; f++) a[f] = *(const bf16x8*)(wp[f] + kn);
        fi_step<NF, RT>(an, H, rbase, lane, k0 + 16, acc);
    }
}
__device__ __forceinline__ unsigned int bf16_bits(float v)
{
    return (unsigned int)__builtin_bit_cast(unsigned short, (__bf16)v);
}
// hidden-layer epilogue: bias, SiLU, BF16 into the activation tile.  Lane l
// holds features f + 8 (j / 4) + 4 (l / 32) + j % 4 of row (l % 32) of each
// tile, so each group of four j is four consecutive features of one row:
// one 8-byte LDS store
template <int RT>
__device__ __forceinline__ void fi_silu_store(const floatx16 (&acc)[2][RT], const float* __restrict__ bias,
                                              __bf16 (*H)[FI_LDS], int f0, int rbase, int lane)
{
#pragma unroll
    for (int ft = 0; ft < 2; ft++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int m = f0 + ft * 32 + 8 * g + 4 * (lane >> 5);
            const float b0 = bias[m], b1 = bias[m + 1], b2 = bias[m + 2], b3 = bias[m + 3];
#pragma unroll
            for (int rt = 0; rt < RT; rt++) {
                float z0 = acc[ft][rt][4 * g] + b0, z1 = acc[ft][rt][4 * g + 1] + b1;
                float z2 = acc[ft][rt][4 * g + 2] + b2, z3 = acc[ft][rt][4 * g + 3] + b3;
                uint2 pk;
                pk.x = bf16_bits(z0 * sigm(z0)) | (bf16_bits(z1 * sigm(z1)) << 16);
                pk.y = bf16_bits(z2 * sigm(z2)) | (bf16_bits(z3 * sigm(z3)) << 16);
                *(uint2*)&H[rbase + rt * 32 + (lane & 31)][m] = pk;
            }
        }
}
// ---------------------------------------------------------------------------
// Fused inference for the warm-start shape (16 -> 4 x 256 SiLU -> 3): a block
// takes FI_ROWS rows through all five layers; the activations stay in LDS as
// BF16 (the products' input precision) and never touch HBM.  Weights are read
// from L2 as BF16 transposed per layer, Wt[n][k] (the last layer padded to 32
// rows of zeros).  In the hidden layers wave w owns output features
// 64 (w % 4) .. +64 (two 32-feature MFMA tiles) of its FI_RT 32-row tiles:
// each weight fragment feeds FI_RT MFMAs, each activation fragment two.  The
// output layer's 32-row tiles are shared out over all the block's waves.
// The weight fragments come from L1/L2, so rows per wave set the loads per
// product: 200k rows take 199 us with 2 row tiles (108 VGPRs, 4 waves per
// SIMD), 159 us with 3, 151 us with 4 (192 VGPRs, 2 waves per SIMD, two
// 128-row blocks per CU), all with the same bits; 5 or more row tiles fit only
// one wave per SIMD.  The accumulation order over k is the generic GEMM's.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(FI_THREADS) __attribute__((amdgpu_waves_per_eu(FI_RT <= 2 ? 4 : 2)))
k_mlp_infer(int n, const float* __restrict__ X, const __bf16* __restrict__ Wt, const float* __restrict__ params,
            const long* __restrict__ boff, float* __restrict__ Y)
{
    __shared__ __attribute__((aligned(16))) __bf16 H[FI_ROWS][FI_LDS];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int f0 = (w & 3) * 64, rbase = (w >> 2) * FI_WROWS;
    const long r0 = (long)blockIdx.x * FI_ROWS;
    // input tile: FI_ROWS rows x 16 features
    for (int e = tid; e < FI_ROWS * 16; e += blockDim.x) {
        int r = e >> 4, k = e & 15;
        long gr = r0 + r;
        H[r][k] = (__bf16)(gr < n ? X[gr * 16 + k] : 0.f);
    }
    __syncthreads();
    const __bf16* W = Wt;
    for (int l = 0; l < 4; l++) {
        floatx16 acc[2][FI_RT];
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int rt = 0; rt < FI_RT; rt++)
#pragma unroll
                for (int j = 0; j < 16; j++) acc[f][rt][j] = 0.f;
        if (l == 0) fi_layer<16, 2, FI_RT>(W, H, f0, rbase, lane, acc);
        else fi_layer<FI_W, 2, FI_RT>(W, H, f0, rbase, lane, acc);
        __syncthreads();  // every wave has read this layer's input
        fi_silu_store<FI_RT>(acc, params + boff[l], H, f0, rbase, lane);
        __syncthreads();
        W += (long)FI_W * (l == 0 ? 16 : FI_W);
    }
    // output layer: features 0..2 of the first 32-feature tile, one 32-row
    // tile per wave at a time
    const float* bias = params + boff[4];
    for (int t = w; t < FI_ROWS / 32; t += FI_THREADS / 64) {
        floatx16 acc[1][1];
#pragma unroll
        for (int j = 0; j < 16; j++) acc[0][0][j] = 0.f;
        fi_layer<FI_W, 1, 1>(W, H, 0, t * 32, lane, acc);
        const long gr = r0 + t * 32 + lane;
        if (lane < 32 && gr < n)
#pragma unroll
            for (int i = 0; i < 3; i++) Y[gr * 3 + i] = acc[0][0][i] + bias[i];
    }
}
// Wt (BF16, transposed, last layer padded to 32 rows) from the FP32 parameters
__global__ void k_pack_wt(int fi, int fo, int rows, const float* __restrict__ W, __bf16* __restrict__ Wt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * fi) return;
    int nrow = i / fi, k = i % fi;
    Wt[i] = (__bf16)(nrow < fo ? W[(long)k * fo + nrow] : 0.f);
}

// ---------------------------------------------------------------------------
// the network object
// ---------------------------------------------------------------------------
struct fipc_mlp {
    int n_in = 0, width = 0, depth = 0, n_out = 0, L = 0;
    std::vector<int> dim;          // dim[0] = n_in, ..., dim[L] = n_out
    std::vector<long> woff, boff;  // offsets of W_l and b_l in the packed parameters
    long nparams = 0;
    DVec<float> params, grads, am, av;
    long adam_t = 0;
    int cap = 0;                   // rows the layer outputs H hold
    int tcap = 0;                  // rows the training buffers (Z, gradients, X, T) hold
    std::vector<DVec<float>> H, Z; // H[l] (l >= 1) layer outputs, Z[l] hidden pre-activations
    DVec<float> X, T, dA, dB, part, loss;
    // fused inference (warm-start shape only): packed BF16 weights, stale
    // after any parameter change
    bool fused = false;
    bool wt_ok = false;
    DVec<__bf16> wt;
    DVec<long> dboff;
    // dataset
    DVec<float> dsX, dsT;
    long ds_n = 0;
    hipStream_t s = nullptr;
};

static int mlp_reserve(fipc_mlp* m, int rows, bool train)
{
    if (rows > m->cap) {
        for (int l = 1; l <= m->L; l++) STCHK(ensure(m->H[l], (size_t)rows * m->dim[l]));
        m->cap = rows;
    }
    if (train && rows > m->tcap) {
        for (int l = 1; l < m->L; l++) STCHK(ensure(m->Z[l], (size_t)rows * m->dim[l]));
        int wmax = *std::max_element(m->dim.begin(), m->dim.end());
        STCHK(ensure(m->dA, (size_t)rows * wmax));
        STCHK(ensure(m->dB, (size_t)rows * wmax));
        STCHK(ensure(m->X, (size_t)rows * m->n_in));
        STCHK(ensure(m->T, (size_t)rows * m->n_out));
        STCHK(ensure(m->loss, (size_t)nblk((long)rows * m->n_out) + 8));
        m->tcap = rows;
    }
    return FIPC_OK;
}

// forward of n rows from device X; Y (device, n x n_out) in the network's buffer.
// With train != 0 the hidden pre-activations are kept for the backward pass.
static int mlp_forward(fipc_mlp* m, hipStream_t s, int n, const float* X, bool train)
{
    STCHK(mlp_reserve(m, n, train));
    const float* in = X;
    for (int l = 0; l < m->L; l++) {
        int fi = m->dim[l], fo = m->dim[l + 1];
        const float* W = m->params.p + m->woff[l];
        const float* b = m->params.p + m->boff[l];
        bool last = l == m->L - 1;
        STCHK(gemm(s, n, fo, fi, in, fi, 1, W, fo, 1, m->H[l + 1].p, fo, last ? EPI_BIAS : EPI_BIAS_SILU, b, nullptr,
                   (!last && train) ? m->Z[l + 1].p : nullptr));
        in = m->H[l + 1].p;
    }
    return FIPC_OK;
}

// gradient of the mean squared error over n rows (X, T on the device) into
// m->grads; the loss partials are left in m->loss (nblk(n n_out) floats)
static int mlp_backward(fipc_mlp* m, hipStream_t s, int n, const float* X, const float* T)
{
    STCHK(mlp_forward(m, s, n, X, true));
    long ny = (long)n * m->n_out;
    float* dZ = m->dA.p;   // gradient at the current layer's pre-activation
    float* dH = m->dB.p;
    k_mse<<<nblk(ny), TPB, 0, s>>>(ny, 1.f / (float)ny, m->H[m->L].p, T, dZ, m->loss.p);
    HIPCHK(hipGetLastError());
    // split-K partials of dW: slices of at most 1024 rows
    int S = std::max(1, std::min(32, (n + 1023) / 1024));
    for (int l = m->L - 1; l >= 0; l--) {
        int fi = m->dim[l], fo = m->dim[l + 1];
        const float* Hin = l == 0 ? X : m->H[l].p;
        // dW_l = H_in^T dZ  (fi x fo), K = n rows
        STCHK(ensure(m->part, (size_t)S * fi * fo));
        STCHK(gemm(s, fi, fo, n, Hin, 1, fi, dZ, fo, 1, m->part.p, fo, EPI_NONE, nullptr, nullptr, nullptr, S,
                   (long)fi * fo));
        k_sum_slices<<<nblk((long)fi * fo), TPB, 0, s>>>((long)fi * fo, S, m->part.p, m->grads.p + m->woff[l]);
        k_colsum<<<(fo + 63) / 64, 256, 0, s>>>(n, fo, dZ, m->grads.p + m->boff[l]);
        HIPCHK(hipGetLastError());
        if (l > 0) {
            // dZ_{l} = (dZ W_l^T) * silu'(Z_l)
            const float* W = m->params.p + m->woff[l];
            STCHK(gemm(s, n, fi, fo, dZ, fo, 1, W, 1, fo, dH, fi, EPI_DSILU, nullptr, m->Z[l].p, nullptr));
            std::swap(dZ, dH);
        }
    }
    return FIPC_OK;
}

int mlp_dims(const fipc_mlp* m, int* n_in, int* n_out)
{
    *n_in = m->n_in;
    *n_out = m->n_out;
    return FIPC_OK;
}
static int mlp_pack(fipc_mlp* m, hipStream_t s)
{
    if (m->wt_ok) return FIPC_OK;
    long total = 0;
    for (int l = 0; l < m->L; l++) total += (long)(l == m->L - 1 ? 32 : m->dim[l + 1]) * m->dim[l];
    STCHK(ensure(m->wt, total));
    long o = 0;
    for (int l = 0; l < m->L; l++) {
        int fi = m->dim[l], fo = m->dim[l + 1], rows = l == m->L - 1 ? 32 : fo;
        k_pack_wt<<<nblk((long)rows * fi), TPB, 0, s>>>(fi, fo, rows, m->params.p + m->woff[l], m->wt.p + o);
        o += (long)rows * fi;
    }
    HIPCHK(hipGetLastError());
    m->wt_ok = true;
    return FIPC_OK;
}
// forward of n rows of device X on stream s: the fused kernel for the
// warm-start shape, the layer-by-layer GEMMs otherwise
static int mlp_infer(fipc_mlp* m, hipStream_t s, int n, const float* X, const float** Y)
{
    if (!m->fused) {
        STCHK(mlp_forward(m, s, n, X, false));
        *Y = m->H[m->L].p;
        return FIPC_OK;
    }
    STCHK(mlp_pack(m, s));
    STCHK(ensure(m->H[m->L], (size_t)n * m->n_out));
    k_mlp_infer<<<(n + FI_ROWS - 1) / FI_ROWS, FI_THREADS, 0, s>>>(n, X, m->wt.p, m->params.p, m->dboff.p,
                                                                  m->H[m->L].p);
    HIPCHK(hipGetLastError());
    *Y = m->H[m->L].p;
    return FIPC_OK;
}
int mlp_forward_dev(fipc_mlp* m, hipStream_t s, int n, const float* X, const float** Y)
{
    return mlp_infer(m, s, n, X, Y);
}
int mlp_dataset_append_dev(fipc_mlp* m, hipStream_t s, int n, const float* X, const float* T)
{
    if (n <= 0) return FIPC_OK;
    long need = m->ds_n + n;
    if ((size_t)need * m->n_in > m->dsX.cap) {
        // grow, keeping the rows already captured
        DVec<float> nx, nt;
        STCHK(ensure(nx, (size_t)need * 2 * m->n_in));
        STCHK(ensure(nt, (size_t)need * 2 * m->n_out));
        if (m->ds_n) {
            HIPCHK(hipMemcpyAsync(nx.p, m->dsX.p, (size_t)m->ds_n * m->n_in * 4, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(nt.p, m->dsT.p, (size_t)m->ds_n * m->n_out * 4, hipMemcpyDeviceToDevice, s));
        }
        HIPCHK(hipStreamSynchronize(s));
        release(m->dsX);
        release(m->dsT);
        m->dsX = nx;
        m->dsT = nt;
    }
    HIPCHK(hipMemcpyAsync(m->dsX.p + m->ds_n * m->n_in, X, (size_t)n * m->n_in * 4, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(m->dsT.p + m->ds_n * m->n_out, T, (size_t)n * m->n_out * 4, hipMemcpyDeviceToDevice, s));
    m->ds_n = need;
    return FIPC_OK;
}
hipStream_t mlp_stream(fipc_mlp* m) { return m->s; }

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" int fipc_mlp_create(fipc_mlp** out, int n_in, int width, int depth, int n_out, unsigned long long seed)
{
    if (!out || n_in <= 0 || width <= 0 || depth < 0 || n_out <= 0) return FIPC_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FIPC_ERR_HIP;
    fipc_mlp* m = new fipc_mlp();
    m->n_in = n_in; m->width = width; m->depth = depth; m->n_out = n_out;
    m->L = depth + 1;
    m->dim.push_back(n_in);
    for (int i = 0; i < depth; i++) m->dim.push_back(width);
    m->dim.push_back(n_out);
    long off = 0;
    for (int l = 0; l < m->L; l++) {
        m->woff.push_back(off);
        off += (long)m->dim[l] * m->dim[l + 1];
        m->boff.push_back(off);
        off += m->dim[l + 1];
    }
    m->nparams = off;
    m->H.resize(m->L + 1);
    m->Z.resize(m->L + 1);
    if (hipStreamCreate(&m->s) != hipSuccess) { delete m; return FIPC_ERR_HIP; }
    int st = FIPC_OK;
    if (!st) st = ensure(m->params, off);
    if (!st) st = ensure(m->grads, off);
    if (!st) st = ensure(m->am, off);
    if (!st) st = ensure(m->av, off);
    if (st) { fipc_mlp_destroy(m); return st; }
    for (int l = 0; l < m->L; l++) {
        long nw = (long)m->dim[l] * m->dim[l + 1];
        // variance 1 / fan_in; the output layer starts 10x smaller so an
        // untrained network predicts almost no correction
        float a = sqrtf(3.f / (float)m->dim[l]) * (l == m->L - 1 ? 0.1f : 1.f);
        k_init_layer<<<nblk(std::max<long>(nw, m->dim[l + 1])), TPB, 0, m->s>>>(
            nw, m->dim[l + 1], m->params.p + m->woff[l], m->params.p + m->boff[l], a, seed + 7919ull * l);
    }
    if (hipGetLastError() != hipSuccess) { fipc_mlp_destroy(m); return FIPC_ERR_HIP; }
    m->fused = n_in == 16 && width == FI_W && depth == 4 && n_out == 3;
    if (m->fused) {
        if (ensure(m->dboff, m->L) != FIPC_OK) { fipc_mlp_destroy(m); return FIPC_ERR_HIP; }
        (void)hipMemcpyAsync(m->dboff.p, m->boff.data(), m->L * sizeof(long), hipMemcpyHostToDevice, m->s);
    }
    (void)hipMemsetAsync(m->am.p, 0, off * 4, m->s);
    (void)hipMemsetAsync(m->av.p, 0, off * 4, m->s);
    if (hipStreamSynchronize(m->s) != hipSuccess) { fipc_mlp_destroy(m); return FIPC_ERR_HIP; }
    *out = m;
    return FIPC_OK;
}

extern "C" int fipc_mlp_destroy(fipc_mlp* m)
{
    if (!m) return FIPC_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    release(m->params); release(m->grads); release(m->am); release(m->av);
    for (auto& v : m->H) release(v);
    for (auto& v : m->Z) release(v);
    release(m->X); release(m->T); release(m->dA); release(m->dB); release(m->part); release(m->loss);
    release(m->dsX); release(m->dsT); release(m->wt); release(m->dboff);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return FIPC_OK;
}

extern "C" int fipc_mlp_nparams(fipc_mlp* m, long* n)
{
    if (!m || !n) return FIPC_ERR_ARG;
    *n = m->nparams;
    return FIPC_OK;
}
extern "C" int fipc_mlp_get_params(fipc_mlp* m, float* w)
{
    if (!m || !w) return FIPC_ERR_ARG;
    HIPCHK(hipMemcpyAsync(w, m->params.p, m->nparams * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    return FIPC_OK;
}
extern "C" int fipc_mlp_set_params(fipc_mlp* m, const float* w)
{
    if (!m || !w) return FIPC_ERR_ARG;
    HIPCHK(hipMemcpyAsync(m->params.p, w, m->nparams * 4, hipMemcpyHostToDevice, m->s));
    m->wt_ok = false;
    HIPCHK(hipMemsetAsync(m->am.p, 0, m->nparams * 4, m->s));
    HIPCHK(hipMemsetAsync(m->av.p, 0, m->nparams * 4, m->s));
    m->adam_t = 0;
    HIPCHK(hipStreamSynchronize(m->s));
    return FIPC_OK;
}

extern "C" int fipc_mlp_forward(fipc_mlp* m, int n, const float* X, float* Y)
{
    if (!m || n < 0 || (n && (!X || !Y))) return FIPC_ERR_ARG;
    if (!n) return FIPC_OK;
    STCHK(mlp_reserve(m, n, true));
    HIPCHK(hipMemcpyAsync(m->X.p, X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->s));
    const float* Yd = nullptr;
    STCHK(mlp_infer(m, m->s, n, m->X.p, &Yd));
    HIPCHK(hipMemcpyAsync(Y, Yd, (size_t)n * m->n_out * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    return FIPC_OK;
}

extern "C" int fipc_mlp_grad(fipc_mlp* m, int n, const float* X, const float* T, float* loss, float* grad)
{
    if (!m || n <= 0 || !X || !T) return FIPC_ERR_ARG;
    STCHK(mlp_reserve(m, n, true));
    HIPCHK(hipMemcpyAsync(m->X.p, X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(m->T.p, T, (size_t)n * m->n_out * 4, hipMemcpyHostToDevice, m->s));
    STCHK(mlp_backward(m, m->s, n, m->X.p, m->T.p));
    int nb = nblk((long)n * m->n_out);
    std::vector<float> lp(nb);
    HIPCHK(hipMemcpyAsync(lp.data(), m->loss.p, nb * 4, hipMemcpyDeviceToHost, m->s));
    if (grad) HIPCHK(hipMemcpyAsync(grad, m->grads.p, m->nparams * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    double l = 0.0;
    for (float v : lp) l += v;
    if (loss) *loss = (float)l;
    return FIPC_OK;
}

extern "C" int fipc_mlp_dataset_clear(fipc_mlp* m)
{
    if (!m) return FIPC_ERR_ARG;
    m->ds_n = 0;
    return FIPC_OK;
}
extern "C" int fipc_mlp_dataset_size(fipc_mlp* m, long* n)
{
    if (!m || !n) return FIPC_ERR_ARG;
    *n = m->ds_n;
    return FIPC_OK;
}
extern "C" int fipc_mlp_dataset_add(fipc_mlp* m, int n, const float* X, const float* T)
{
    if (!m || n < 0 || (n && (!X || !T))) return FIPC_ERR_ARG;
    if (!n) return FIPC_OK;
    DVec<float> dx, dt;
    STCHK(ensure(dx, (size_t)n * m->n_in));
    STCHK(ensure(dt, (size_t)n * m->n_out));
    HIPCHK(hipMemcpyAsync(dx.p, X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(dt.p, T, (size_t)n * m->n_out * 4, hipMemcpyHostToDevice, m->s));
    int st = mlp_dataset_append_dev(m, m->s, n, dx.p, dt.p);
    (void)hipStreamSynchronize(m->s);
    release(dx);
    release(dt);
    return st;
}
extern "C" int fipc_mlp_dataset_get(fipc_mlp* m, long first, long n, float* X, float* T)
{
    if (!m || first < 0 || n < 0 || first + n > m->ds_n) return FIPC_ERR_ARG;
    if (!n) return FIPC_OK;
    if (X) HIPCHK(hipMemcpyAsync(X, m->dsX.p + first * m->n_in, (size_t)n * m->n_in * 4, hipMemcpyDeviceToHost, m->s));
    if (T) HIPCHK(hipMemcpyAsync(T, m->dsT.p + first * m->n_out, (size_t)n * m->n_out * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    return FIPC_OK;
}

// iters Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) on batches drawn from
// the dataset; losses[i] = mean squared error of batch i before its update
extern "C" int fipc_mlp_train(fipc_mlp* m, int iters, int batch, float lr, unsigned long long seed, float* losses)
{
    if (!m || iters < 0 || batch <= 0 || m->ds_n == 0) return FIPC_ERR_ARG;
    STCHK(mlp_reserve(m, batch, true));
    int nb = nblk((long)batch * m->n_out);
    DVec<float> lossall;
    STCHK(ensure(lossall, (size_t)iters * nb + 1));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < iters; it++) {
        k_gather_batch<<<nblk(batch), TPB, 0, m->s>>>(batch, m->ds_n, seed + 1000003ull * (unsigned long long)it,
                                                      m->n_in, m->n_out, m->dsX.p, m->dsT.p, m->X.p, m->T.p);
        HIPCHK(hipGetLastError());
        STCHK(mlp_backward(m, m->s, batch, m->X.p, m->T.p));
        HIPCHK(hipMemcpyAsync(lossall.p + (size_t)it * nb, m->loss.p, nb * 4, hipMemcpyDeviceToDevice, m->s));
        m->adam_t++;
        float c1 = 1.f - powf(b1, (float)m->adam_t), c2 = 1.f - powf(b2, (float)m->adam_t);
        k_adam<<<nblk(m->nparams), TPB, 0, m->s>>>(m->nparams, m->params.p, m->grads.p, m->am.p, m->av.p, lr, b1, b2,
                                                    eps, c1, c2);
        m->wt_ok = false;
        HIPCHK(hipGetLastError());
    }
    std::vector<float> lp((size_t)iters * nb);
    if (iters) HIPCHK(hipMemcpyAsync(lp.data(), lossall.p, lp.size() * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    release(lossall);
    if (losses)
        for (int it = 0; it < iters; it++) {
            double l = 0.0;
            for (int b = 0; b < nb; b++) l += lp[(size_t)it * nb + b];
            losses[it] = (float)l;
        }
    return FIPC_OK;
}
