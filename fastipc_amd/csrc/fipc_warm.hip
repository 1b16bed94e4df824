// fipc_warm.hip -- the learned Newton warm start on a context: per-vertex
// features from the step state, the predicted increment written as the
// solution vector (so that the caller filters it with the CCD step bound and
// applies it like a Newton direction), and capture of training rows from the
// plain solver's trajectories.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <vector>
#include "fipc_ctx.h"
#include "fipc_rows.h"
#include "fipc_mlp.h"

using namespace fipc;

#define TPB 256
#define WS_NF 16
static inline int nblk(long n, int t = TPB) { return (int)((n + t - 1) / t); }

struct RowSet {
    const int* rows[NTYPES];
    int n[NTYPES];
};

// per vertex: the smallest (squared distance, row) key over the contact rows
// it belongs to.  Key = float bits of d^2 (order-preserving for d^2 >= 0)
// << 32 | kind << 28 | row.
// (Dirichlet vertices are skipped: their start is prescribed, and the rows of
// a moving plate would otherwise all contend for its few vertices)
template <int KIND>
__global__ void k_ws_keys(int n, const int* __restrict__ rows, const double* __restrict__ x,
                          const unsigned char* __restrict__ dmask, unsigned long long* keys)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    constexpr int K = KindInfo<KIND>::K;
    int id[4];
    v3 px[4], p0[4];
    load_row<KIND>(rows, r, x, x, id, px, p0);
    float d2 = (float)constraint_dist(KIND, px);
    unsigned long long key = ((unsigned long long)__float_as_uint(d2) << 32) | ((unsigned)KIND << 28) | (unsigned)r;
#pragma unroll
    for (int a = 0; a < K; a++)
        if (!dmask[id[a]]) atomicMin(&keys[id[a]], key);
}

// the two primitives of a row: PP {0}|{1}, PE {0}|{1,2}, PT {0}|{1,2,3},
// edge-edge kinds {0,1}|{2,3}
__device__ __forceinline__ int side_split(int kind) { return kind <= 2 ? 1 : 2; }

__global__ void k_ws_features(int nv, const double* __restrict__ x, const double* __restrict__ xt,
                              const double* __restrict__ v, const double* __restrict__ dprev,
                              const unsigned char* __restrict__ dmask, const unsigned long long* __restrict__ keys,
                              RowSet rs, double dt, double grav, double inv_s, double inv_dhat,
                              float* __restrict__ feat)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    float* f = feat + (size_t)i * WS_NF;
    v3 vd = mk(v[3 * i] * dt * inv_s, v[3 * i + 1] * dt * inv_s, v[3 * i + 2] * dt * inv_s);
    v3 dp = mk(dprev[3 * i] * inv_s, dprev[3 * i + 1] * inv_s, dprev[3 * i + 2] * inv_s);
    f[0] = (float)vd.x; f[1] = (float)vd.y; f[2] = (float)vd.z;
    f[3] = (float)dp.x; f[4] = (float)dp.y; f[5] = (float)dp.z;
    f[6] = 0.f; f[7] = (float)(grav * dt * dt * inv_s); f[8] = 0.f;
    float dist = 1.f, nx = 0.f, ny = 0.f, nz = 0.f, rel_n = 0.f, rel_t = 0.f;
    unsigned long long key = keys[i];
    if (key != ~0ull) {
        float d2 = __uint_as_float((unsigned)(key >> 32));
        int kind = (int)((key >> 28) & 7u), r = (int)(key & 0x0fffffffu);
        int K = kind_points(kind), W = kind == 0 ? 2 : (kind == 1 ? 3 : 4), sp = side_split(kind);
        const int* row = rs.rows[kind] + (size_t)W * r;
        int pos = 0;
        for (int a = 0; a < K; a++)
            if (row[a] == i) pos = a;
        int lo = pos < sp ? sp : 0, hi = pos < sp ? K : sp;
        // the other primitive: centroid, and mean inertial (or prescribed) increment
        v3 c = mk(0, 0, 0), u = mk(0, 0, 0);
        for (int a = lo; a < hi; a++) {
            c = c + ldv(x, row[a]);
            u = u + (ldv(xt, row[a]) - ldv(x, row[a]));
        }
        double w = 1.0 / (hi - lo);
        v3 d = ldv(x, i) - mk(c.x * w, c.y * w, c.z * w);
        double len = sqrt(n2(d));
        v3 n = len > 0 ? mk(d.x / len, d.y / len, d.z / len) : mk(0, 0, 0);
        nx = (float)n.x; ny = (float)n.y; nz = (float)n.z;
        dist = fminf(sqrtf(d2) * (float)inv_dhat, 1.f);
        // partner's increment relative to this vertex's inertial increment:
        // the component along n (approach when negative) and the tangential size
        v3 rel = mk(u.x * w, u.y * w, u.z * w) - (ldv(xt, i) - ldv(x, i));
        rel = mk(rel.x * inv_s, rel.y * inv_s, rel.z * inv_s);
        double rn = rel.x * n.x + rel.y * n.y + rel.z * n.z;
        v3 rt = rel - mk(rn * n.x, rn * n.y, rn * n.z);
        rel_n = (float)(-rn);
        rel_t = (float)sqrt(n2(rt));
    }
    f[9] = dist; f[10] = nx; f[11] = ny; f[12] = nz;
    f[13] = dmask[i] ? 1.f : 0.f;
    f[14] = rel_n;
    f[15] = rel_t;
}

// p = xTilde - x + dprev + s y (free vertices): the inertial increment, the
// previous step's correction repeated, and the network's change of it;
// Dirichlet vertices: their prescribed increment xTilde - x
__global__ void k_ws_apply(int nv, const float* __restrict__ Y, int ldy, double s, const double* __restrict__ x,
                           const double* __restrict__ xt, const double* __restrict__ dprev,
                           const unsigned char* __restrict__ dmask, double* p)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    bool fixed = dmask[i];
    for (int d = 0; d < 3; d++) {
        double inc = xt[3 * i + d] - x[3 * i + d];
        p[3 * i + d] = fixed ? inc : inc + dprev[3 * i + d] + s * (double)Y[(size_t)i * ldy + d];
    }
}

// free vertices within d-hat of a contact primitive (feature 9 = dist / d-hat
// below 1), appended to near_ids; *nnear counts them
__global__ void k_ws_near(int nfree, const int* __restrict__ free_ids, const float* __restrict__ feat, int* near_ids,
                          int* nnear)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool near = false;
    int v = 0;
    if (k < nfree) {
        v = free_ids[k];
        near = feat[(size_t)v * WS_NF + 9] < 1.f;
    }
    // one counter update per wavefront
    unsigned long long mask = __ballot(near);
    if (!mask) return;
    int lane = __lane_id(), leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(nnear, __popcll(mask));
    base = __shfl(base, leader, 64);
    if (near) near_ids[base + __popcll(mask & ((1ull << lane) - 1))] = v;
}
// sample k < nnear_samples from the near vertices (when there are any), the
// others from all free vertices, uniformly with replacement
__global__ void k_ws_pick(int n, int nfree, const int* __restrict__ free_ids, int nnear_samples,
                          const int* __restrict__ near_ids, const int* __restrict__ nnear, unsigned long long seed,
                          const float* __restrict__ feat, int* ids, float* X)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    unsigned long long z = seed * 0x9e3779b97f4a7c15ull + (unsigned long long)k + 0x632be59bd9b4e019ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    int nn = nnear_samples > 0 ? *nnear : 0;
    int v = (k < nnear_samples && nn > 0) ? near_ids[z % (unsigned long long)nn]
                                          : free_ids[z % (unsigned long long)nfree];
    ids[k] = v;
    for (int j = 0; j < WS_NF; j++) X[(size_t)k * WS_NF + j] = feat[(size_t)v * WS_NF + j];
}
// target: the change of the correction, ((x_{n+1} - xTilde) - dprev) / s;
// dprev still holds the previous step's correction (end_step replaces it)
__global__ void k_ws_target(int n, const int* __restrict__ ids, const double* __restrict__ x,
                            const double* __restrict__ xt, const double* __restrict__ dprev, double inv_s, float* T)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int v = ids[k];
    for (int d = 0; d < 3; d++)
        T[(size_t)k * 3 + d] = (float)((x[3 * v + d] - xt[3 * v + d] - dprev[3 * v + d]) * inv_s);
}

static int ws_features(fipc_ctx* c, double scale)
{
    if (!(scale > 0)) return FIPC_ERR_ARG;
    int nv = c->nv;
    STCHK(ensure(c->wsX, (size_t)nv * WS_NF));
    STCHK(ensure(c->ckey, nv));
    HIPCHK(hipMemsetAsync(c->ckey.p, 0xff, (size_t)nv * 8, c->s));
    RowSet rs;
    for (int t = 0; t < NTYPES; t++) {
        rs.rows[t] = c->rows[t].p;
        rs.n[t] = c->count[t];
        int n = c->count[t];
        if (!n) continue;
#define LAUNCH_K(K) k_ws_keys<K><<<nblk(n), TPB, 0, c->s>>>(n, c->rows[t].p, c->x.p, c->dmask.p, c->ckey.p)
        switch (t) {
        case 0: LAUNCH_K(0); break;
        case 1: LAUNCH_K(1); break;
        case 2: LAUNCH_K(2); break;
        case 3: LAUNCH_K(3); break;
        case 4: LAUNCH_K(4); break;
        case 5: LAUNCH_K(5); break;
        default: LAUNCH_K(6); break;
        }
#undef LAUNCH_K
    }
    k_ws_features<<<nblk(nv), TPB, 0, c->s>>>(nv, c->x.p, c->xt.p, c->v.p, c->dprev.p, c->dmask.p, c->ckey.p, rs,
                                              c->dt, c->gravity, 1.0 / scale, 1.0 / c->dHat, c->wsX.p);
    HIPCHK(hipGetLastError());
    return FIPC_OK;
}

extern "C" int fipc_ws_features(fipc_ctx* c, double scale, float* feat)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    STCHK(ws_features(c, scale));
    if (feat) HIPCHK(hipMemcpyAsync(feat, c->wsX.p, (size_t)c->nv * WS_NF * 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return FIPC_OK;
}

static int ws_check_net(fipc_mlp* m)
{
    int ni = 0, no = 0;
    mlp_dims(m, &ni, &no);
    return ni == WS_NF && no == 3 ? FIPC_OK : FIPC_ERR_ARG;
}

extern "C" int fipc_ws_predict(fipc_ctx* c, fipc_mlp* m, double scale)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    if (!m) return FIPC_ERR_ARG;
    STCHK(ws_check_net(m));
    HIPCHK(hipStreamSynchronize(mlp_stream(m)));
    STCHK(ws_features(c, scale));
    const float* Y = nullptr;
    STCHK(mlp_forward_dev(m, c->s, c->nv, c->wsX.p, &Y));
    k_ws_apply<<<nblk(c->nv), TPB, 0, c->s>>>(c->nv, Y, 3, scale, c->x.p, c->xt.p, c->dprev.p, c->dmask.p, c->sol.p);
    c->restart_ready = false;  // sol changed
    c->sol_version++;
    HIPCHK(hipGetLastError());
    // no host wait: later work on the network's own stream (training,
    // parameter changes, destruction) waits for this stream's use of its
    // buffers instead
    if (!c->ws_done) HIPCHK(hipEventCreateWithFlags(&c->ws_done, hipEventDisableTiming));
    HIPCHK(hipEventRecord(c->ws_done, c->s));
    HIPCHK(hipStreamWaitEvent(mlp_stream(m), c->ws_done, 0));
    return FIPC_OK;
}

extern "C" int fipc_ws_capture_begin(fipc_ctx* c, fipc_mlp* m, double scale, int nsample, unsigned long long seed,
                                     double near_frac)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    if (!m || nsample <= 0 || !(near_frac >= 0.0 && near_frac <= 1.0)) return FIPC_ERR_ARG;
    STCHK(ws_check_net(m));
    if (c->nfree < 0) {
        // free vertices, from the Dirichlet mask
        std::vector<unsigned char> mask(c->nv);
        HIPCHK(hipMemcpyAsync(mask.data(), c->dmask.p, c->nv, hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        std::vector<int> ids;
        for (int i = 0; i < c->nv; i++)
            if (!mask[i]) ids.push_back(i);
        c->nfree = (int)ids.size();
        if (c->nfree) {
            STCHK(ensure(c->free_ids, ids.size()));
            HIPCHK(hipMemcpyAsync(c->free_ids.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, c->s));
        }
    }
    c->cap_n = 0;
    if (c->nfree == 0) return FIPC_OK;
    STCHK(ws_features(c, scale));
    STCHK(ensure(c->cap_ids, nsample));
    STCHK(ensure(c->cap_X, (size_t)nsample * WS_NF));
    STCHK(ensure(c->cap_T, (size_t)nsample * 3));
    int nnear_samples = (int)(near_frac * nsample);
    if (nnear_samples > 0) {
        STCHK(ensure(c->near_ids, c->nfree));
        STCHK(ensure(c->nnear, 1));
        HIPCHK(hipMemsetAsync(c->nnear.p, 0, 4, c->s));
        k_ws_near<<<nblk(c->nfree), TPB, 0, c->s>>>(c->nfree, c->free_ids.p, c->wsX.p, c->near_ids.p, c->nnear.p);
    }
    k_ws_pick<<<nblk(nsample), TPB, 0, c->s>>>(nsample, c->nfree, c->free_ids.p, nnear_samples, c->near_ids.p,
                                               c->nnear.p, seed, c->wsX.p, c->cap_ids.p, c->cap_X.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    c->cap_n = nsample;
    return FIPC_OK;
}

extern "C" int fipc_ws_capture_end(fipc_ctx* c, fipc_mlp* m, double scale)
{
    if (!c || !c->nv) return FIPC_ERR_STATE;
    if (!m || !(scale > 0)) return FIPC_ERR_ARG;
    STCHK(ws_check_net(m));
    if (!c->cap_n) return FIPC_OK;
    HIPCHK(hipStreamSynchronize(mlp_stream(m)));
    k_ws_target<<<nblk(c->cap_n), TPB, 0, c->s>>>(c->cap_n, c->cap_ids.p, c->x.p, c->xt.p, c->dprev.p, 1.0 / scale,
                                                   c->cap_T.p);
    HIPCHK(hipGetLastError());
    STCHK(mlp_dataset_append_dev(m, c->s, c->cap_n, c->cap_X.p, c->cap_T.p));
    HIPCHK(hipStreamSynchronize(c->s));
    c->cap_n = 0;
    return FIPC_OK;
}
