// fipc_noalloc.hip -- an entry point that allocates no device memory.
#include <hip/hip_runtime.h>
#include <mutex>
#include "fipc_ctx.h"

#define TPB 256
static inline int nblk(long n, int t = TPB) { return (int)((n + t - 1) / t); }

// the batched CCD kernel of fipc_broad.hip (one thread per pair)
__global__ void k_batch_ccd(int type, int n, const double* pts, const double* disp, double eta, const double* d2,
                            double* toc);

// ---- the same batch without device allocation -------------------------------
// For callers that cannot call hipMalloc (graph capture, serving loops): the
// device copies of the inputs and the output live in a file-scope device
// array, scratch, of FIPC_CCD_NOALLOC_MAX pairs (26 doubles each: 12 points,
// 12 displacements, the squared distance and the result; 3.25 MiB per GPU,
// held until the process ends).  Every call on one GPU uses that GPU's copy:
// calls that overlap on one GPU race.  Nothing in it needs clearing (each
// call overwrites what it reads).
#define FIPC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[FIPC_CCD_NOALLOC_MAX * 26];
static std::mutex g_scratch_mu;
static double* g_scratch_dev[FIPC_MAX_DEVICES];
// the device address of scratch on the current device, looked up once per
// device (the array's name is a host-side stand-in, never a device address)
static int scratch_of_current_device(double** out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= FIPC_MAX_DEVICES) return FIPC_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (double*)ptr;
    }
    *out = g_scratch_dev[dev];
    return FIPC_OK;
}
extern "C" int fipc_batch_ccd_noalloc(int type, int n, const double* pts, const double* disp, double eta,
                                      const double* dist2, double* toc)
{
    if (n < 0 || (type != 0 && type != 1)) return FIPC_ERR_ARG;
    if (n > FIPC_CCD_NOALLOC_MAX) return FIPC_ERR_CAPACITY;
    if (n == 0) return FIPC_OK;
    if (!pts || !disp || !dist2 || !toc) return FIPC_ERR_ARG;
    double* sp = nullptr;
    STCHK(scratch_of_current_device(&sp));
    double *dp = sp, *dd = sp + (size_t)12 * n, *d2 = sp + (size_t)24 * n, *dt = sp + (size_t)25 * n;
    HIPCHK(hipMemcpy(dp, pts, (size_t)n * 96, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dd, disp, (size_t)n * 96, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d2, dist2, (size_t)n * 8, hipMemcpyHostToDevice));
    k_batch_ccd<<<nblk(n), TPB>>>(type, n, dp, dd, eta, d2, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(toc, dt, (size_t)n * 8, hipMemcpyDeviceToHost));
    return FIPC_OK;
}
