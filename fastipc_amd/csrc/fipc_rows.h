// fipc_rows.h -- device helpers shared by the kernels that read contact rows:
// vertex loads and the per-kind row geometry (row width W, points K, dofs N,
// lanes per row G for the projected-Hessian kernels).
#pragma once
#include "fipc_math.h"

namespace fipc {

__device__ __forceinline__ v3 ldv(const double* a, int i) { return v3{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

template <int KIND>
struct KindInfo {
    static constexpr int K = kind_points(KIND);
    static constexpr int N = 3 * K;
    // lanes per row of the contact assembly: one per Hessian row, rounded up
    // to a power of two
    static constexpr int G = N <= 8 ? 8 : 16;
};

template <int KIND>
__device__ __forceinline__ void load_row(const int* rows, int r, const double* x, const double* x0, int (&id)[4], v3 (&px)[4], v3 (&p0)[4])
{
    constexpr int K = KindInfo<KIND>::K;
    constexpr int W = KIND == 0 ? 2 : (KIND == 1 ? 3 : 4);
#pragma unroll
    for (int a = 0; a < 4; a++) id[a] = a < K ? rows[W * r + a] : 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        px[a] = a < K ? ldv(x, id[a]) : mk(0, 0, 0);
        p0[a] = (KIND >= 4 && a < K) ? ldv(x0, id[a]) : mk(0, 0, 0);
    }
}


}  // namespace fipc
