// Plain interpreter instantiations, MSE finalize and small utility kernels.
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

hipError_t launch_batched_mse(int dpt, int xc, const BatchArgs &a, int grid, hipStream_t s);
hipError_t launch_batched_pred(int dpt, int xc, const BatchArgs &a, int grid, hipStream_t s);
hipError_t launch_batched_col(int dpt, int xc, const BatchArgs &a, int grid, hipStream_t s);
hipError_t launch_batched_mseref(int dpt, int xc, const BatchArgs &a, int grid, hipStream_t s);
int batched_occupancy_mse(int dpt, int xc);
int batched_occupancy_pred(int dpt, int xc);
int batched_occupancy_col(int dpt, int xc);
int batched_occupancy_mseref(int dpt, int xc);

hipError_t launch_batched(int mode, int dpt, int xc, const BatchArgs &a, int grid, hipStream_t s)
{
    if (mode == MODE_MSE) return launch_batched_mse(dpt, xc, a, grid, s);
    if (mode == MODE_PRED) return launch_batched_pred(dpt, xc, a, grid, s);
    if (mode == MODE_COLUMN) return launch_batched_col(dpt, xc, a, grid, s);
    if (mode == MODE_MSE_REF) return launch_batched_mseref(dpt, xc, a, grid, s);
    return hipErrorInvalidValue;
}

int batched_occupancy(int mode, int dpt, int xc)
{
    if (mode == MODE_MSE) return batched_occupancy_mse(dpt, xc);
    if (mode == MODE_PRED) return batched_occupancy_pred(dpt, xc);
    if (mode == MODE_COLUMN) return batched_occupancy_col(dpt, xc);
    if (mode == MODE_MSE_REF) return batched_occupancy_mseref(dpt, xc);
    return 0;
}

hipError_t launch_plain(int s_slots, int xc, const Ins *prog, int len, const float *X, int64_t N,
                        int nv, double *psum, long long *pcnt, hipStream_t s)
{
    const unsigned grid = (unsigned)((N + kBlock - 1) / kBlock);
#define SR_PLAIN(SV, XCV)                                                                  \
    if (s_slots <= SV && xc == XCV) {                                                      \
        hipLaunchKernelGGL((k_plain<SV, XCV>), dim3(grid), dim3(kBlock), 0, s, prog, len, X, \
                           N, nv, psum, pcnt);                                             \
        return hipGetLastError();                                                          \
    }
    SR_PLAIN(16, 4)
    SR_PLAIN(16, 16)
    SR_PLAIN(16, 32)
    SR_PLAIN(64, 4)
    SR_PLAIN(64, 16)
    SR_PLAIN(64, 32)
#undef SR_PLAIN
    return hipErrorInvalidValue;
}

// One wave per expression; partials summed lane-strided then by a fixed
// shuffle tree, so the result is deterministic.
__global__ __launch_bounds__(kBlock) void k_finalize(const double *psum, const long long *pcnt,
                                                     int P, int E, const int *E_dev, float *mse,
                                                     long long *cnt)
{
    const int e = (int)((blockIdx.x * (unsigned)kBlock + threadIdx.x) / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    if (E_dev && e >= *E_dev) return;
    if (e >= E) return;
    double s = 0.0;
    long long c = 0;
    for (int k = lane; k < P; k += kWave) {
        s += psum[(int64_t)e * P + k];
        c += pcnt[(int64_t)e * P + k];
    }
    s = wave_sum(s);
    c = wave_sum(c);
    if (lane == 0) {
        mse[e] = c > 0 ? (float)(s / (double)c) : qnan();
        if (cnt) cnt[e] = c;
    }
}

hipError_t launch_finalize(const double *psum, const long long *pcnt, int P, int E, float *mse,
                           long long *cnt, hipStream_t s, const int *E_dev)
{
    const int per_block = kBlock / kWave;
    hipLaunchKernelGGL(k_finalize, dim3((E + per_block - 1) / per_block), dim3(kBlock), 0, s, psum,
                       pcnt, P, E, E_dev, mse, cnt);
    return hipGetLastError();
}

__global__ void k_zero_u32(uint32_t *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace sr
