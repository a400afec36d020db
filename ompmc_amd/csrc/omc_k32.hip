// FP32-transport instantiations (FP64 tallies).
#include "omc_launch.hpp"

namespace omc {

__global__ void k_accumulate(double *e, double *sum, double *sum2, long n, double scale) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double v = e[i];
        if (scale != 1.0) v /= scale;
        sum[i] += v;
        sum2[i] += v * v;
        e[i] = 0.0;
    }
}

hipError_t launch_accumulate(double *e, double *sum, double *sum2, long n, double scale, hipStream_t st) {
    int block = 256;
    long g = (n + block - 1) / block;
    int grid = (int)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(block), 0, st, e, sum, sum2, n, scale);
    return hipGetLastError();
}

template <class R> hipError_t launch_history(const HistArgs<R> &a, int rng, bool count, int grid, int block,
                                             hipStream_t st) {
    if (rng == RNG_REPLAY) {
        if (count) hipLaunchKernelGGL((k_history<R, Replay, true>), dim3(grid), dim3(block), 0, st, a);
        else hipLaunchKernelGGL((k_history<R, Replay, false>), dim3(grid), dim3(block), 0, st, a);
    } else {
        if (count) hipLaunchKernelGGL((k_history<R, Philox, true>), dim3(grid), dim3(block), 0, st, a);
        else hipLaunchKernelGGL((k_history<R, Philox, false>), dim3(grid), dim3(block), 0, st, a);
    }
    return hipGetLastError();
}

template <class R> hipError_t launch_sampler(const SampArgs<R> &a, hipStream_t st) {
    int block = 64;
    int grid = (a.ncase + block - 1) / block;
    hipLaunchKernelGGL((k_sampler<R>), dim3(grid), dim3(block), 0, st, a);
    return hipGetLastError();
}

template hipError_t launch_history<float>(const HistArgs<float> &, int, bool, int, int, hipStream_t);
template hipError_t launch_sampler<float>(const SampArgs<float> &, hipStream_t);

}  // namespace omc
