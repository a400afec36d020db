// FP32-transport instantiations (FP64 tallies).
#include "omc_launch_impl.hpp"

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

template hipError_t launch_history<float>(const HistArgs<float> &, int, bool, int, int, hipStream_t);
template hipError_t launch_sampler<float>(const SampArgs<float> &, hipStream_t);
template hipError_t launch_rounds<float>(const EvLaunch<float> &, long, int, bool, hipStream_t);
template hipError_t launch_round_part<float>(const EvLaunch<float> &, long, int, bool, hipStream_t);

hipError_t launch_job_start(EvCtl *c, long long nhist, long long target, long long cap, int dest, hipStream_t st) {
    hipLaunchKernelGGL(k_job_start, dim3(1), dim3(64), 0, st, c, nhist, target, cap, dest);
    return hipGetLastError();
}

}  // namespace omc
