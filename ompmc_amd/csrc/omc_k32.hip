// FP32-transport instantiations (FP64 tallies).
#include "omc_launch_impl.hpp"

namespace omc {

__global__ void k_accumulate(double *e, double *sum, double *sum2, long n, double scale, double *dep) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    double raw = 0.0;
    for (; i < n; i += stride) {
        double v = e[i];
        raw += v;
        if (scale != 1.0) v /= scale;
        sum[i] += v;
        sum2[i] += v * v;
        e[i] = 0.0;
    }
    wave_add(dep, raw);
}

// FP64 atomic throughput probe: every lane adds 1.0 at `nper` addresses of a
// splitmix64 sequence over mask + 1 doubles, with the hardware FP64 atomic
// add the tallies use (global_atomic_add_f64).
__global__ void __launch_bounds__(256) k_atomic_probe(double *t, unsigned long long mask, int nper,
                                                      unsigned long long seed) {
    unsigned long long x = seed + (unsigned long long)(blockIdx.x * blockDim.x + threadIdx.x) * 0xD1B54A32D192ED03ull;
    for (int k = 0; k < nper; k++) {
        x += 0x9E3779B97F4A7C15ull;
        unsigned long long z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        atomicAdd(t + (z & mask), 1.0);
    }
}

hipError_t launch_atomic_probe(double *t, unsigned long long mask, int grid, int nper, unsigned long long seed,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_atomic_probe, dim3(grid), dim3(256), 0, st, t, mask, nper, seed);
    return hipGetLastError();
}

hipError_t launch_accumulate(double *e, double *sum, double *sum2, long n, double scale, double *dep,
                             hipStream_t st) {
    int block = 256;
    long g = (n + block - 1) / block;
    int grid = (int)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(block), 0, st, e, sum, sum2, n, scale, dep);
    return hipGetLastError();
}

template hipError_t launch_history<float>(const HistArgs<float> &, int, bool, int, int, hipStream_t);
template hipError_t launch_sampler<float>(const SampArgs<float> &, hipStream_t);
template hipError_t launch_rounds<float>(const EvLaunch<float> &, long, int, bool, hipStream_t);
template hipError_t launch_round_part<float>(const EvLaunch<float> &, long, int, bool, hipStream_t);
template hipError_t build_rounds_graph<float>(const EvLaunch<float> &, long, int, bool, hipGraph_t *);

hipError_t launch_job_start(EvCtl *c, long long nhist, long long target, long long cap, long long inj_max, int dest,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_job_start, dim3(1), dim3(64), 0, st, c, nhist, target, cap, inj_max, dest);
    return hipGetLastError();
}

}  // namespace omc
