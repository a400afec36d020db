// FP64-transport instantiations: the GPU oracle run and the injected-random-
// number parity path.  This file is compiled with -ffp-contract=off.
#include "omc_launch.hpp"

namespace omc {

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

template hipError_t launch_history<double>(const HistArgs<double> &, int, bool, int, int, hipStream_t);
template hipError_t launch_sampler<double>(const SampArgs<double> &, hipStream_t);

}  // namespace omc
