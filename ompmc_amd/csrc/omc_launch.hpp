// Launchers of the templated kernels, instantiated per precision in
// omc_k32.hip (float) and omc_k64.hip (double, built without FP contraction
// so that it rounds like the CPU reference).
#pragma once
#include <hip/hip_runtime.h>
#include "omc_kernels.hpp"

namespace omc {

template <class R> hipError_t launch_history(const HistArgs<R> &a, int rng, bool count, int grid, int block,
                                             hipStream_t st);
template <class R> hipError_t launch_sampler(const SampArgs<R> &a, hipStream_t st);
hipError_t launch_accumulate(double *e, double *sum, double *sum2, long n, double scale, hipStream_t st);

}  // namespace omc
