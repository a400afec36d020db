// FP64-transport instantiations: the GPU oracle run and the injected-random-
// number parity path.  This file is compiled with -ffp-contract=off.
#include "omc_launch_impl.hpp"

namespace omc {

template hipError_t launch_history<double>(const HistArgs<double> &, int, bool, int, int, hipStream_t);
template hipError_t launch_sampler<double>(const SampArgs<double> &, hipStream_t);
template hipError_t launch_rounds<double>(const EvLaunch<double> &, long, int, bool, hipStream_t);
template hipError_t launch_round_part<double>(const EvLaunch<double> &, long, int, bool, hipStream_t);

}  // namespace omc

// ------------------------------------------------------------------ Dij post-processing
#include "omc_dij.hpp"

namespace omc {

__global__ void k_accum_slots(double *e, double *sum, double *sum2, long n, long stride, const double *nper) {
    const long s = blockIdx.y;
    const double np = nper[s];
    double *es = e + s * stride, *ss = sum + s * stride, *s2 = sum2 + s * stride;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double v = es[i];
        if (v != 0.0) {
            v /= np;
            ss[i] += v;
            s2[i] += v * v;
            es[i] = 0.0;
        }
    }
}

__device__ __forceinline__ double block_max(double v, double *lds) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double m = lds[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) m = fmax(m, lds[k]);
    return m;
}

__global__ void __launch_bounds__(DIJ_BLOCK)
k_dij_finish(double *sum, double *sum2, double *carry, int nslot, long stride, long nvox, const double *mass,
             const double *rho, int nbatch, double *blockmax, double *tot_dose, double *tot_var) {
    __shared__ double lds[DIJ_BLOCK / 64];
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = v < nvox;
    const long irl = v + 1;
    double m = in ? mass[v] : 1.0, r = in ? rho[v] : 1.0;
    double c = in ? carry[irl] : 0.0, td = 0.0, tv = 0.0;
    const double factor = 1.602E-10 / (m);
    for (int s = 0; s < nslot; s++) {
        double dose = 0.0;
        if (in) {
            double *ps = sum + (long)s * stride + irl, *ps2 = sum2 + (long)s * stride + irl;
            double e = *ps, e2raw = *ps2, e2 = e2raw + c, u, u0;
            e *= factor;
            e2 *= factor * factor;
            e /= (double)nbatch;
            e2 /= (double)(nbatch - 1);
            if (e != 0.0) {
                u = e2 - e * e;
                u /= nbatch;
                // plan statistics: the batch estimate of the variance of the
                // mean, (sum x^2 / n - mean^2) / (n - 1), without the
                // reference's carry-over and without its 1/(n-1) on the sum
                // of squares (which leaves a floor of mean^2 / (n (n-1)))
                double m2 = e2raw * (factor * factor) / (double)nbatch;
                u0 = (m2 - e * e) / (double)(nbatch - 1);
                if (u0 < 0.0) u0 = 0.0;
            } else {
                e = 0.0;
                u = 0.0;
                u0 = 0.0;
            }
            if (r < 0.044) { e = 0.0; u = 0.0; u0 = 0.0; }
            *ps = e;
            *ps2 = u;
            c = u;
            dose = e;
            td += e;
            tv += u0;
        }
        double bm = block_max(dose, lds);
        if (threadIdx.x == 0) blockmax[(long)s * gridDim.x + blockIdx.x] = bm;
    }
    if (in) {
        carry[irl] = c;
        if (tot_dose) tot_dose[irl] += td;
        if (tot_var) tot_var[irl] += tv;
    }
}

__global__ void __launch_bounds__(DIJ_BLOCK) k_dij_thresh(const double *blockmax, int nblk, double rel, double *thresh) {
    __shared__ double lds[DIJ_BLOCK / 64];
    const int s = blockIdx.x;
    double m = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) m = fmax(m, blockmax[(long)s * nblk + b]);
    m = block_max(m, lds);
    if (threadIdx.x == 0) thresh[s] = m * rel;
}

__global__ void __launch_bounds__(DIJ_BLOCK)
k_dij_count(const double *dose, long stride, long nvox, const double *thresh, int nchunk, int *counts) {
    __shared__ int lds[DIJ_BLOCK / 64];
    const int s = blockIdx.y, c = blockIdx.x;
    const double th = thresh[s];
    const double *d = dose + (long)s * stride + 1;
    int n = 0;
    for (int k = 0; k < DIJ_CHUNK / DIJ_BLOCK; k++) {
        long v = (long)c * DIJ_CHUNK + k * DIJ_BLOCK + threadIdx.x;
        if (v < nvox && d[v] > th) n++;
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < DIJ_BLOCK / 64; w++) t += lds[w];
        counts[(long)s * nchunk + c] = t;
    }
}

__global__ void __launch_bounds__(DIJ_BLOCK)
k_dij_compact(const double *dose, const double *var, long stride, long nvox, const double *thresh, int nchunk,
              const int64_t *off, int64_t *ir, double *val, double *vval) {
    __shared__ int lds[DIJ_BLOCK / 64];
    const int s = blockIdx.y, c = blockIdx.x;
    const double th = thresh[s];
    const double *d = dose + (long)s * stride + 1, *dv = var + (long)s * stride + 1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t base = off[(long)s * nchunk + c];
    for (int k = 0; k < DIJ_CHUNK / DIJ_BLOCK; k++) {
        long v = (long)c * DIJ_CHUNK + k * DIJ_BLOCK + threadIdx.x;
        bool keep = v < nvox && d[v] > th;
        unsigned long long mask = __ballot(keep);
        __syncthreads();
        if (lane == 0) lds[w] = __popcll(mask);
        __syncthreads();
        int before = 0, total = 0;
        for (int q = 0; q < DIJ_BLOCK / 64; q++) {
            if (q < w) before += lds[q];
            total += lds[q];
        }
        if (keep) {
            int64_t at = base + before + __popcll(mask & ((1ull << lane) - 1ull));
            ir[at] = v;
            val[at] = d[v];
            if (vval) vval[at] = dv[v];
        }
        base += total;
    }
}

hipError_t launch_accum_slots(double *e, double *sum, double *sum2, long n, long stride, int nslot, const double *nper,
                              hipStream_t st) {
    long g = (n + 255) / 256;
    int gx = (int)(g > 1024 ? 1024 : g);
    hipLaunchKernelGGL(k_accum_slots, dim3(gx, nslot), dim3(256), 0, st, e, sum, sum2, n, stride, nper);
    return hipGetLastError();
}

hipError_t launch_dij_finish(double *sum, double *sum2, double *carry, int nslot, long stride, long nvox,
                             const double *mass, const double *rho, int nbatch, double *blockmax, double *tot_dose,
                             double *tot_var, int *nblk, hipStream_t st) {
    int g = (int)((nvox + DIJ_BLOCK - 1) / DIJ_BLOCK);
    *nblk = g;
    hipLaunchKernelGGL(k_dij_finish, dim3(g), dim3(DIJ_BLOCK), 0, st, sum, sum2, carry, nslot, stride, nvox, mass, rho,
                       nbatch, blockmax, tot_dose, tot_var);
    return hipGetLastError();
}

hipError_t launch_dij_select(const double *dose, long stride, long nvox, int nslot, const double *blockmax, int nblk,
                             double rel, double *thresh, int *counts, hipStream_t st) {
    hipLaunchKernelGGL(k_dij_thresh, dim3(nslot), dim3(DIJ_BLOCK), 0, st, blockmax, nblk, rel, thresh);
    int nchunk = (int)((nvox + DIJ_CHUNK - 1) / DIJ_CHUNK);
    hipLaunchKernelGGL(k_dij_count, dim3(nchunk, nslot), dim3(DIJ_BLOCK), 0, st, dose, stride, nvox, thresh, nchunk,
                       counts);
    return hipGetLastError();
}

hipError_t launch_dij_compact(const double *dose, const double *var, long stride, long nvox, int nslot,
                              const double *thresh, const int64_t *off, int64_t *ir, double *val, double *vval,
                              hipStream_t st) {
    int nchunk = (int)((nvox + DIJ_CHUNK - 1) / DIJ_CHUNK);
    hipLaunchKernelGGL(k_dij_compact, dim3(nchunk, nslot), dim3(DIJ_BLOCK), 0, st, dose, var, stride, nvox, thresh,
                       nchunk, off, ir, val, vval);
    return hipGetLastError();
}

}  // namespace omc
