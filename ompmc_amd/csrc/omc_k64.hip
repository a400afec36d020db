// FP64-transport instantiations: the GPU oracle run and the injected-random-
// number parity path.  This file is compiled with -ffp-contract=off.
#include "omc_launch_impl.hpp"

namespace omc {

template hipError_t launch_history<double>(const HistArgs<double> &, int, bool, int, int, hipStream_t);
template hipError_t launch_sampler<double>(const SampArgs<double> &, hipStream_t);
template hipError_t launch_rounds<double>(const EvLaunch<double> &, long, int, bool, hipStream_t);
template hipError_t launch_round_part<double>(const EvLaunch<double> &, long, int, bool, hipStream_t);
template hipError_t build_rounds_graph<double>(const EvLaunch<double> &, long, int, bool, hipGraph_t *);

}  // namespace omc

// ------------------------------------------------------------------ Dij post-processing
#include "omc_dij.hpp"

namespace omc {

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

// The reference's finish of one voxel of one beamlet (accumulateResults with
// iout = 1, omc_matrad.c:920-1012) from its accumulators: acc = sum of the
// batch means, acc2 = the carried variance plus the sum of their squares (in
// that order), acc2raw = the sum of squares alone.  Writes the dose and the
// reference's (Q11) variance of the mean, and u0, the batch estimate of the
// variance of the mean for plan statistics, (sum x^2 / n - mean^2) / (n - 1)
// without the carry-over (the Q11 formula keeps a floor of mean^2 / (n (n-1))).
__device__ __forceinline__ void dij_finish_voxel(double acc, double acc2, double acc2raw, double factor, double rho,
                                                 int nbatch, double &dose, double &var, double &u0) {
    double e = acc, e2 = acc2, u;
    e *= factor;
    e2 *= factor * factor;
    e /= (double)nbatch;
    e2 /= (double)(nbatch - 1);
    if (e != 0.0) {
        u = e2 - e * e;
        u /= nbatch;
        double m2 = acc2raw * (factor * factor) / (double)nbatch;
        u0 = (m2 - e * e) / (double)(nbatch - 1);
        if (u0 < 0.0) u0 = 0.0;
    } else {
        e = 0.0;
        u = 0.0;
        u0 = 0.0;
    }
    if (rho < 0.044) { e = 0.0; u = 0.0; u0 = 0.0; }
    dose = e;
    var = u;
}

// Fused batch statistics and finish of a group whose jobs carried all its
// batches (slot k * ng + s = beamlet s in batch k).  One wavefront per tile of
// TALLY_TILE regions, four per block; for every beamlet in order it folds the
// batch slots the transport marked as touched (accumEndep, :897-918: the
// sums start from zero and from the carried variance, and take the batches in
// batch order, so the arithmetic is the reference's), clears what it read,
// and finishes the voxel.  Skipping an unmarked slot is exact: its energies
// are zero, and the reference adds 0 and 0^2.  Dose and variance are written
// only in tiles some batch touched (`any` marks them for the selection); in
// the others the dose is zero, and so is the variance carried on.  Also
// per-block maxima, the energy deposited (raw tallies, the EB_DEPOSIT term)
// and the beamlet-summed dose and plan variance.
__global__ void __launch_bounds__(DIJ_BLOCK)
k_dij_fold_finish(double *e, uint8_t *touched, int ng, int nbatch, long nreg, long ntile, const double *nper,
                  double *dose, double *var, uint8_t *any, double *carry, const double *mass, const double *rho,
                  double *blockmax, double *tot_dose, double *tot_var, double *dep) {
    __shared__ double lds[DIJ_BLOCK / 64];
    const int lane = threadIdx.x & 63;
    const long t = (long)blockIdx.x * (DIJ_BLOCK / 64) + (threadIdx.x >> 6);
    const bool tile_in = t < ntile;
    const long irl = t * TALLY_TILE + lane;
    const bool in = irl >= 1 && irl < nreg;
    const double m = in ? mass[irl - 1] : 1.0, r = in ? rho[irl - 1] : 1.0;
    const double factor = 1.602E-10 / (m);
    double c = in ? carry[irl] : 0.0, td = 0.0, tv = 0.0, raw = 0.0;
    for (int s = 0; s < ng; s++) {
        const double np = nper[s];
        double acc = 0.0, acc2 = c, acc2raw = 0.0;
        bool hit = false;
        for (int k = 0; k < nbatch; k++) {
            const long slot = (long)k * ng + s;
            uint8_t *mark = touched + slot * ntile + t;
            if (tile_in && *mark) {                    // the same for every lane of the wave
                hit = true;
                double x = 0.0;
                if (irl < nreg) {
                    double *pe = e + slot * nreg + irl;
                    x = *pe;
                    if (x != 0.0) *pe = 0.0;
                }
                raw += x;
                x /= np;
                acc += x;
                acc2 += x * x;
                acc2raw += x * x;
                if (lane == 0) *mark = 0;
            }
        }
        double dd = 0.0;
        if (in) {
            double vv, u0;
            dij_finish_voxel(acc, acc2, acc2raw, factor, r, nbatch, dd, vv, u0);
            c = vv;
            td += dd;
            tv += u0;
            if (hit) {
                dose[(long)s * nreg + irl] = dd;
                var[(long)s * nreg + irl] = vv;
            }
        }
        if (tile_in && lane == 0) any[(long)s * ntile + t] = hit ? 1 : 0;
        const double bm = block_max(dd, lds);
        if (threadIdx.x == 0) blockmax[(long)s * gridDim.x + blockIdx.x] = bm;
    }
    if (in) {
        carry[irl] = c;
        if (tot_dose && td != 0.0) tot_dose[irl] += td;
        if (tot_var && tv != 0.0) tot_var[irl] += tv;
    }
    wave_add(dep, raw);
}

// Batch statistics of one job of a group whose batches span several jobs:
// for slot s, x = endep / nper[s]; sum += x; sum2 += x^2; endep = 0; the raw
// energies are summed into *dep.
__global__ void k_accum_slots(double *e, double *sum, double *sum2, long n, long stride, const double *nper,
                              double *dep) {
    const long s = blockIdx.y;
    const double np = nper[s];
    double *es = e + s * stride, *ss = sum + s * stride, *s2 = sum2 + s * stride;
    double raw = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double v = es[i];
        if (v != 0.0) {
            raw += v;
            v /= np;
            ss[i] += v;
            s2[i] += v * v;
            es[i] = 0.0;
        }
    }
    wave_add(dep, raw);
}

// Finish of a group after multi-job accumulation (k_accum_slots): the same
// voxel finish, with the carried variance added to the finished sum of
// squares (the one rounding in which this order differs from the
// reference's), in place: sum <- dose, sum2 <- variance.  Blocks of
// DIJ_BLOCK regions, like k_dij_fold_finish; every tile counts as touched.
__global__ void __launch_bounds__(DIJ_BLOCK)
k_dij_finish(double *sum, double *sum2, double *carry, int ng, long nreg, const double *mass, const double *rho,
             int nbatch, double *blockmax, double *tot_dose, double *tot_var) {
    __shared__ double lds[DIJ_BLOCK / 64];
    const long irl = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = irl >= 1 && irl < nreg;
    const double m = in ? mass[irl - 1] : 1.0, r = in ? rho[irl - 1] : 1.0;
    const double factor = 1.602E-10 / (m);
    double c = in ? carry[irl] : 0.0, td = 0.0, tv = 0.0;
    for (int s = 0; s < ng; s++) {
        double dd = 0.0;
        if (in) {
            double *ps = sum + (long)s * nreg + irl, *ps2 = sum2 + (long)s * nreg + irl;
            const double e2raw = *ps2;
            double vv, u0;
            dij_finish_voxel(*ps, e2raw + c, e2raw, factor, r, nbatch, dd, vv, u0);
            *ps = dd;
            *ps2 = vv;
            c = vv;
            td += dd;
            tv += u0;
        }
        const double bm = block_max(dd, lds);
        if (threadIdx.x == 0) blockmax[(long)s * gridDim.x + blockIdx.x] = bm;
    }
    if (in) {
        carry[irl] = c;
        if (tot_dose && td != 0.0) tot_dose[irl] += td;
        if (tot_var && tv != 0.0) tot_var[irl] += tv;
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

// Entries strictly above the threshold per chunk of DIJ_CHUNK regions
// (:1617-1624); a wave takes one tile at a time and skips tiles no batch
// touched.
__global__ void __launch_bounds__(DIJ_BLOCK)
k_dij_count(const double *dose, const uint8_t *any, long nreg, long ntile, const double *thresh, int nchunk,
            int *counts) {
    __shared__ int lds[DIJ_BLOCK / 64];
    const int s = blockIdx.y, c = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double th = thresh[s];
    const double *d = dose + (long)s * nreg;
    int n = 0;
    for (int k = 0; k < DIJ_CHUNK / DIJ_BLOCK; k++) {
        const long t = (long)c * (DIJ_CHUNK / TALLY_TILE) + k * (DIJ_BLOCK / 64) + w;
        if (t < ntile && any[(long)s * ntile + t]) {
            const long irl = t * TALLY_TILE + lane;
            if (irl >= 1 && irl < nreg && d[irl] > th) n++;
        }
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if (lane == 0) lds[w] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tt = 0;
        for (int q = 0; q < DIJ_BLOCK / 64; q++) tt += lds[q];
        counts[(long)s * nchunk + c] = tt;
    }
}

// Order-preserving compaction into CSC rows (voxel index irl - 1, 0-based)
// and values (:1667-1680): the regions of a chunk in increasing order, four
// tiles (one per wave) at a time.
__global__ void __launch_bounds__(DIJ_BLOCK)
k_dij_compact(const double *dose, const double *var, const uint8_t *any, long nreg, long ntile, const double *thresh,
              int nchunk, const int64_t *off, int64_t *ir, double *val, double *vval) {
    __shared__ int lds[DIJ_BLOCK / 64];
    const int s = blockIdx.y, c = blockIdx.x;
    const double th = thresh[s];
    const double *d = dose + (long)s * nreg, *dv = var + (long)s * nreg;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t base = off[(long)s * nchunk + c];
    for (int k = 0; k < DIJ_CHUNK / DIJ_BLOCK; k++) {
        const long t = (long)c * (DIJ_CHUNK / TALLY_TILE) + k * (DIJ_BLOCK / 64) + w;
        const long irl = t * TALLY_TILE + lane;
        const bool keep = t < ntile && any[(long)s * ntile + t] && irl >= 1 && irl < nreg && d[irl] > th;
        const unsigned long long mask = __ballot(keep);
        __syncthreads();
        if (lane == 0) lds[w] = __popcll(mask);
        __syncthreads();
        int before = 0, total = 0;
        for (int q = 0; q < DIJ_BLOCK / 64; q++) {
            if (q < w) before += lds[q];
            total += lds[q];
        }
        if (keep) {
            const int64_t at = base + before + __popcll(mask & ((1ull << lane) - 1ull));
            ir[at] = irl - 1;
            val[at] = d[irl];
            if (vval) vval[at] = dv[irl];
        }
        base += total;
    }
}

hipError_t launch_dij_fold_finish(double *e, uint8_t *touched, int ng, int nbatch, long nreg, const double *nper,
                                  double *dose, double *var, uint8_t *any, double *carry, const double *mass,
                                  const double *rho, double *blockmax, double *tot_dose, double *tot_var, double *dep,
                                  hipStream_t st) {
    const long ntile = dij_ntile(nreg);
    hipLaunchKernelGGL(k_dij_fold_finish, dim3(dij_nblk(nreg)), dim3(DIJ_BLOCK), 0, st, e, touched, ng, nbatch, nreg,
                       ntile, nper, dose, var, any, carry, mass, rho, blockmax, tot_dose, tot_var, dep);
    return hipGetLastError();
}

hipError_t launch_accum_slots(double *e, double *sum, double *sum2, long n, long stride, int nslot, const double *nper,
                              double *dep, hipStream_t st) {
    long g = (n + 255) / 256;
    int gx = (int)(g > 1024 ? 1024 : g);
    hipLaunchKernelGGL(k_accum_slots, dim3(gx, nslot), dim3(256), 0, st, e, sum, sum2, n, stride, nper, dep);
    return hipGetLastError();
}

hipError_t launch_dij_finish(double *sum, double *sum2, double *carry, int ng, long nreg, const double *mass,
                             const double *rho, int nbatch, double *blockmax, double *tot_dose, double *tot_var,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_dij_finish, dim3(dij_nblk(nreg)), dim3(DIJ_BLOCK), 0, st, sum, sum2, carry, ng, nreg, mass,
                       rho, nbatch, blockmax, tot_dose, tot_var);
    return hipGetLastError();
}

hipError_t launch_dij_select(const double *dose, const uint8_t *any, long nreg, int ng, const double *blockmax,
                             double rel, double *thresh, int *counts, hipStream_t st) {
    hipLaunchKernelGGL(k_dij_thresh, dim3(ng), dim3(DIJ_BLOCK), 0, st, blockmax, dij_nblk(nreg), rel, thresh);
    const int nchunk = dij_nchunk(nreg);
    hipLaunchKernelGGL(k_dij_count, dim3(nchunk, ng), dim3(DIJ_BLOCK), 0, st, dose, any, nreg, dij_ntile(nreg), thresh,
                       nchunk, counts);
    return hipGetLastError();
}

hipError_t launch_dij_compact(const double *dose, const double *var, const uint8_t *any, long nreg, int ng,
                              const double *thresh, const int64_t *off, int64_t *ir, double *val, double *vval,
                              hipStream_t st) {
    const int nchunk = dij_nchunk(nreg);
    hipLaunchKernelGGL(k_dij_compact, dim3(nchunk, ng), dim3(DIJ_BLOCK), 0, st, dose, var, any, nreg, dij_ntile(nreg),
                       thresh, nchunk, off, ir, val, vval);
    return hipGetLastError();
}

}  // namespace omc
