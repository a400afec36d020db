// Dose runs in the history form without any device allocation
// (omc_run_dosxyz_noalloc): every buffer a run needs beyond its inputs and
// outputs -- the batch tally, the source's inverse CDF, counters, error
// flags, the energy balance and the work counter -- lies in one fixed device
// array, `scratch`, that the library holds on each device it runs on until
// the process ends.  For callers that cannot call hipMalloc: code under HIP
// graph capture, tight serving loops, harnesses that forbid it.
//
// In host code `scratch` names a host-side stand-in, not the device array:
// its device address is taken once per device at context creation
// (omc_noalloc_init, hipGetSymbolAddress with that device current) and only
// that address is handed to kernels and copies.  Calls that overlap on one
// device share the array and race; each call clears what must start at zero
// with a fill kernel (no hipMemsetAsync), so a call never sees the previous
// call's contents.
#include "omc_capi_internal.hpp"

#include <chrono>
#include <mutex>

using namespace omc;

namespace {

// Layout of the scratch array (bytes): the batch tally of up to
// OMC_NOALLOC_MAX_REGIONS doubles, then a block for the source's inverse CDF
// (two arrays of OMC_INVDIM values of the transport precision), then small
// fields: 8 event counters, 4 error flags, the energy balance, the work counter.
constexpr size_t TALLY_BYTES = (size_t)OMC_NOALLOC_MAX_REGIONS * sizeof(double);
constexpr size_t CDF_OFF = TALLY_BYTES;
constexpr size_t CDF_BYTES = 2 * (size_t)OMC_INVDIM * sizeof(double);
constexpr size_t SMALL_OFF = CDF_OFF + ((CDF_BYTES + 255) / 256) * 256;
constexpr size_t CNT_OFF = SMALL_OFF;                                   // 8 x u64
constexpr size_t ERR_OFF = CNT_OFF + 8 * sizeof(unsigned long long);    // 4 x int
constexpr size_t EBAL_OFF = ERR_OFF + 4 * sizeof(int) + 16;             // EB_N x double
constexpr size_t NEXT_OFF = EBAL_OFF + 8 * sizeof(double);              // 1 x u64
constexpr size_t SMALL_END = NEXT_OFF + sizeof(unsigned long long);
constexpr size_t SCRATCH_BYTES = SMALL_OFF + 4096;
static_assert(EB_N <= 8, "energy balance block");
static_assert(SMALL_END <= SCRATCH_BYTES, "scratch layout");
static_assert(OMC_NOALLOC_SCRATCH_BYTES == SCRATCH_BYTES, "header constant");

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char *g_ptr[MAX_DEVICES] = {};    // device address of `scratch` on each device

// Zeroes n 64-bit words (the fills a captured graph can replay).
__global__ void k_zero64(unsigned long long *p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0ull;
}

hipError_t zero64(void *p, size_t bytes, hipStream_t st) {
    const long n = (long)(bytes / 8);
    const int block = 256;
    const long g = (n + block - 1) / block;
    const int grid = (int)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
    hipLaunchKernelGGL(k_zero64, dim3(grid), dim3(block), 0, st, (unsigned long long *)p, n);
    return hipGetLastError();
}

template <class R>
int run_noalloc(omc_ctx *c, PhantomDev *pd, const omc_dosxyz_source_desc *src, int64_t nper, int nbatch,
                const omc_run_opts *o, double *d_sum, double *d_sum2, omc_run_stats *st, unsigned char *base) {
    int rc = OMC_OK;
    const hipStream_t stream = c->stream;
    auto wall0 = std::chrono::steady_clock::now();
    const long nreg = pd->nreg;
    double *d_e = (double *)base;
    R *d_cdf = (R *)(base + CDF_OFF);
    unsigned long long *d_cnt = (unsigned long long *)(base + CNT_OFF);
    int *d_err = (int *)(base + ERR_OFF);
    double *d_ebal = (double *)(base + EBAL_OFF);
    unsigned long long *d_next = (unsigned long long *)(base + NEXT_OFF);
    int err[4] = {0, 0, 0, 0};
    HistArgs<R> a;
    memset(&a, 0, sizeof a);
    a.t = tables_of<R>(c);
    a.g = geom_of<R>(pd);
    {
        // the source as make_dosxyz_src builds it, its inverse CDF in scratch
        omc_dosxyz_source hs;
        memset(&hs, 0, sizeof hs);
        rc = omc_dosxyz_source_init(&hs, &pd->host, src->collimator, src->ssd, src->charge);
        if (rc) return rc;
        SrcDosxyz<R> &S = a.dos;
        S.spectrum = src->spec.spectrum;
        S.energy = (R)src->spec.energy;
        S.deltak = (R)OMC_INVDIM;
        S.cdfinv1 = S.cdfinv2 = nullptr;
        if (src->spec.spectrum) {
            if (!src->spec.cdfinv1 || !src->spec.cdfinv2)
                return omc_seterr(OMC_ERR_ARG, "spectrum source without inverse CDF");
            R h[2 * OMC_INVDIM];
            for (int k = 0; k < OMC_INVDIM; k++) {
                h[k] = (R)src->spec.cdfinv1[k];
                h[OMC_INVDIM + k] = (R)src->spec.cdfinv2[k];
            }
            HIPCHK(hipMemcpyAsync(d_cdf, h, sizeof h, hipMemcpyHostToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));     // h lives on this stack frame
            S.cdfinv1 = d_cdf;
            S.cdfinv2 = d_cdf + OMC_INVDIM;
        }
        S.charge = hs.charge;
        S.ssd = (R)hs.ssd;
        S.xinl = (R)hs.xinl;
        S.yinl = (R)hs.yinl;
        S.xsize = (R)hs.xsize;
        S.ysize = (R)hs.ysize;
        S.ixinl = hs.ixinl;
        S.iyinl = hs.iyinl;
    }
    // the tally, counters, flags and balance of the previous call are cleared;
    // the caller's sums start at zero
    HIPCHK(zero64(d_e, sizeof(double) * nreg, stream));
    HIPCHK(zero64(base + SMALL_OFF, SMALL_END - SMALL_OFF, stream));
    HIPCHK(zero64(d_sum, sizeof(double) * nreg, stream));
    HIPCHK(zero64(d_sum2, sizeof(double) * nreg, stream));
    a.endep = d_e;
    a.ebal = d_ebal;
    a.counters = d_cnt;
    a.err = d_err;
    a.next = d_next;
    a.seed = o->seed;
    a.chunk = 1;
    a.max_steps = hist_max_steps(o);
    a.nsplit = o->nsplit;
    for (int b = 0; b < nbatch; b++) {
        a.nhist = nper;
        a.hist_base = (long)b * nper;
        HIPCHK(zero64(d_next, sizeof(unsigned long long), stream));
        const int block = o->block > 0 ? o->block : 256;
        const long want = (nper + block - 1) / block;
        const int grid = o->grid > 0 ? o->grid : (int)(want < 2048 ? want : 2048);
        HIPCHK(launch_history<R>(a, RNG_STREAMS, (o->count_events & 1) != 0, grid, block, stream));
        HIPCHK(launch_accumulate(d_e, d_sum, d_sum2, nreg, 1.0, d_ebal + EB_DEPOSIT, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(err, d_err, sizeof err, hipMemcpyDeviceToHost));
    if (err[0]) rc = omc_seterr(OMC_ERR_STACK, "particle stack overflow on the device");
    else if (err[1]) rc = omc_seterr(OMC_ERR_ARG, "replay stream underrun");
    else if (err[2]) rc = omc_seterr(OMC_ERR_LIMIT, "history step guard reached");
    if (st) {
        unsigned long long cnt[8];
        double eb[EB_N];
        HIPCHK(hipMemcpy(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(eb, d_ebal, sizeof eb, hipMemcpyDeviceToHost));
        st->e_source = eb[EB_SOURCE];
        st->e_escape = eb[EB_ESCAPE];
        st->e_lost = eb[EB_LOST];
        st->e_deposit = eb[EB_DEPOSIT];
        st->histories = (double)nper * nbatch;
        st->n_howfar = cnt[0]; st->n_hownear = cnt[1]; st->n_cross = cnt[2]; st->n_dep = cnt[3];
        st->n_rn = cnt[4]; st->n_photon = cnt[5]; st->n_electron = cnt[6];
        st->errors = (err[0] ? 1 : 0) | (err[1] ? 2 : 0) | (err[2] ? 4 : 0);
        st->kernel_launches[OMC_K_HISTORY] = (uint64_t)nbatch;
        st->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
        st->gpu_seconds = st->wall_seconds;     // no timing events: they are runtime objects of their own
    }
done:
    return rc;
}

}  // namespace

// Device address of `scratch` on `device`, taken once per device with that
// device current.
int omc_noalloc_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return omc_seterr(OMC_ERR_ARG, "device %d", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return OMC_OK;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess || cur != device)
        return omc_seterr(OMC_ERR_DEVICE, "scratch set-up needs device %d current", device);
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return omc_seterr(OMC_ERR_DEVICE, "address of the scratch array on device %d: %s", device,
                          hipGetErrorString(e));
    g_ptr[device] = (unsigned char *)p;
    return OMC_OK;
}

extern "C" int omc_run_dosxyz_noalloc(omc_ctx *c, int phantom_id, const omc_dosxyz_source_desc *src,
                                      int64_t nperbatch, int nbatch, const omc_run_opts *o, double *sum, double *sum2,
                                      omc_run_stats *st) {
    if (!c) return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz_noalloc: null context");
    PhantomDev *pd = omc_get_phantom(c, phantom_id);
    if (!pd) return omc_seterr(OMC_ERR_ARG, "no phantom %d", phantom_id);
    if (!src || !sum || !sum2 || nperbatch < 1 || nbatch < 1)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz_noalloc: bad arguments");
    int rc = omc_check_opts(o);
    if (rc) return rc;
    if (pd->nreg > OMC_NOALLOC_MAX_REGIONS)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz_noalloc: %ld regions exceed the scratch array's %d",
                          (long)pd->nreg, OMC_NOALLOC_MAX_REGIONS);
    if (o->nsplit > HISTORY_SCRATCH_NSPLIT)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz_noalloc: nsplit %d (at most %d: deeper stacks need device "
                          "memory)", o->nsplit, HISTORY_SCRATCH_NSPLIT);
    if (o->rng != OMC_RNG_STREAMS)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz_noalloc: keyed random streams only");
    if (o->form != OMC_FORM_HISTORY)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz_noalloc: the history form only (the event form's pools "
                          "are sized per run)");
    if (hipSetDevice(c->device) != hipSuccess) return omc_seterr(OMC_ERR_DEVICE, "hipSetDevice(%d)", c->device);
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = c->device < MAX_DEVICES ? g_ptr[c->device] : nullptr;
    }
    if (!base) return omc_seterr(OMC_ERR_DEVICE, "scratch array of device %d not set up", c->device);
    if (st) memset(st, 0, sizeof *st);
    return o->precision == OMC_FP32 ? run_noalloc<float>(c, pd, src, nperbatch, nbatch, o, sum, sum2, st, base)
                                    : run_noalloc<double>(c, pd, src, nperbatch, nbatch, o, sum, sum2, st, base);
}
