// Internal host-side types shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/omc_gpu.h"
#include "omc_host.h"
#include "omc_launch.hpp"

// Device allocations owned together and released together.  An arena set to
// use_cache(stream) takes its blocks from a per-device cache and hands them
// back on release, after synchronising that stream.  hipFree synchronises the
// whole device, which would stall every other stream of a run or of a
// server.  And mapping ~100 GiB anew costs seconds, so run-scoped arenas
// reuse blocks across calls.
struct DevArena {
    std::vector<std::pair<void *, size_t>> blocks;
    size_t bytes = 0;
    bool ok = true;
    bool cached = false;
    hipStream_t stream = nullptr;     // cached arenas: the stream their blocks are used on
    void use_cache(hipStream_t s) {
        cached = true;
        stream = s;
    }
    void *alloc(size_t n);
    template <class T> T *alloc_n(size_t n) {
        T *p = (T *)alloc(sizeof(T) * n);
        if (!p) ok = false;
        return p;
    }
    template <class T> const T *upload(const std::vector<T> &v) {
        T *p = alloc_n<T>(v.size() ? v.size() : 1);
        if (!p) return nullptr;
        if (!v.empty() && hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess) ok = false;
        return p;
    }
    void release();
};

// Block cache of the current device (see DevArena).  devcache_held() is the
// memory it holds for reuse; devcache_trim() returns all of it to the runtime.
void *devcache_alloc(size_t n, size_t *got);
void devcache_free(void *p, size_t n);
size_t devcache_held();
void devcache_trim();
// Small pinned host buffers (control-block snapshots), kept for reuse: hipHostFree
// also synchronises the device.
void *pinned_small_alloc();
void pinned_small_free(void *p);
constexpr size_t PINNED_SMALL_BYTES = 4096;
// Large pinned host blocks for results handed to the caller (downloaded at
// DMA speed into pages that are already mapped), kept for reuse when
// released: at most PINNED_CACHE_MAX bytes are held.
void *pinned_block_alloc(size_t n);
bool pinned_block_free(void *p);     // false: p is not a block of the pool (left untouched)
constexpr size_t PINNED_CACHE_MAX = (size_t)4 << 30;

struct PhantomDev {
    omc_phantom host;            // owned copy (dose conversion, source set-up)
    double ecut = 0, pcut = 0;
    long nreg = 0;               // voxels + 1
    DevArena mem;
    const uint8_t *dmed = nullptr;
    const double *dmass = nullptr;   // voxel mass (g): the reference's dx*dy*dz*rho, for the Dij finish
    const double *drho = nullptr;
    omc::Geom<float> g32;            // host copies of the descriptors
    omc::Geom<double> g64;
    const omc::Geom<float> *dg32 = nullptr;   // device copies passed to kernels
    const omc::Geom<double> *dg64 = nullptr;
};

struct omc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    omc_tables host;
    DevArena tabmem;
    DevArena scratch;
    omc::Tables<float> t32;          // host copies of the descriptors
    omc::Tables<double> t64;
    const omc::Tables<float> *dt32 = nullptr;  // device copies passed to kernels
    const omc::Tables<double> *dt64 = nullptr;
    std::map<int, PhantomDev *> phantoms;
    int next_id = 1;
};

// device-resident table and geometry descriptors of a precision
template <class R> const omc::Tables<R> *tables_of(const omc_ctx *c);
template <> inline const omc::Tables<float> *tables_of<float>(const omc_ctx *c) { return c->dt32; }
template <> inline const omc::Tables<double> *tables_of<double>(const omc_ctx *c) { return c->dt64; }
// the host copies of the same descriptors (device pointers inside)
template <class R> const omc::Tables<R> &host_tables_of(const omc_ctx *c);
template <> inline const omc::Tables<float> &host_tables_of<float>(const omc_ctx *c) { return c->t32; }
template <> inline const omc::Tables<double> &host_tables_of<double>(const omc_ctx *c) { return c->t64; }
template <class R> const omc::Geom<R> *geom_of(const PhantomDev *p);
template <> inline const omc::Geom<float> *geom_of<float>(const PhantomDev *p) { return p->dg32; }
template <> inline const omc::Geom<double> *geom_of<double>(const PhantomDev *p) { return p->dg64; }

int omc_seterr(int code, const char *fmt, ...);
// device address of the no-allocation scratch array on `device` (current), once per device
int omc_noalloc_init(int device);
int omc_check_opts(const omc_run_opts *o);
PhantomDev *omc_get_phantom(omc_ctx *c, int id);
// cache_stream: the phantom's device memory comes from the block cache and is
// used on that stream only (server requests); null: plain allocations
int omc_phantom_make(omc_ctx *c, const omc_phantom_desc *d, PhantomDev **out, hipStream_t cache_stream = nullptr);
void omc_phantom_destroy(PhantomDev *pd);
size_t omc_phantom_bytes(const omc_phantom_desc *d);

// event-queue form (omc_event.hip)
template <class R>
int omc_run_dosxyz_event(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_dosxyz_source_desc *src,
                         int64_t nper, int nbatch, const omc_run_opts *o, double *sum, double *sum2, omc_run_stats *st);
// Dij run on a given stream (omc_event.hip)
int omc_dij_run(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_spectrum_desc *spec,
                const omc_beamlets_desc *bix, int nbatch, double rel, const omc_run_opts *o, omc_csc *dij,
                omc_csc *var, double *plan_dose, double *plan_var, omc_run_stats *st);
// dose run on a given stream (omc_capi.hip); the server runs requests through it
int omc_dosxyz_run(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_dosxyz_source_desc *src,
                   int64_t nperbatch, int nbatch, const omc_run_opts *o, double *sum, double *sum2, omc_run_stats *st);

template <class R> inline void fill_spectrum(DevArena &A, const omc_spectrum_desc &sp, int &spectrum, R &energy,
                                             R &deltak, const R *&c1, const R *&c2) {
    spectrum = sp.spectrum;
    energy = (R)sp.energy;
    deltak = (R)OMC_INVDIM;
    c1 = c2 = nullptr;
    if (sp.spectrum) {
        std::vector<R> a(OMC_INVDIM), b(OMC_INVDIM);
        for (int k = 0; k < OMC_INVDIM; k++) { a[k] = (R)sp.cdfinv1[k]; b[k] = (R)sp.cdfinv2[k]; }
        c1 = A.upload(a);
        c2 = A.upload(b);
    }
}

template <class R> inline int make_dosxyz_src(DevArena &A, const PhantomDev *pd, const omc_dosxyz_source_desc *src,
                                              omc::SrcDosxyz<R> &S) {
    omc_dosxyz_source hs;
    memset(&hs, 0, sizeof hs);
    int rc = omc_dosxyz_source_init(&hs, &pd->host, src->collimator, src->ssd, src->charge);
    if (rc) return rc;
    if (src->spec.spectrum && (!src->spec.cdfinv1 || !src->spec.cdfinv2))
        return omc_seterr(OMC_ERR_ARG, "spectrum source without inverse CDF");
    fill_spectrum<R>(A, src->spec, S.spectrum, S.energy, S.deltak, S.cdfinv1, S.cdfinv2);
    S.charge = hs.charge;
    S.ssd = (R)hs.ssd;
    S.xinl = (R)hs.xinl;
    S.yinl = (R)hs.yinl;
    S.xsize = (R)hs.xsize;
    S.ysize = (R)hs.ysize;
    S.ixinl = hs.ixinl;
    S.iyinl = hs.iyinl;
    return OMC_OK;
}

#define HIPCHK(x)                                                                                         \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) { rc = omc_seterr(OMC_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); goto done; } \
    } while (0)

// History-form step guard: transport calls (one photon flight, or an
// electron's transport up to its next secondary) one history may take before
// it counts as stuck; max_particle_steps sets it, 0 means 10^8.
inline long hist_max_steps(const omc_run_opts *o) {
    return o->max_particle_steps > 0 ? (long)o->max_particle_steps : 100000000L;
}

// Device-memory particle stacks for history-form splitting beyond the
// scratch stack: split_stack_depth(nsplit) entries for every lane of a
// launch of up to max_grid blocks of `block` lanes.  The stacks take at most
// 8 GiB; max_grid comes back lowered to what fits, and launches must not
// exceed it.
template <class R> int alloc_split_stacks(DevArena &A, omc::HistArgs<R> &a, int nsplit, int block, int &max_grid) {
    a.stack = nullptr;
    a.stack_depth = 0;
    if (nsplit <= omc::HISTORY_SCRATCH_NSPLIT) return OMC_OK;
    const int depth = omc::split_stack_depth(nsplit);
    const size_t per_block = (size_t)block * depth * sizeof(omc::Part<R>);
    size_t nb = ((size_t)8 << 30) / per_block;
    if (nb < 1) nb = 1;
    if (max_grid < 1) max_grid = 1;
    if ((size_t)max_grid > nb) max_grid = (int)nb;
    a.stack = A.alloc_n<omc::Part<R>>((size_t)max_grid * block * depth);
    if (!A.ok) return omc_seterr(OMC_ERR_NOMEM, "particle stacks for nsplit %d", nsplit);
    a.stack_depth = depth;
    return OMC_OK;
}
