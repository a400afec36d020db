// Host implementation of the C ABI in include/omc_gpu.h: device context,
// table and phantom upload, dose and sampler entry points.  The Dij and
// serving entry points live in omc_dij.hip and omc_server.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "omc_capi_internal.hpp"

using namespace omc;

// ------------------------------------------------------------------ errors
extern "C" const char *omc_last_error(void) { return omc_host_last_error(); }
extern "C" int omc_api_version(void) { return OMC_API_VERSION; }

int omc_seterr(int code, const char *fmt, ...) {
    char buf[480];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return omc_host_set_error(code, "%s", buf);
}

// ------------------------------------------------------------------ device memory
namespace {

// Released run blocks of one device, by size.  A request takes the smallest
// cached block of at least its size and at most 5/4 of it; blocks are
// rounded to 2 MiB.  The cache holds at most CACHE_MAX bytes; a block that
// would exceed it goes back to the runtime.
struct DevCache {
    std::mutex mu;
    std::multimap<size_t, void *> free;
    size_t held = 0;
};
constexpr size_t CACHE_GRAIN = (size_t)2 << 20;
constexpr size_t CACHE_MAX = (size_t)128 << 30;
constexpr int CACHE_DEVICES = 64;

DevCache &cache_here() {
    static DevCache caches[CACHE_DEVICES];
    int d = 0;
    (void)hipGetDevice(&d);
    return caches[(unsigned)d % CACHE_DEVICES];
}

std::mutex pinned_mu;
std::vector<void *> pinned_free;

}  // namespace

void *devcache_alloc(size_t n, size_t *got) {
    n = (n + CACHE_GRAIN - 1) / CACHE_GRAIN * CACHE_GRAIN;
    DevCache &C = cache_here();
    {
        std::lock_guard<std::mutex> lk(C.mu);
        auto it = C.free.lower_bound(n);
        if (it != C.free.end() && it->first <= n + n / 4) {
            void *p = it->second;
            *got = it->first;
            C.held -= it->first;
            C.free.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess) {
        (void)hipGetLastError();          // recovered below: keep it from surfacing at the next launch check
        devcache_trim();
        if (hipMalloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    *got = n;
    return p;
}

void devcache_free(void *p, size_t n) {
    DevCache &C = cache_here();
    {
        std::lock_guard<std::mutex> lk(C.mu);
        if (C.held + n <= CACHE_MAX) {
            C.free.emplace(n, p);
            C.held += n;
            return;
        }
    }
    (void)hipFree(p);
}

size_t devcache_held() {
    DevCache &C = cache_here();
    std::lock_guard<std::mutex> lk(C.mu);
    return C.held;
}

void devcache_trim() {
    DevCache &C = cache_here();
    std::multimap<size_t, void *> out;
    {
        std::lock_guard<std::mutex> lk(C.mu);
        out.swap(C.free);
        C.held = 0;
    }
    for (auto &kv : out) (void)hipFree(kv.second);
}

void *pinned_small_alloc() {
    {
        std::lock_guard<std::mutex> lk(pinned_mu);
        if (!pinned_free.empty()) {
            void *p = pinned_free.back();
            pinned_free.pop_back();
            return p;
        }
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, PINNED_SMALL_BYTES, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

namespace {
std::mutex pinned_big_mu;
std::multimap<size_t, void *> pinned_big_free;       // released blocks by size
std::map<void *, size_t> pinned_big_live;            // blocks handed out, with their sizes
size_t pinned_big_held = 0;
}  // namespace

void *pinned_block_alloc(size_t n) {
    n = (std::max<size_t>(n, 1) + CACHE_GRAIN - 1) / CACHE_GRAIN * CACHE_GRAIN;
    std::lock_guard<std::mutex> lk(pinned_big_mu);
    auto it = pinned_big_free.lower_bound(n);
    void *p = nullptr;
    size_t got = n;
    if (it != pinned_big_free.end() && it->first <= n + n / 4) {
        p = it->second;
        got = it->first;
        pinned_big_held -= got;
        pinned_big_free.erase(it);
    } else if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    pinned_big_live[p] = got;
    return p;
}

bool pinned_block_free(void *p) {
    if (!p) return false;
    size_t n = 0;
    {
        std::lock_guard<std::mutex> lk(pinned_big_mu);
        auto it = pinned_big_live.find(p);
        if (it == pinned_big_live.end()) return false;  // not a block of this pool
        n = it->second;
        pinned_big_live.erase(it);
        if (pinned_big_held + n <= PINNED_CACHE_MAX) {
            pinned_big_free.emplace(n, p);
            pinned_big_held += n;
            return true;
        }
    }
    (void)hipHostFree(p);
    return true;
}

void pinned_small_free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(pinned_mu);
    pinned_free.push_back(p);
}

void DevArena::release() {
    if (cached && !blocks.empty()) {
        // a block goes back to the cache only once the work using it is done
        if (stream) (void)hipStreamSynchronize(stream);
        else (void)hipDeviceSynchronize();
        for (auto &b : blocks) devcache_free(b.first, b.second);
    } else {
        for (auto &b : blocks) (void)hipFree(b.first);
    }
    blocks.clear();
    bytes = 0;
}

void *DevArena::alloc(size_t n) {
    void *p = nullptr;
    size_t got = n ? n : 1;
    if (cached) {
        p = devcache_alloc(got, &got);
        if (!p) return nullptr;
    } else if (hipMalloc(&p, got) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    blocks.emplace_back(p, got);
    bytes += n;
    return p;
}

// ------------------------------------------------------------------ tables
template <class R> static int upload_tables(omc_ctx *c, DevArena &A, Tables<R> &T) {
    const omc_tables &h = c->host;
    int nm = h.nmed;
    T.nmed = nm;
    std::vector<Med<R>> med(nm);
    for (int i = 0; i < nm; i++) {
        Med<R> &m = med[i];
        memset(&m, 0, sizeof m);
        m.ge0 = (R)h.ge0[i]; m.ge1 = (R)h.ge1[i]; m.eke0 = (R)h.eke0[i]; m.eke1 = (R)h.eke1[i];
        m.rho = (R)h.rho[i]; m.te = (R)h.te[i]; m.thmoll = (R)h.thmoll[i]; m.ap = (R)h.ap[i];
        m.delcm = (R)h.delcm[i]; m.zbrang = (R)h.zbrang[i]; m.bpar0 = (R)h.bpar0[i]; m.bpar1 = (R)h.bpar1[i];
        m.blcc = (R)h.blcc[i]; m.xcc = (R)h.xcc[i]; m.esig_e = (R)h.esig_e[i]; m.psig_e = (R)h.psig_e[i];
        m.e_array0 = (R)h.e_array[(size_t)i * OMC_MXEKE];
        for (int k = 0; k < 8; k++) {
            m.dl1[k] = (R)h.dl1[8 * i + k]; m.dl2[k] = (R)h.dl2[8 * i + k]; m.dl3[k] = (R)h.dl3[8 * i + k];
            m.dl4[k] = (R)h.dl4[8 * i + k]; m.dl5[k] = (R)h.dl5[8 * i + k]; m.dl6[k] = (R)h.dl6[8 * i + k];
        }
        m.sig_mono[0] = h.sig_ismonotone[i];
        m.sig_mono[1] = h.sig_ismonotone[nm + i];
        m.meke = h.meke[i];
    }
    T.med = A.upload(med);
    auto pw = [&](const double *c0, const double *c1, size_t n) {
        std::vector<PW<R>> v(n);
        for (size_t k = 0; k < n; k++) { v[k].c1 = (R)c1[k]; v[k].c0 = (R)c0[k]; }
        return A.upload(v);
    };
    auto vec = [&](const double *p, size_t n) {
        std::vector<R> v(n);
        for (size_t k = 0; k < n; k++) v[k] = (R)p[k];
        return A.upload(v);
    };
    size_t nge = (size_t)nm * OMC_MXGE, nek = (size_t)nm * OMC_MXEKE;
    T.gmfp = pw(h.gmfp0, h.gmfp1, nge);
    T.gbr1 = pw(h.gbr10, h.gbr11, nge);
    T.gbr2 = pw(h.gbr20, h.gbr21, nge);
    T.cohe = pw(h.cohe0, h.cohe1, nge);
    T.pmax = pw(h.ray_pmax0, h.ray_pmax1, nge);
    // the sampler reads the form-factor CDF of medium 0 only (reference quirk Q1)
    T.ray_xgrid = vec(h.ray_xgrid, OMC_MXRAYFF);
    T.ray_fcum = vec(h.ray_fcum, OMC_MXRAYFF);
    T.ray_b = vec(h.ray_b, OMC_MXRAYFF);
    T.ray_c = vec(h.ray_c, OMC_MXRAYFF);
    T.ray_iarr = A.upload(std::vector<int>(h.ray_iarr, h.ray_iarr + OMC_RAYCDFSIZE));
    T.esig = pw(h.esig0, h.esig1, nek);
    T.psig = pw(h.psig0, h.psig1, nek);
    T.ededx = pw(h.ededx0, h.ededx1, nek);
    T.pdedx = pw(h.pdedx0, h.pdedx1, nek);
    T.ebr1 = pw(h.ebr10, h.ebr11, nek);
    T.pbr1 = pw(h.pbr10, h.pbr11, nek);
    T.pbr2 = pw(h.pbr20, h.pbr21, nek);
    T.tmxs = pw(h.tmxs0, h.tmxs1, nek);
    T.blcce = pw(h.blcce0, h.blcce1, nek);
    T.etae = pw(h.etae_ms0, h.etae_ms1, nek);
    T.etap = pw(h.etap_ms0, h.etap_ms1, nek);
    T.q1ce = pw(h.q1ce_ms0, h.q1ce_ms1, nek);
    T.q1cp = pw(h.q1cp_ms0, h.q1cp_ms1, nek);
    T.q2ce = pw(h.q2ce_ms0, h.q2ce_ms1, nek);
    T.q2cp = pw(h.q2cp_ms0, h.q2cp_ms1, nek);
    T.range_ep = vec(h.range_ep, 2 * nek);
    T.e_array = vec(h.e_array, nek);
    T.spin_rej = vec(h.spin_rej, (size_t)nm * OMC_NSPIN_MED + 1);   // + the zero entry of quirk Q17
    // the coefficients of each charged bin interleaved, [2][nmed][stride]
    // with stride = the largest MEKE + 1 (bins 0 .. MEKE of every medium)
    int stride = 0;
    for (int i = 0; i < nm; i++) stride = std::max(stride, std::min(h.meke[i] + 1, OMC_MXEKE));
    std::vector<CBin<R>> cb((size_t)2 * nm * stride);
    for (int qel = 0; qel < 2; qel++)
        for (int i = 0; i < nm; i++)
            for (int k = 0; k < stride; k++) {
                const size_t s = (size_t)i * OMC_MXEKE + k;
                CBin<R> &b = cb[((size_t)qel * nm + i) * stride + k];
                memset(&b, 0, sizeof b);
                auto set = [&](PW<R> &q, const double *c0, const double *c1) { q.c1 = (R)c1[s]; q.c0 = (R)c0[s]; };
                set(b.dedx, qel ? h.pdedx0 : h.ededx0, qel ? h.pdedx1 : h.ededx1);
                set(b.sig, qel ? h.psig0 : h.esig0, qel ? h.psig1 : h.esig1);
                set(b.tmxs, h.tmxs0, h.tmxs1);
                set(b.eta, qel ? h.etap_ms0 : h.etae_ms0, qel ? h.etap_ms1 : h.etae_ms1);
                set(b.blcc, h.blcce0, h.blcce1);
                set(b.q1c, qel ? h.q1cp_ms0 : h.q1ce_ms0, qel ? h.q1cp_ms1 : h.q1ce_ms1);
                set(b.q2c, qel ? h.q2cp_ms0 : h.q2ce_ms0, qel ? h.q2cp_ms1 : h.q2ce_ms1);
                set(b.br1, qel ? h.pbr10 : h.ebr10, qel ? h.pbr11 : h.ebr11);
                if (qel) set(b.br2, h.pbr20, h.pbr21);
                b.e_lo = (R)h.e_array[s];
                b.range = (R)h.range_ep[(size_t)qel * nek + s];
            }
    T.cbin = A.upload(cb);
    T.cb_stride = stride;
    std::vector<MsBin<R>> msb(OMC_NMS);
    for (size_t n = 0; n < (size_t)OMC_NMS; n++) msb[n] = MsBin<R>{(R)h.wms[n], h.ims[n], (R)h.fms[n], (R)h.ums[n]};
    T.ms = A.upload(msb);
    T.b2spin_min = (R)h.b2spin_min; T.dbeta2i = (R)h.dbeta2i; T.espml = (R)h.espml; T.dleneri = (R)h.dleneri;
    T.dqq1i = (R)h.dqq1i; T.dllambi = (R)h.dllambi; T.dqmsi = (R)h.dqmsi;
    if (!A.ok) return omc_seterr(OMC_ERR_DEVICE, "table upload failed: %s", hipGetErrorString(hipGetLastError()));
    return OMC_OK;
}

extern "C" int omc_ctx_create(omc_ctx **out, int nmed, const char *const *media, const char *pegs_file,
                              const char *pgs4form_file, const char *data_folder, int device) {
    if (!out || !media || nmed < 1 || nmed > OMC_MXMED) return omc_seterr(OMC_ERR_ARG, "omc_ctx_create: bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return omc_seterr(OMC_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return omc_seterr(OMC_ERR_ARG, "device %d of %d", device, ndev);
    omc_ctx *c = new omc_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return omc_seterr(OMC_ERR_DEVICE, "hipSetDevice(%d)", device); }
    int rc0 = omc_noalloc_init(device);      // the fixed scratch array of omc_run_dosxyz_noalloc on this device
    if (rc0) { delete c; return rc0; }
    char names[OMC_MXMED][60];
    for (int i = 0; i < nmed; i++) snprintf(names[i], 60, "%s", media[i]);
    omc_paths p = {pegs_file, pgs4form_file, data_folder};
    int rc = omc_tables_build(&c->host, nmed, (const char(*)[60])names, &p);
    if (rc) { delete c; return rc; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        omc_tables_free(&c->host);
        delete c;
        return omc_seterr(OMC_ERR_DEVICE, "stream creation failed");
    }
    rc = upload_tables<float>(c, c->tabmem, c->t32);
    if (!rc) rc = upload_tables<double>(c, c->tabmem, c->t64);
    if (!rc) {
        // the descriptors themselves live in device memory: kernels take
        // pointers to them (no per-lane copy of a by-value argument)
        c->dt32 = c->tabmem.upload(std::vector<Tables<float>>{c->t32});
        c->dt64 = c->tabmem.upload(std::vector<Tables<double>>{c->t64});
        if (!c->tabmem.ok) rc = omc_seterr(OMC_ERR_NOMEM, "table descriptors");
    }
    if (rc) { omc_ctx_destroy(c); return rc; }
    *out = c;
    return OMC_OK;
}

extern "C" void omc_ctx_destroy(omc_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (auto &kv : c->phantoms) {
        kv.second->mem.release();
        omc_phantom_free(&kv.second->host);
        delete kv.second;
    }
    c->phantoms.clear();
    c->tabmem.release();
    c->scratch.release();
    devcache_trim();          // blocks cached by this context's runs (none of them in use)
    if (c->stream) (void)hipStreamDestroy(c->stream);
    omc_tables_free(&c->host);
    delete c;
}

extern "C" int omc_ctx_table(omc_ctx *c, const char *name, const double **data, long *n) {
    if (!c || !name || !data || !n) return omc_seterr(OMC_ERR_ARG, "omc_ctx_table: bad arguments");
    const omc_tables &h = c->host;
    long nm = h.nmed;
    struct E { const char *nm; const double *p; long n; };
    const E tab[] = {
        {"ge0", h.ge0, nm}, {"ge1", h.ge1, nm},
        {"gmfp0", h.gmfp0, nm * OMC_MXGE}, {"gmfp1", h.gmfp1, nm * OMC_MXGE},
        {"gbr10", h.gbr10, nm * OMC_MXGE}, {"gbr11", h.gbr11, nm * OMC_MXGE},
        {"gbr20", h.gbr20, nm * OMC_MXGE}, {"gbr21", h.gbr21, nm * OMC_MXGE},
        {"cohe0", h.cohe0, nm * OMC_MXGE}, {"cohe1", h.cohe1, nm * OMC_MXGE},
        {"ray_xgrid", h.ray_xgrid, nm * OMC_MXRAYFF}, {"ray_fcum", h.ray_fcum, nm * OMC_MXRAYFF},
        {"ray_b", h.ray_b, nm * OMC_MXRAYFF}, {"ray_c", h.ray_c, nm * OMC_MXRAYFF},
        {"ray_pmax0", h.ray_pmax0, nm * OMC_MXGE}, {"ray_pmax1", h.ray_pmax1, nm * OMC_MXGE},
        {"dl1", h.dl1, nm * 8}, {"dl2", h.dl2, nm * 8}, {"dl3", h.dl3, nm * 8},
        {"dl4", h.dl4, nm * 8}, {"dl5", h.dl5, nm * 8}, {"dl6", h.dl6, nm * 8},
        {"bpar0", h.bpar0, nm}, {"bpar1", h.bpar1, nm}, {"pair_delcm", h.delcm, nm}, {"zbrang", h.zbrang, nm},
        {"esig0", h.esig0, nm * OMC_MXEKE}, {"esig1", h.esig1, nm * OMC_MXEKE},
        {"psig0", h.psig0, nm * OMC_MXEKE}, {"psig1", h.psig1, nm * OMC_MXEKE},
        {"ededx0", h.ededx0, nm * OMC_MXEKE}, {"ededx1", h.ededx1, nm * OMC_MXEKE},
        {"pdedx0", h.pdedx0, nm * OMC_MXEKE}, {"pdedx1", h.pdedx1, nm * OMC_MXEKE},
        {"ebr10", h.ebr10, nm * OMC_MXEKE}, {"ebr11", h.ebr11, nm * OMC_MXEKE},
        {"pbr10", h.pbr10, nm * OMC_MXEKE}, {"pbr11", h.pbr11, nm * OMC_MXEKE},
        {"pbr20", h.pbr20, nm * OMC_MXEKE}, {"pbr21", h.pbr21, nm * OMC_MXEKE},
        {"tmxs0", h.tmxs0, nm * OMC_MXEKE}, {"tmxs1", h.tmxs1, nm * OMC_MXEKE},
        {"blcce0", h.blcce0, nm * OMC_MXEKE}, {"blcce1", h.blcce1, nm * OMC_MXEKE},
        {"etae_ms0", h.etae_ms0, nm * OMC_MXEKE}, {"etae_ms1", h.etae_ms1, nm * OMC_MXEKE},
        {"etap_ms0", h.etap_ms0, nm * OMC_MXEKE}, {"etap_ms1", h.etap_ms1, nm * OMC_MXEKE},
        {"q1ce_ms0", h.q1ce_ms0, nm * OMC_MXEKE}, {"q1ce_ms1", h.q1ce_ms1, nm * OMC_MXEKE},
        {"q1cp_ms0", h.q1cp_ms0, nm * OMC_MXEKE}, {"q1cp_ms1", h.q1cp_ms1, nm * OMC_MXEKE},
        {"q2ce_ms0", h.q2ce_ms0, nm * OMC_MXEKE}, {"q2ce_ms1", h.q2ce_ms1, nm * OMC_MXEKE},
        {"q2cp_ms0", h.q2cp_ms0, nm * OMC_MXEKE}, {"q2cp_ms1", h.q2cp_ms1, nm * OMC_MXEKE},
        {"e_array", h.e_array, nm * OMC_MXEKE}, {"range_ep", h.range_ep, 2 * nm * OMC_MXEKE},
        {"eke0", h.eke0, nm}, {"eke1", h.eke1, nm}, {"esig_e", h.esig_e, nm}, {"psig_e", h.psig_e, nm},
        {"xcc", h.xcc, nm}, {"blcc", h.blcc, nm}, {"expeke1", h.expeke1, nm},
        {"spin_rej", h.spin_rej, nm * OMC_NSPIN_MED},
        {"ums", h.ums, OMC_NMS}, {"fms", h.fms, OMC_NMS}, {"wms", h.wms, OMC_NMS},
        {"rho", h.rho, nm}, {"ae", h.ae, nm}, {"ap", h.ap, nm}, {"ue", h.ue, nm}, {"up", h.up, nm},
        {"te", h.te, nm}, {"thmoll", h.thmoll, nm}, {"rlc", h.rlc, nm}, {"delcm", h.delcm, nm},
    };
    for (const E &e : tab)
        if (!strcmp(e.nm, name)) { *data = e.p; *n = e.n; return OMC_OK; }
    return omc_seterr(OMC_ERR_ARG, "no table named %s", name);
}

// ------------------------------------------------------------------ phantoms
template <class R> static void fill_geom(omc_ctx *c, PhantomDev *pd, Geom<R> &G) {
    const omc_phantom &ph = pd->host;
    const omc_tables &h = c->host;
    long nvox = (long)ph.nx * ph.ny * ph.nz;
    G.nx = ph.nx; G.ny = ph.ny; G.nz = ph.nz;
    auto vec = [&](const double *p, size_t n) {
        std::vector<R> v(n);
        for (size_t k = 0; k < n; k++) v[k] = (R)p[k];
        return pd->mem.upload(v);
    };
    G.xb = vec(ph.xb, ph.nx + 1);
    G.yb = vec(ph.yb, ph.ny + 1);
    G.zb = vec(ph.zb, ph.nz + 1);
    std::vector<R> rhof(nvox);
    for (long v = 0; v < nvox; v++) {
        int im = ph.med[v] - 1;
        rhof[v] = im < 0 ? (R)0 : (ph.rho[v] == 0.0F ? (R)1.0 : (R)(ph.rho[v] / h.rho[im]));
    }
    G.rhof = pd->mem.upload(rhof);
    G.med = pd->dmed;
    for (int m = 0; m < OMC_MXMED; m++) {
        G.pcut[m] = m < h.nmed ? (R)(h.ap[m] <= pd->pcut ? pd->pcut : h.ap[m]) : (R)0;
        G.ecut[m] = m < h.nmed ? (R)(h.ae[m] <= pd->ecut ? pd->ecut : h.ae[m]) : (R)0;
        G.rhof_max[m] = (R)0;
    }
    // largest density ratio per medium (majorant of the photon delta tracking)
    for (long v = 0; v < nvox; v++) {
        int im = ph.med[v] - 1;
        if (im >= 0 && rhof[v] > G.rhof_max[im]) G.rhof_max[im] = rhof[v];
    }
    // equal spacing per axis: voxel lookup by multiplication
    auto even = [](const double *b, int n, double &inv) {
        double d = (b[n] - b[0]) / n;
        inv = 1.0 / d;
        for (int i = 0; i < n; i++)
            if (fabs((b[i + 1] - b[i]) - d) > 1e-9 * fabs(d)) return false;
        return true;
    };
    double ix, iy, iz;
    // every axis is evaluated: each call also sets that axis' 1 / spacing
    const bool ex = even(ph.xb, ph.nx, ix), ey = even(ph.yb, ph.ny, iy), ez = even(ph.zb, ph.nz, iz);
    G.uniform = ex && ey && ez ? 1 : 0;
    G.inv_dx = (R)ix;
    G.inv_dy = (R)iy;
    G.inv_dz = (R)iz;
}

size_t omc_phantom_bytes(const omc_phantom_desc *d) {
    size_t nvox = (size_t)d->nx * d->ny * d->nz;
    return nvox * (1 + 4 + 8 + 16) + 16 * (size_t)(d->nx + d->ny + d->nz + 3);
}

int omc_phantom_make(omc_ctx *c, const omc_phantom_desc *d, PhantomDev **out, hipStream_t cache_stream) {
    if (!c || !d || !d->xb || !d->yb || !d->zb || !d->med || !d->rho || d->nx < 1 || d->ny < 1 || d->nz < 1)
        return omc_seterr(OMC_ERR_ARG, "omc_phantom_upload: bad descriptor");
    long nvox = (long)d->nx * d->ny * d->nz;
    for (long v = 0; v < nvox; v++)
        if (d->med[v] < 0 || d->med[v] > c->host.nmed)
            return omc_seterr(OMC_ERR_FORMAT, "voxel %ld has medium %d (context has %d)", v, d->med[v], c->host.nmed);
    PhantomDev *pd = new PhantomDev();
    if (cache_stream) pd->mem.use_cache(cache_stream);
    omc_phantom &ph = pd->host;
    ph.nmed = c->host.nmed;
    ph.nx = d->nx; ph.ny = d->ny; ph.nz = d->nz;
    ph.owns = 1;
    ph.xb = (double *)malloc(sizeof(double) * (d->nx + 1));
    ph.yb = (double *)malloc(sizeof(double) * (d->ny + 1));
    ph.zb = (double *)malloc(sizeof(double) * (d->nz + 1));
    ph.med = (int *)malloc(sizeof(int) * nvox);
    ph.rho = (double *)malloc(sizeof(double) * nvox);
    memcpy(ph.xb, d->xb, sizeof(double) * (d->nx + 1));
    memcpy(ph.yb, d->yb, sizeof(double) * (d->ny + 1));
    memcpy(ph.zb, d->zb, sizeof(double) * (d->nz + 1));
    memcpy(ph.med, d->med, sizeof(int) * nvox);
    memcpy(ph.rho, d->rho, sizeof(double) * nvox);
    pd->ecut = d->ecut;
    pd->pcut = d->pcut;
    pd->nreg = nvox + 1;
    std::vector<uint8_t> m8(nvox);
    for (long v = 0; v < nvox; v++) m8[v] = d->med[v] == 0 ? VACUUM : (uint8_t)(d->med[v] - 1);
    pd->dmed = pd->mem.upload(m8);
    fill_geom<float>(c, pd, pd->g32);
    fill_geom<double>(c, pd, pd->g64);
    pd->dg32 = pd->mem.upload(std::vector<Geom<float>>{pd->g32});
    pd->dg64 = pd->mem.upload(std::vector<Geom<double>>{pd->g64});
    {
        // voxel masses as the reference's accumulateResults forms them (omc_matrad.c:944-950)
        std::vector<double> mass(nvox), rho(ph.rho, ph.rho + nvox);
        for (int iz = 0; iz < ph.nz; iz++)
            for (int iy = 0; iy < ph.ny; iy++)
                for (int ix = 0; ix < ph.nx; ix++) {
                    long v = ix + (long)iy * ph.nx + (long)iz * ph.nx * ph.ny;
                    double m = (ph.xb[ix + 1] - ph.xb[ix]) * (ph.yb[iy + 1] - ph.yb[iy]) * (ph.zb[iz + 1] - ph.zb[iz]);
                    m *= ph.rho[v];
                    mass[v] = m;
                }
        pd->dmass = pd->mem.upload(mass);
        pd->drho = pd->mem.upload(rho);
    }
    if (!pd->mem.ok) {
        pd->mem.release();
        omc_phantom_free(&pd->host);
        delete pd;
        return omc_seterr(OMC_ERR_NOMEM, "phantom upload failed (%ld voxels)", nvox);
    }
    *out = pd;
    return OMC_OK;
}

void omc_phantom_destroy(PhantomDev *pd) {
    if (!pd) return;
    pd->mem.release();
    omc_phantom_free(&pd->host);
    delete pd;
}

extern "C" int omc_phantom_upload(omc_ctx *c, const omc_phantom_desc *d, int *id) {
    if (!id) return omc_seterr(OMC_ERR_ARG, "omc_phantom_upload: null id");
    (void)hipSetDevice(c->device);
    PhantomDev *pd = nullptr;
    int rc = omc_phantom_make(c, d, &pd);
    if (rc) return rc;
    *id = c->next_id++;
    c->phantoms[*id] = pd;
    return OMC_OK;
}

extern "C" int omc_phantom_release(omc_ctx *c, int id) {
    auto it = c ? c->phantoms.find(id) : decltype(c->phantoms.end())();
    if (!c || it == c->phantoms.end()) return omc_seterr(OMC_ERR_ARG, "no phantom %d", id);
    omc_phantom_destroy(it->second);
    c->phantoms.erase(it);
    return OMC_OK;
}

PhantomDev *omc_get_phantom(omc_ctx *c, int id) {
    auto it = c->phantoms.find(id);
    return it == c->phantoms.end() ? nullptr : it->second;
}

// ------------------------------------------------------------------ runs
extern "C" int omc_spectrum_file(const char *path, double *cdfinv1, double *cdfinv2) {
    if (!path || !cdfinv1 || !cdfinv2) return omc_seterr(OMC_ERR_ARG, "omc_spectrum_file: bad arguments");
    omc_spectrum sp;
    int rc = omc_spectrum_read(&sp, path);
    if (rc) return rc;
    memcpy(cdfinv1, sp.cdfinv1, sizeof(double) * OMC_INVDIM);
    memcpy(cdfinv2, sp.cdfinv2, sizeof(double) * OMC_INVDIM);
    return OMC_OK;
}

extern "C" void omc_run_opts_default(omc_run_opts *o) {
    memset(o, 0, sizeof *o);
    o->precision = OMC_FP32;
    o->form = OMC_FORM_EVENT;
    o->rng = OMC_RNG_STREAMS;
    o->seed = 1802;
    o->nsplit = 1;
}

int omc_check_opts(const omc_run_opts *o) {
    if (!o) return omc_seterr(OMC_ERR_ARG, "null run options");
    if (o->precision != OMC_FP32 && o->precision != OMC_FP64) return omc_seterr(OMC_ERR_ARG, "precision %d", o->precision);
    if (o->nsplit < 1 || o->nsplit > 64) return omc_seterr(OMC_ERR_ARG, "nsplit %d (1..64)", o->nsplit);
    if (o->form != OMC_FORM_HISTORY && o->form != OMC_FORM_EVENT) return omc_seterr(OMC_ERR_ARG, "form %d", o->form);
    if (o->rng == OMC_RNG_REPLAY && (!o->replay || !o->replay_offsets))
        return omc_seterr(OMC_ERR_ARG, "replay random numbers requested without a stream");
    if (o->rng == OMC_RNG_REPLAY && o->form != OMC_FORM_HISTORY)
        return omc_seterr(OMC_ERR_ARG, "replayed random numbers need the history form (one stream per history)");
    if (o->event_target < 0 || o->electron_budget < 0 || o->group < 0 || o->hbm_budget < 0 || o->grid < 0 ||
        o->block < 0 || o->batches_per_job < 0 || o->max_particle_steps < 0)
        return omc_seterr(OMC_ERR_ARG, "negative launch option");
    if (o->electron_turns < -1) return omc_seterr(OMC_ERR_ARG, "electron_turns %d (-1 no limit, 0 automatic)", o->electron_turns);
    if (o->streams < 0 || o->streams > 4) return omc_seterr(OMC_ERR_ARG, "streams %d (0 automatic, 1..4)", o->streams);
    if (o->graph != 0 && o->graph != -1) return omc_seterr(OMC_ERR_ARG, "graph %d (0 graphs, -1 eager launches)", o->graph);
    if (o->lds_tables != 0 && o->lds_tables != -1)
        return omc_seterr(OMC_ERR_ARG, "lds_tables %d (0 staged in LDS, -1 read from device memory)", o->lds_tables);
    return OMC_OK;
}

template <class R>
static int run_dosxyz_history(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_dosxyz_source_desc *src,
                              int64_t nper, int nbatch, const omc_run_opts *o, double *sum, double *sum2,
                              omc_run_stats *st) {
    int rc = OMC_OK;
    DevArena A;
    A.use_cache(stream);
    HistArgs<R> a;
    memset(&a, 0, sizeof a);
    a.t = tables_of<R>(c);
    a.g = geom_of<R>(pd);
    long nreg = pd->nreg;
    double *d_e = A.alloc_n<double>(nreg), *d_s = A.alloc_n<double>(nreg), *d_s2 = A.alloc_n<double>(nreg);
    unsigned long long *d_cnt = A.alloc_n<unsigned long long>(8), *d_next = A.alloc_n<unsigned long long>(1);
    int *d_err = A.alloc_n<int>(4);
    double *d_ebal = A.alloc_n<double>(EB_N);
    uint32_t *d_rn = nullptr;
    long *d_off = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> kev;
    float ms = 0.f;
    const int hblock = o->block > 0 ? o->block : 256;
    // the largest launch of the run (the particle stacks of nsplit > 8 cover it)
    int max_grid = o->grid > 0 ? o->grid : (int)std::min<int64_t>((nper + hblock - 1) / hblock, 2048);
    int err[4] = {0, 0, 0, 0};
    auto wall0 = std::chrono::steady_clock::now();
    if (!A.ok) { rc = omc_seterr(OMC_ERR_NOMEM, "device allocation failed"); goto done; }
    rc = make_dosxyz_src<R>(A, pd, src, a.dos);
    if (rc) goto done;
    if (o->rng == OMC_RNG_REPLAY) {
        long nh = (long)nper * nbatch;
        d_rn = A.alloc_n<uint32_t>(o->replay_len);
        d_off = A.alloc_n<long>(nh + 1);
        if (!A.ok) { rc = omc_seterr(OMC_ERR_NOMEM, "replay stream allocation failed"); goto done; }
        HIPCHK(hipMemcpy(d_rn, o->replay, sizeof(uint32_t) * o->replay_len, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_off, o->replay_offsets, sizeof(long) * (nh + 1), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemsetAsync(d_e, 0, sizeof(double) * nreg, stream));
    HIPCHK(hipMemsetAsync(d_s, 0, sizeof(double) * nreg, stream));
    HIPCHK(hipMemsetAsync(d_s2, 0, sizeof(double) * nreg, stream));
    HIPCHK(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long) * 8, stream));
    HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int) * 4, stream));
    HIPCHK(hipMemsetAsync(d_ebal, 0, sizeof(double) * EB_N, stream));
    a.endep = d_e;
    a.ebal = d_ebal;
    a.counters = d_cnt;
    a.err = d_err;
    a.next = d_next;
    a.seed = o->seed;
    a.replay = d_rn;
    a.chunk = 1;
    a.max_steps = hist_max_steps(o);
    a.nsplit = o->nsplit;
    rc = alloc_split_stacks<R>(A, a, o->nsplit, hblock, max_grid);
    if (rc) goto done;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    if (o->count_events) {
        kev.assign(2 * nbatch, nullptr);
        for (auto &e : kev) HIPCHK(hipEventCreate(&e));
    }
    HIPCHK(hipEventRecord(ev0, stream));
    for (int b = 0; b < nbatch; b++) {
        a.nhist = nper;
        a.hist_base = (long)b * nper;
        a.replay_off = d_off ? d_off + (long)b * nper : nullptr;
        HIPCHK(hipMemsetAsync(d_next, 0, sizeof(unsigned long long), stream));
        int block = hblock;
        long want = (nper + block - 1) / block;
        int grid = o->grid > 0 ? o->grid : (int)(want < 2048 ? want : 2048);
        if (grid > max_grid) grid = max_grid;
        if (o->count_events) HIPCHK(hipEventRecord(kev[2 * b], stream));
        HIPCHK(launch_history<R>(a, o->rng == OMC_RNG_REPLAY ? RNG_REPLAY : RNG_STREAMS, (o->count_events & 1) != 0,
                                 grid, block, stream));
        if (o->count_events) HIPCHK(hipEventRecord(kev[2 * b + 1], stream));
        HIPCHK(launch_accumulate(d_e, d_s, d_s2, nreg, 1.0, d_ebal + EB_DEPOSIT, stream));
    }
    HIPCHK(hipEventRecord(ev1, stream));
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    HIPCHK(hipMemcpy(sum, d_s, sizeof(double) * nreg, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sum2, d_s2, sizeof(double) * nreg, hipMemcpyDeviceToHost));
    // the device's error flags decide the result whether or not statistics
    // were asked for: a history cut short leaves a biased dose
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
        st->gpu_seconds = ms * 1e-3;
        st->histories = (double)nper * nbatch;
        st->n_howfar = cnt[0]; st->n_hownear = cnt[1]; st->n_cross = cnt[2]; st->n_dep = cnt[3];
        st->n_rn = cnt[4]; st->n_photon = cnt[5]; st->n_electron = cnt[6];
        st->errors = (err[0] ? 1 : 0) | (err[1] ? 2 : 0) | (err[2] ? 4 : 0);
        double tk = 0;
        for (int b = 0; b < nbatch && o->count_events; b++) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, kev[2 * b], kev[2 * b + 1]));
            tk += t * 1e-3;
        }
        st->kernel_seconds[OMC_K_HISTORY] = o->count_events ? tk : ms * 1e-3;
        st->kernel_launches[OMC_K_HISTORY] = (uint64_t)nbatch;
    }
done:
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (auto e : kev)
        if (e) (void)hipEventDestroy(e);
    A.release();
    if (st) st->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    return rc;
}

int omc_dosxyz_run(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_dosxyz_source_desc *src,
                   int64_t nperbatch, int nbatch, const omc_run_opts *o, double *sum, double *sum2, omc_run_stats *st) {
    if (!c || !pd || !src || !sum || !sum2 || nperbatch < 1 || nbatch < 1)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz: bad arguments");
    int rc = omc_check_opts(o);
    if (rc) return rc;
    (void)hipSetDevice(c->device);
    if (st) memset(st, 0, sizeof *st);
    if (o->form == OMC_FORM_EVENT)
        return o->precision == OMC_FP32
                   ? omc_run_dosxyz_event<float>(c, stream, pd, src, nperbatch, nbatch, o, sum, sum2, st)
                   : omc_run_dosxyz_event<double>(c, stream, pd, src, nperbatch, nbatch, o, sum, sum2, st);
    return o->precision == OMC_FP32 ? run_dosxyz_history<float>(c, stream, pd, src, nperbatch, nbatch, o, sum, sum2, st)
                                    : run_dosxyz_history<double>(c, stream, pd, src, nperbatch, nbatch, o, sum, sum2, st);
}

extern "C" int omc_run_dosxyz(omc_ctx *c, int phantom_id, const omc_dosxyz_source_desc *src, int64_t nperbatch,
                              int nbatch, const omc_run_opts *o, double *sum, double *sum2, omc_run_stats *st) {
    if (!c) return omc_seterr(OMC_ERR_ARG, "omc_run_dosxyz: null context");
    PhantomDev *pd = omc_get_phantom(c, phantom_id);
    if (!pd) return omc_seterr(OMC_ERR_ARG, "no phantom %d", phantom_id);
    return omc_dosxyz_run(c, c->stream, pd, src, nperbatch, nbatch, o, sum, sum2, st);
}

extern "C" int omc_dosxyz_dose(omc_ctx *c, int phantom_id, const double *sum, const double *sum2, int64_t nperbatch,
                               int nbatch, double *dose, double *unc) {
    PhantomDev *pd = c ? omc_get_phantom(c, phantom_id) : nullptr;
    if (!pd || !sum || !sum2 || !dose || !unc) return omc_seterr(OMC_ERR_ARG, "omc_dosxyz_dose: bad arguments");
    omc_dosxyz_finish(&pd->host, sum, sum2, (int)nperbatch, nbatch, dose, unc);
    return OMC_OK;
}

extern "C" int omc_write_3ddose_file(omc_ctx *c, int phantom_id, const char *path, const double *dose, const double *unc) {
    PhantomDev *pd = c ? omc_get_phantom(c, phantom_id) : nullptr;
    if (!pd || !path) return omc_seterr(OMC_ERR_ARG, "omc_write_3ddose_file: bad arguments");
    return omc_write_3ddose(path, &pd->host, dose, unc);
}

// ------------------------------------------------------------------ sampler parity
template <class R>
static int sampler_eval(omc_ctx *c, PhantomDev *pd, int ncase, const int *kind, const double *in, int in_stride,
                        const uint32_t *rn, const int64_t *rn_off, double *out, int out_stride, int64_t *rn_used) {
    int rc = OMC_OK;
    DevArena A;
    SampArgs<R> a;
    memset(&a, 0, sizeof a);
    a.t = tables_of<R>(c);
    a.g = geom_of<R>(pd);
    a.ncase = ncase;
    long nrn = rn_off[ncase];
    int *d_kind = A.alloc_n<int>(ncase);
    double *d_in = A.alloc_n<double>((size_t)ncase * in_stride);
    uint32_t *d_rn = A.alloc_n<uint32_t>(nrn > 0 ? nrn : 1);
    long *d_off = A.alloc_n<long>(ncase + 1);
    double *d_out = A.alloc_n<double>((size_t)ncase * out_stride);
    long *d_used = A.alloc_n<long>(ncase);
    int *d_err = A.alloc_n<int>(1);
    if (!A.ok) { rc = omc_seterr(OMC_ERR_NOMEM, "sampler buffers"); goto done; }
    HIPCHK(hipMemcpy(d_kind, kind, sizeof(int) * ncase, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_in, in, sizeof(double) * ncase * in_stride, hipMemcpyHostToDevice));
    if (nrn > 0) HIPCHK(hipMemcpy(d_rn, rn, sizeof(uint32_t) * nrn, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_off, rn_off, sizeof(long) * (ncase + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(d_out, 0, sizeof(double) * ncase * out_stride));
    HIPCHK(hipMemset(d_err, 0, sizeof(int)));
    a.kind = d_kind; a.in = d_in; a.in_stride = in_stride; a.rn = d_rn; a.rn_off = d_off;
    a.out = d_out; a.out_stride = out_stride; a.rn_used = d_used; a.err = d_err;
    HIPCHK(launch_sampler<R>(a, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, d_out, sizeof(double) * ncase * out_stride, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rn_used, d_used, sizeof(long) * ncase, hipMemcpyDeviceToHost));
done:
    A.release();
    return rc;
}

extern "C" int omc_sampler_eval(omc_ctx *c, int phantom_id, int precision, int ncase, const int *kind, const double *in,
                                int in_stride, const uint32_t *rn, const int64_t *rn_off, double *out, int out_stride,
                                int64_t *rn_used) {
    PhantomDev *pd = c ? omc_get_phantom(c, phantom_id) : nullptr;
    if (!pd || ncase < 1 || !kind || !in || !rn_off || !out || !rn_used || in_stride < 17 || out_stride < 64)
        return omc_seterr(OMC_ERR_ARG, "omc_sampler_eval: bad arguments");
    (void)hipSetDevice(c->device);
    return precision == OMC_FP64 ? sampler_eval<double>(c, pd, ncase, kind, in, in_stride, rn, rn_off, out, out_stride, rn_used)
                                 : sampler_eval<float>(c, pd, ncase, kind, in, in_stride, rn, rn_off, out, out_stride, rn_used);
}

// ------------------------------------------------------------------ measurement
extern "C" int omc_atomic_probe(int device, int64_t nslot, int64_t nops, double *ops_per_second) {
    if (!ops_per_second || nslot < 1024 || nslot > (1LL << 30) || (nslot & (nslot - 1)) || nops < 1)
        return omc_seterr(OMC_ERR_ARG, "omc_atomic_probe: nslot a power of two in 2^10..2^30, nops > 0");
    int rc = OMC_OK;
    if (hipSetDevice(device) != hipSuccess) return omc_seterr(OMC_ERR_DEVICE, "device %d", device);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    const int grid = cus * 16;
    const long lanes = (long)grid * 256;
    const int nper = (int)std::max<long>(1, (long)((nops + lanes - 1) / lanes));
    DevArena A;
    double *t = A.alloc_n<double>((size_t)nslot);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    if (!A.ok) { rc = omc_seterr(OMC_ERR_NOMEM, "probe table of %lld doubles", (long long)nslot); goto done; }
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemset(t, 0, sizeof(double) * nslot));
    HIPCHK(launch_atomic_probe(t, (unsigned long long)(nslot - 1), grid, nper, 1, nullptr));
    HIPCHK(hipEventRecord(e0, nullptr));
    HIPCHK(launch_atomic_probe(t, (unsigned long long)(nslot - 1), grid, nper, 2, nullptr));
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *ops_per_second = (double)lanes * nper / (ms * 1e-3);
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    A.release();
    return rc;
}
