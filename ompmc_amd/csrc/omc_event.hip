// Host side of the event-queue form, the dose-influence (Dij) runs and the
// request server.
#include "omc_capi_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

using namespace omc;

namespace {

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return omc_seterr(OMC_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

struct ArenaGuard {
    DevArena &a;
    ~ArenaGuard() { a.release(); }
};

struct Events {
    std::vector<hipEvent_t> ev;
    explicit Events(int n) : ev(n, nullptr) {
        for (auto &e : ev) (void)hipEventCreate(&e);
    }
    ~Events() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    float ms(int a, int b) const {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, ev[a], ev[b]);
        return t;
    }
};

// A pinned host buffer of PINNED_SMALL_BYTES from the reuse pool.
struct PinnedSmall {
    void *p = pinned_small_alloc();
    ~PinnedSmall() { pinned_small_free(p); }
};

int device_cus(int dev) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    return n;
}

// ------------------------------------------------------------------ event-form state
// Condensed-history steps between lane refills.  Whole 1960-beamlet plan,
// interleaved-bin build, best of three: 8 -> 17.2, 12 -> 18.2, 16 -> 17.7 M
// histories/s (with 1 M or 2 M in flight, and 16, 24 or 32 turns, all
// within 1% of each other at 12).  With the two-band charged queue, 16 against
// 12: 1960-beamlet plan FP32 22.9 against 22.5 M, FP64 12.4 against 12.6 M;
// 10,044-beamlet plan at 3e8 histories FP32 14.2 against 15.2 M.
constexpr int DEFAULT_BUDGET = 12;
// Refills an electron may hold its lane in one round before it is requeued.
// Measured on the 1960-beamlet plan (1 M in flight): 12 -> 12.0, 16 -> 12.3,
// 24 -> 12.5, 64 -> 11.2, no limit -> 9.7 M histories/s.
constexpr int DEFAULT_TURNS = 24;
// Workgroups per compute unit of the persistent event kernels.  Whole c3
// plan, best of three (M histories/s), charged x photon: 8 x 8 21.9,
// 2 x 8 22.0, 2 x 4 22.2, 4 x 4 21.9 -- within run-to-run noise.
#ifndef OMC_CHARGED_BLOCKS_PER_CU
#define OMC_CHARGED_BLOCKS_PER_CU 8
#endif
#ifndef OMC_PHOTON_BLOCKS_PER_CU
#define OMC_PHOTON_BLOCKS_PER_CU 8
#endif
constexpr int CHARGED_BLOCKS_PER_CU = OMC_CHARGED_BLOCKS_PER_CU;
constexpr int PHOTON_BLOCKS_PER_CU = OMC_PHOTON_BLOCKS_PER_CU;
// LDS a charged-kernel workgroup may give to the electron bins (12 KB per
// medium at 150 bins; two workgroups share a CU's 160 KB).  Plans with more
// media read them from device memory.
constexpr size_t MAX_LDS_TABLES = 64 * 1024;
// Entries of a charged-kernel lane's deposit list (TallyList), 16 bytes each
// for the 256 lanes of a workgroup; the lists are used where bins and lists
// together leave room for two workgroups per CU.  Eight entries measured the
// same as six on the 1960-beamlet plan (22.55 against 22.64 M histories/s,
// three runs each).
constexpr int TALLY_LIST = 6;
constexpr size_t TALLY_LIST_BYTES = (size_t)TALLY_LIST * 256 * 16;
constexpr size_t MAX_LDS_PER_GROUP = 80 * 1024;
// Round guard of a job: the rounds its injection needs plus 2^16 -- with the
// requeue after DEFAULT_TURNS x DEFAULT_BUDGET steps that is ~1.9e7 steps for
// one particle, thousands of times the longest physical track.  A job that
// has not drained by then is reported as a step-guard error.
long max_rounds(long long nhist, long long target) { return 65536L + 8L * (long)(nhist / std::max(target, 1LL)); }

template <class R> size_t pool_entry_bytes() { return sizeof(R) * 10 + sizeof(int) * 6 + 4 + 8 + 16; }

template <class R> size_t event_bytes(long long target) {
    return (size_t)4 * (size_t)(8 * target) * pool_entry_bytes<R>() + sizeof(EvCtl);
}

// Particles in flight: 4096 per compute unit.  On the 1960-beamlet plan the
// rate is flat from 1 M to 2 M in flight (12.3-12.6 M histories/s) and falls
// below (10.1 M at 256 k, 12.0 M at 512 k) and above (11.6 M at 4 M).
// With photon splitting a history carries nsplit times more electrons, and
// rounds are fewer and fuller with four times the particles in flight: at
// nsplit 20 on every 20th beamlet 1.40 s at 1 M, 1.13 s at 2 M, 1.05 s at
// 4 M and 1.04 s at 8 M (15.3 s against 16.3 s on the whole plan).
long long default_target(int dev, int nsplit = 1) {
    return (long long)device_cus(dev) * 4096 * (nsplit > 1 ? 4 : 1);
}

// run_job's code for a pool overflow: the caller zeroes the job's tallies,
// doubles the pools (EvState::grow) and runs the job again.  Every particle's
// random stream is keyed by its history, so the rerun is the same job.
constexpr int RC_POOL_OVERFLOW = -100;
constexpr int MAX_POOL_GROWTH = 5;

template <class R> struct EvState {
    DevArena A;
    Pool<R> pool[2][2];
    EvCtl *ctl = nullptr;
    long long target = 0, cap = 0, capmul = 8;   // pool capacity = capmul x target (peak in flight measured near 4x)
    ~EvState() { A.release(); }
    int grow() {
        A.release();
        capmul *= 2;
        return alloc(target);
    }
    int alloc(long long tgt) {
        target = tgt < 64 ? 64 : tgt;
        cap = capmul * target;
        // the charged pool's counter holds each band's length in 32 bits
        if (cap > (1ll << 31))
            return omc_seterr(OMC_ERR_NOMEM, "event pools for %lld particles exceed 2^31 entries", cap);
        for (int p = 0; p < 2; p++)
            for (int k = 0; k < 2; k++) {
                Pool<R> &P = pool[p][k];
                R **fr[10] = {&P.e, &P.x, &P.y, &P.z, &P.u, &P.v, &P.w, &P.wt, &P.demfp, &P.sig0};
                for (R **f : fr) *f = A.alloc_n<R>(cap);
                int **fi[6] = {&P.ix, &P.iy, &P.iz, &P.iq, &P.phase, &P.slot};
                for (int **f : fi) *f = A.alloc_n<int>(cap);
                P.ctr = A.alloc_n<uint32_t>(cap);
                P.stream = A.alloc_n<uint64_t>(cap);
                P.rs = A.alloc_n<uint4>(cap);
            }
        ctl = A.alloc_n<EvCtl>(1);
        if (!A.ok) return omc_seterr(OMC_ERR_NOMEM, "event pools for %lld particles do not fit", cap);
        return OMC_OK;
    }
};

struct JobStats {
    double t_source = 0, t_photon = 0, t_charged = 0, t_other = 0;
    uint64_t n_source = 0, n_photon = 0, n_charged = 0, n_other = 0;
    uint64_t ev_photon = 0, ev_charged = 0, rounds = 0, peak = 0;
    // With an origin event (recorded before any of the run's work on every
    // stream), the start and end of each timed launch in ms after it, per
    // kernel class (source, photon, charged, round end): launches of
    // concurrent streams overlap, and the union of these intervals is the
    // time the device ran a class at all.
    hipEvent_t origin = nullptr;
    std::vector<std::pair<float, float>> iv[4];
    uint64_t graph_replays = 0, graph_fallbacks = 0;
    uint64_t lds_bytes = 0;          // electron bins the charged kernel staged in LDS per workgroup
};

// Seconds covered by the union of the intervals of one kernel class over
// all given jobs (concurrent launches count once).
double busy_seconds(const std::vector<const JobStats *> &jobs, int k) {
    std::vector<std::pair<float, float>> v;
    for (const JobStats *j : jobs) v.insert(v.end(), j->iv[k].begin(), j->iv[k].end());
    std::sort(v.begin(), v.end());
    double tot = 0, a = 0, b = -1;
    for (const auto &x : v) {
        if (x.first > b) {
            if (b > a) tot += b - a;
            a = x.first;
            b = x.second;
        } else if (x.second > b) {
            b = x.second;
        }
    }
    if (b > a) tot += b - a;
    return tot * 1e-3;
}

// Runs one job (nhist primaries) to completion.  Rounds are enqueued in
// chunks of K; the host reads a snapshot of the control block after each
// chunk while the next chunk runs, so the device never waits for the host.
// With `timing`, every kernel of every round is bracketed by events on the
// job's stream and its duration is added to the job statistics when the
// chunk has completed.  Without it and with L.graph, the first chunk's
// kernels are built into a HIP graph and every chunk of the job replays
// it: a round's kernels depend on the round only through its parity, and
// chunks start at even rounds, so every chunk is the same sequence of
// launches.  A replay is one runtime call instead of 4 K launches.
struct GraphExec {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~GraphExec() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

template <class R>
int run_job(hipStream_t st, EvState<R> &S, EvLaunch<R> &L, long long nhist, int dest, bool count, bool timing,
            JobStats &js) {
    constexpr int K = 4;
    static_assert(K % 2 == 0, "chunks must start at even rounds");
    bool use_graph = L.graph && !timing;
    js.lds_bytes = (uint64_t)L.eb_words * 16u;
    GraphExec gx;
    // With photon splitting one split puts up to 2 x nsplit charged particles
    // into the queues: a round injects at most target / nsplit primaries, so
    // that a round's photons fit their products into the pools.
    const long long inj_max = L.nsplit > 1 ? std::max<long long>(64, S.target / L.nsplit) : 0;
    CK(launch_job_start(S.ctl, nhist, S.target, S.cap, inj_max, dest, st));
    static_assert(2 * sizeof(EvCtl) <= PINNED_SMALL_BYTES, "control-block snapshots exceed a pinned buffer");
    PinnedSmall snapbuf;
    if (!snapbuf.p) return omc_seterr(OMC_ERR_NOMEM, "pinned host buffer");
    EvCtl *snap = (EvCtl *)snapbuf.p;
    Events chunk_done(2);
    Events kev(timing ? 2 * K * 5 : 0);
    auto enqueue = [&](long r0, int slot) -> hipError_t {
        hipError_t x;
        if (use_graph && !gx.x) {
            // the graph is built node by node rather than captured from the
            // stream: the lanes of a Dij run and the requests of a server
            // enqueue, copy and allocate from threads of their own meanwhile,
            // and a capture invalidated by such a call (measured: the
            // synchronous phantom upload of a concurrent request) leaves the
            // stream unusable.  A graph the runtime refuses leaves the job to
            // eager launches.
            bool ok = build_rounds_graph(L, r0, K, count, &gx.g) == hipSuccess &&
                      hipGraphInstantiate(&gx.x, gx.g, nullptr, nullptr, 0) == hipSuccess;
            if (!ok) {
                (void)hipGetLastError();
                use_graph = false;
                js.graph_fallbacks++;
            }
        }
        if (use_graph) {
            if ((x = hipGraphLaunch(gx.x, st)) != hipSuccess) return x;
            js.graph_replays++;
        } else {
            for (int j = 0; j < K; j++) {
                hipEvent_t *e = timing ? &kev.ev[(size_t)(slot * K + j) * 5] : nullptr;
                if (e && (x = hipEventRecord(e[0], st)) != hipSuccess) return x;
                for (int part = 0; part < 4; part++) {
                    if ((x = launch_round_part(L, r0 + j, part, count, st)) != hipSuccess) return x;
                    if (e && (x = hipEventRecord(e[part + 1], st)) != hipSuccess) return x;
                }
            }
        }
        return hipMemcpyAsync(&snap[slot], S.ctl, sizeof(EvCtl), hipMemcpyDeviceToHost, st) == hipSuccess
                   ? hipEventRecord(chunk_done.ev[slot], st)
                   : hipErrorUnknown;
    };
    auto harvest = [&](int slot) {
        if (!timing) return;
        for (int j = 0; j < K; j++) {
            int b = (slot * K + j) * 5;
            js.t_source += kev.ms(b, b + 1) * 1e-3;
            js.t_photon += kev.ms(b + 1, b + 2) * 1e-3;
            js.t_charged += kev.ms(b + 2, b + 3) * 1e-3;
            js.t_other += kev.ms(b + 3, b + 4) * 1e-3;
            if (js.origin) {
                float t[5];
                for (int q = 0; q < 5; q++) (void)hipEventElapsedTime(&t[q], js.origin, kev.ev[b + q]);
                for (int q = 0; q < 4; q++) js.iv[q].emplace_back(t[q], t[q + 1]);
            }
        }
        js.n_source += K; js.n_photon += K; js.n_charged += K; js.n_other += K;
    };
    const long rmax = max_rounds(nhist, inj_max > 0 ? inj_max : S.target);
    long r = 0;
    int cur = 0;
    CK(enqueue(r, 0));
    r += K;
    for (;;) {
        CK(enqueue(r, 1 - cur));
        r += K;
        CK(hipEventSynchronize(chunk_done.ev[cur]));
        harvest(cur);
        if (snap[cur].err || snap[cur].done_round >= 0 || r > rmax) break;
        cur = 1 - cur;
    }
    CK(hipStreamSynchronize(st));
    harvest(1 - cur);
    const EvCtl *done = &snap[1 - cur];   // the last snapshot, taken after every enqueued round
    js.ev_photon += done->events[0];
    js.ev_charged += done->events[1];
    js.rounds += (uint64_t)(done->done_round >= 0 ? done->done_round : done->rounds);
    js.peak = std::max<uint64_t>(js.peak, done->peak);
    if (done->err & 2) {
        static const char *why[] = {"?", "photon flight", "electron pass", "electron step", "rejection loop",
                                    "lane trips", "requeues"};
        const int w = done->drop_why >= 0 && done->drop_why <= 6 ? done->drop_why : 0;
        return omc_seterr(OMC_ERR_LIMIT,
                          "a particle made no progress within its iteration guard and was dropped (first: %s guard, "
                          "source line %d; iq %d, E %.6g MeV, voxel (%d, %d, %d), phase %d, requeues %d)",
                          why[w], done->drop_site, done->drop_iq, (double)done->drop_e, done->drop_ix, done->drop_iy,
                          done->drop_iz, done->drop_phase, done->drop_age);
    }
    if (done->err & 1) {
        omc_seterr(OMC_ERR_STACK, "event pool overflow (target %lld, capacity %lld)", S.target, S.cap);
        return RC_POOL_OVERFLOW;
    }
    if (done->done_round < 0) return omc_seterr(OMC_ERR_LIMIT, "event job did not drain within %ld rounds", rmax);
    return OMC_OK;
}

// Half of the device memory that is free or held for reuse by the block cache.
size_t default_budget() {
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) freeb = (size_t)8 << 30;
    return (freeb + devcache_held()) / 2;
}

// Batches of a dose run transported in one job: o->batches_per_job, else as
// many as a quarter of the memory budget holds tallies for.
int batches_per_job(const omc_run_opts *o, int nbatch, size_t tally_bytes) {
    if (o->batches_per_job > 0) return std::min(o->batches_per_job, nbatch);
    size_t budget = o->hbm_budget > 0 ? (size_t)o->hbm_budget : default_budget();
    size_t n = budget / 4 / std::max<size_t>(tally_bytes, 1);
    return (int)std::max<size_t>(1, std::min<size_t>(n, (size_t)nbatch));
}

template <class R> void fill_launch(omc_ctx *c, const PhantomDev *pd, EvState<R> &S, const omc_run_opts *o,
                                    EvLaunch<R> &L) {
    memset(&L, 0, sizeof L);
    L.t = tables_of<R>(c);
    L.g = geom_of<R>(pd);
    L.seed = o->seed;
    for (int p = 0; p < 2; p++)
        for (int k = 0; k < 2; k++) L.pool[p][k] = S.pool[p][k];
    L.ctl = S.ctl;
    L.cap = S.cap;
    L.budget = o->electron_budget > 0 ? o->electron_budget : DEFAULT_BUDGET;
    L.turns = o->electron_turns > 0 ? o->electron_turns : (o->electron_turns < 0 ? 0 : DEFAULT_TURNS);
    // a particle still going after 2^20 condensed-history steps (a hundred
    // times the longest tracks seen) is taken as stuck
    const long steps = o->max_particle_steps > 0 ? (long)o->max_particle_steps : (1L << 20);
    L.max_age = L.turns > 0 ? (int)std::min<long>(steps / ((long)L.turns * L.budget) + 1, 1L << 22) : 0;
    L.delta = o->voxel_walk ? 0 : 1;
    L.nsplit = o->nsplit;
    L.graph = o->graph >= 0 ? 1 : 0;
    {
        const Tables<R> &ht = host_tables_of<R>(c);
        L.ebin = ht.cbin;
        const size_t bytes = (size_t)ht.nmed * ht.cb_stride * sizeof(CBin<R>);
        if (std::is_same<R, float>::value && o->lds_tables >= 0 && bytes <= MAX_LDS_TABLES)
            L.eb_words = (int)(bytes / 16);
        // FP32: the deposit lists, where they fit beside the bins in two workgroups per CU
        if (std::is_same<R, float>::value && (size_t)L.eb_words * 16 + TALLY_LIST_BYTES <= MAX_LDS_PER_GROUP)
            L.tlist = TALLY_LIST;
    }
    int cus = device_cus(c->device);
    L.grid_event = o->grid > 0 ? o->grid : cus * PHOTON_BLOCKS_PER_CU;
    L.grid_charged = o->grid > 0 ? o->grid : cus * CHARGED_BLOCKS_PER_CU;
    L.grid_source = cus * 4;
    L.src.t = L.t;
    L.src.g = L.g;
    L.src.seed = o->seed;
}

// Runs a job, growing the pools and rerunning it from zeroed tallies
// (nslot slots of nreg doubles at `endep`) when the queues overflow.  The
// energy balance ebal[0..EB_N) (optional) is copied to ebal[EB_N..2 EB_N)
// before the job and restored before a rerun.  Tile marks of the failed
// attempt stay set: they only make the statistics read tiles of zeros.
template <class R>
int run_job_grow(hipStream_t st, EvState<R> &S, EvLaunch<R> &L, long long nhist, int dest, bool count, bool timing,
                 JobStats &js, double *endep, size_t ndouble, double *ebal) {
    if (ebal) CK(hipMemcpyAsync(ebal + EB_N, ebal, sizeof(double) * EB_N, hipMemcpyDeviceToDevice, st));
    for (int attempt = 0;; attempt++) {
        int rc = run_job(st, S, L, nhist, dest, count, timing, js);
        if (rc != RC_POOL_OVERFLOW) return rc;
        if (attempt >= MAX_POOL_GROWTH) return OMC_ERR_STACK;
        CK(hipMemsetAsync(endep, 0, sizeof(double) * ndouble, st));
        if (ebal) CK(hipMemcpyAsync(ebal, ebal + EB_N, sizeof(double) * EB_N, hipMemcpyDeviceToDevice, st));
        CK(hipStreamSynchronize(st));
        rc = S.grow();
        if (rc) return rc;
        for (int p = 0; p < 2; p++)
            for (int k = 0; k < 2; k++) L.pool[p][k] = S.pool[p][k];
        L.ctl = S.ctl;
        L.cap = S.cap;
    }
}

// Union busy time per kernel class of the given jobs (timed runs).
void stats_busy(omc_run_stats *st, const std::vector<const JobStats *> &jobs) {
    static const int cls[4] = {OMC_K_SOURCE, OMC_K_PHOTON, OMC_K_CHARGED, OMC_K_OTHER};
    for (int q = 0; q < 4; q++) st->kernel_busy_seconds[cls[q]] = busy_seconds(jobs, q);
}

void stats_from_job(omc_run_stats *st, const JobStats &js) {
    if (!st) return;
    st->kernel_seconds[OMC_K_SOURCE] += js.t_source;
    st->kernel_seconds[OMC_K_PHOTON] += js.t_photon;
    st->kernel_seconds[OMC_K_CHARGED] += js.t_charged;
    st->kernel_seconds[OMC_K_OTHER] += js.t_other;
    st->kernel_launches[OMC_K_SOURCE] += js.n_source;
    st->kernel_launches[OMC_K_PHOTON] += js.n_photon;
    st->kernel_launches[OMC_K_CHARGED] += js.n_charged;
    st->kernel_launches[OMC_K_OTHER] += js.n_other;
    st->events_photon += js.ev_photon;
    st->events_charged += js.ev_charged;
    st->rounds += js.rounds;
    st->graph_replays += js.graph_replays;
    st->graph_fallbacks += js.graph_fallbacks;
    st->lds_table_bytes = std::max<uint64_t>(st->lds_table_bytes, js.lds_bytes);
    st->peak_inflight = std::max<uint64_t>(st->peak_inflight, js.peak);
}

void stats_energy(omc_run_stats *st, const double *eb) {
    st->e_source += eb[EB_SOURCE];
    st->e_escape += eb[EB_ESCAPE];
    st->e_lost += eb[EB_LOST];
    st->e_deposit += eb[EB_DEPOSIT];
}

void stats_counters(omc_run_stats *st, const unsigned long long *cnt) {
    // cnt[0..7] photon kernel (or history kernel), cnt[8..15] charged kernel
    st->n_howfar = cnt[0] + cnt[8];
    st->n_hownear = cnt[1] + cnt[9];
    st->n_cross = cnt[2] + cnt[10];
    st->n_dep = cnt[3] + cnt[11];
    st->n_rn = cnt[4] + cnt[12];
    st->n_photon = cnt[5] + cnt[13];
    st->n_electron = cnt[6] + cnt[14];
    for (int k = 0; k < 2; k++) {
        st->ev_howfar[k] = cnt[8 * k + 0];
        st->ev_hownear[k] = cnt[8 * k + 1];
        st->ev_cross[k] = cnt[8 * k + 2];
        st->ev_dep[k] = cnt[8 * k + 3];
    }
}

}  // namespace

// ------------------------------------------------------------------ dose runs, event form
template <class R>
int omc_run_dosxyz_event(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_dosxyz_source_desc *src,
                         int64_t nper, int nbatch, const omc_run_opts *o, double *sum, double *sum2,
                         omc_run_stats *st) {
    auto wall0 = std::chrono::steady_clock::now();
    DevArena A;
    ArenaGuard ag{A};
    A.use_cache(stream);
    EvState<R> S;
    S.A.use_cache(stream);
    long nreg = pd->nreg;
    // Batches transported together: each tallies into its own copy of the
    // region array (slot = batch) and the batch statistics are taken in batch
    // order afterwards, so the result is the batch-by-batch one while the
    // job drains its longest tracks once instead of once per batch.
    const int nbj = batches_per_job(o, nbatch, (size_t)nreg * sizeof(double));
    long long tgt = o->event_target > 0 ? o->event_target : default_target(c->device, o->nsplit);
    if (tgt > 4 * (long long)nper * nbj) tgt = std::max<long long>(4 * (long long)nper * nbj, 64);
    int rc = S.alloc(tgt);
    if (rc) return rc;
    double *d_e = A.alloc_n<double>((size_t)nbj * nreg), *d_s = A.alloc_n<double>(nreg),
           *d_s2 = A.alloc_n<double>(nreg);
    long *d_soff = A.alloc_n<long>(nbj + 1);
    unsigned long long *d_cnt = A.alloc_n<unsigned long long>(16);
    double *d_ebal = A.alloc_n<double>(2 * EB_N);
    if (!A.ok) return omc_seterr(OMC_ERR_NOMEM, "device allocation failed");
    std::vector<long> soff(nbj + 1);
    for (int k = 0; k <= nbj; k++) soff[k] = (long)k * nper;
    CK(hipMemcpy(d_soff, soff.data(), sizeof(long) * (nbj + 1), hipMemcpyHostToDevice));
    EvLaunch<R> L;
    fill_launch(c, pd, S, o, L);
    rc = make_dosxyz_src<R>(A, pd, src, L.src.dos);
    if (rc) return rc;
    L.src.use_beamlets = 0;
    L.src.slot_off = d_soff;
    L.endep = d_e;
    L.slot_stride = nreg;
    L.ebal = d_ebal;
    L.counters = o->count_events ? d_cnt : nullptr;
    CK(hipMemsetAsync(d_ebal, 0, sizeof(double) * 2 * EB_N, stream));
    CK(hipMemsetAsync(d_e, 0, sizeof(double) * nbj * nreg, stream));
    CK(hipMemsetAsync(d_s, 0, sizeof(double) * nreg, stream));
    CK(hipMemsetAsync(d_s2, 0, sizeof(double) * nreg, stream));
    CK(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long) * 16, stream));
    int dest = L.src.dos.charge != 0 ? 1 : 0;
    JobStats js;
    Events ev(2);
    CK(hipEventRecord(ev.ev[0], stream));
    js.origin = ev.ev[0];
    for (int b0 = 0; b0 < nbatch; b0 += nbj) {
        const int nb = std::min(nbj, nbatch - b0);
        L.src.hist_base = (long)b0 * nper;       // streams stay keyed by the history's place in the run
        L.src.nslot = nb;
        rc = run_job_grow(stream, S, L, (long long)nper * nb, dest, (o->count_events & 1) != 0,
                          o->count_events != 0, js, d_e, (size_t)nb * nreg, d_ebal);
        if (rc) return rc;
        for (int k = 0; k < nb; k++)
            CK(launch_accumulate(d_e + (size_t)k * nreg, d_s, d_s2, nreg, 1.0, d_ebal + EB_DEPOSIT, stream));
    }
    CK(hipEventRecord(ev.ev[1], stream));
    CK(hipStreamSynchronize(stream));
    CK(hipMemcpy(sum, d_s, sizeof(double) * nreg, hipMemcpyDeviceToHost));
    CK(hipMemcpy(sum2, d_s2, sizeof(double) * nreg, hipMemcpyDeviceToHost));
    if (st) {
        unsigned long long cnt[16];
        double eb[EB_N];
        CK(hipMemcpy(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
        CK(hipMemcpy(eb, d_ebal, sizeof eb, hipMemcpyDeviceToHost));
        stats_energy(st, eb);
        st->gpu_seconds = ev.ms(0, 1) * 1e-3;
        st->histories = (double)nper * nbatch;
        st->batches_per_job = nbj;
        stats_counters(st, cnt);
        stats_from_job(st, js);
        if (o->count_events) stats_busy(st, {&js});
        st->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    }
    return OMC_OK;
}

template int omc_run_dosxyz_event<float>(omc_ctx *, hipStream_t, PhantomDev *, const omc_dosxyz_source_desc *, int64_t,
                                         int, const omc_run_opts *, double *, double *, omc_run_stats *);
template int omc_run_dosxyz_event<double>(omc_ctx *, hipStream_t, PhantomDev *, const omc_dosxyz_source_desc *,
                                          int64_t, int, const omc_run_opts *, double *, double *, omc_run_stats *);

// ------------------------------------------------------------------ Dij runs
namespace {

// Histories per batch of each beamlet as the reference derives them
// (omc_matrad.c:1490-1498): nperbatch = nhist / nbatch, at least 1.
std::vector<long> beamlet_nper(const omc_beamlets_desc *bix, int nbatch) {
    std::vector<long> v(bix->nbixels);
    for (int b = 0; b < bix->nbixels; b++) {
        long nh = bix->nhist ? (long)bix->nhist[b] : (long)bix->nhist_each;
        if (nh / nbatch == 0) nh = nbatch;
        v[b] = nh / nbatch;
    }
    return v;
}

// Device bytes of one beamlet of a group whose jobs carry nbj batches: nbj
// batch tallies with their tile marks, plus two arrays for the running sum
// and sum of squares (or the dose and variance), the beamlet's tile map and
// its compaction counts.
size_t dij_slot_bytes(long nreg, int nbj) {
    return (size_t)(nbj + 2) * nreg * sizeof(double) + (size_t)(nbj + 1) * dij_ntile(nreg) +
           (size_t)dij_nchunk(nreg) * (sizeof(int) + sizeof(int64_t)) + (size_t)dij_nblk(nreg) * sizeof(double) +
           (size_t)64 * nbj;
}

// Device bytes of a Dij run besides its beamlet slots: carry and plan
// arrays, and the event pools or, for the history form with nsplit above
// HISTORY_SCRATCH_NSPLIT, the particle stacks in device memory.
size_t dij_fixed_bytes(long nreg, const omc_run_opts *o, long long target) {
    size_t b = (size_t)3 * nreg * sizeof(double) + ((size_t)64 << 20);
    if (o->form == OMC_FORM_EVENT)
        b += o->precision == OMC_FP32 ? event_bytes<float>(target) : event_bytes<double>(target);
    else if (o->nsplit > HISTORY_SCRATCH_NSPLIT)    // particle stacks, capped as alloc_split_stacks caps them
        b += std::min<size_t>((size_t)(o->grid > 0 ? o->grid : 2048) * (o->block > 0 ? o->block : 256) *
                                  split_stack_depth(o->nsplit) *
                                  (o->precision == OMC_FP32 ? sizeof(Part<float>) : sizeof(Part<double>)),
                              (size_t)8 << 30);
    return b;
}

// Shape of a Dij run: G beamlets per group and nbj batches per transport job
// (event form; the history form runs one batch per launch).  Every job ends
// with the slowest track of its particles, so the plan takes the fewest jobs,
// ceil(nbix / G) x ceil(nbatch / nbj), whose tallies fit the budget; the
// options group and batches_per_job pin either factor.  Returns the bytes the
// plan uses.
struct DijPlan {
    int G = 1, nbj = 1;
};

// Default memory budget of a Dij run's tallies: half the free device memory,
// at most 96 GiB.  Past ~100 GiB held by the process, allocations cost
// seconds to map and unmap (measured on the 1960-beamlet plan: 128-144 GiB
// runs lost 2-5 s per call, 96 GiB none) while saving only a few jobs.
size_t dij_default_budget() { return std::min(default_budget(), (size_t)96 << 30); }

size_t dij_plan(const omc_run_opts *o, long nreg, int nbix, int nbatch, long long target, size_t budget, DijPlan &P) {
    if (!budget) budget = o->hbm_budget > 0 ? (size_t)o->hbm_budget : dij_default_budget();
    const size_t fixed = dij_fixed_bytes(nreg, o, target);
    const size_t room = budget > fixed ? budget - fixed : 0;
    const int gmax = std::min(nbix, 4096);
    long best_jobs = -1;
    int lo = 1, hi = nbatch;
    if (o->batches_per_job > 0) lo = hi = std::min(o->batches_per_job, nbatch);
    for (int nbj = hi; nbj >= lo; nbj--) {
        long g = o->group > 0 ? std::min(o->group, nbix) : (long)(room / dij_slot_bytes(nreg, nbj));
        g = std::max(1L, std::min<long>(g, gmax));
        if (nbj > lo && (size_t)g * dij_slot_bytes(nreg, nbj) > room) continue;   // fewer batches per job
        long jobs = (long)((nbix + g - 1) / g) * (long)((nbatch + nbj - 1) / nbj);
        if (best_jobs < 0 || jobs < best_jobs) {
            best_jobs = jobs;
            P.G = (int)g;
            P.nbj = nbj;
        }
    }
    return fixed + (size_t)P.G * dij_slot_bytes(nreg, P.nbj);
}

template <class R>
int upload_beamlets(DevArena &A, const omc_spectrum_desc *spec, const omc_beamlets_desc *b, SrcBeamlets<R> &S) {
    fill_spectrum<R>(A, *spec, S.spectrum, S.energy, S.deltak, S.cdfinv1, S.cdfinv2);
    std::vector<int> ib(b->ibeam, b->ibeam + b->nbixels);
    for (int v : ib)
        if (v < 0 || v >= b->nbeams) return omc_seterr(OMC_ERR_ARG, "beamlet beam index %d of %d beams", v, b->nbeams);
    auto vec = [&](const double *p, size_t n) {
        std::vector<R> v(n);
        for (size_t k = 0; k < n; k++) v[k] = (R)p[k];
        return A.upload(v);
    };
    S.ibeam = A.upload(ib);
    S.src = vec(b->src, 3 * (size_t)b->nbeams);
    S.corner = vec(b->corner, 3 * (size_t)b->nbixels);
    S.side1 = vec(b->side1, 3 * (size_t)b->nbixels);
    S.side2 = vec(b->side2, 3 * (size_t)b->nbixels);
    S.gaussian = b->gaussian;
    S.sigma = (R)b->sigma;
    if (!A.ok) return omc_seterr(OMC_ERR_NOMEM, "beamlet arrays");
    return OMC_OK;
}

// One lane of a Dij run: a stream with its own event pools, batch tallies,
// finish and compaction buffers.  With two lanes, alternate beamlet groups
// are transported concurrently, so the drain at the end of one group's job
// (a few long electron tracks that leave most of the device idle) overlaps
// the bulk of the next group's job on the other stream.  The finish of a
// group reads and writes the carry of the previous beamlet's variance (a
// behaviour of the reference kept for parity) and the plan totals, so group
// finishes run in group order across the lanes, chained by events.
template <class R> struct DijLane {
    hipStream_t st = nullptr;
    bool own_stream = false;
    DevArena A;
    EvState<R> S;
    EvLaunch<R> L;
    HistArgs<R> H;
    double *d_e = nullptr, *d_s = nullptr, *d_s2 = nullptr, *d_bmax = nullptr, *d_thresh = nullptr, *d_nper = nullptr;
    double *d_ebal = nullptr;          // [2 * EB_N]: the lane's energy balance and its copy before a job
    uint8_t *d_touch = nullptr;        // [nslot][ntile] tile marks of the batch slots (fused finish)
    uint8_t *d_any = nullptr;          // [G][ntile] tiles a beamlet received energy in
    int *d_counts = nullptr, *d_slot_bix = nullptr, *d_slot_bt = nullptr;
    int64_t *d_off = nullptr;
    long *d_slot_off = nullptr;
    hipEvent_t fin = nullptr;          // finish of the lane's latest group
    std::vector<DevArena> gouts;       // csc_on_device: the compacted columns of this lane's groups
    std::vector<uint8_t> hmark;        // given tallies: tile marks computed on the host
    JobStats js;
    double t_hist = 0, t_other = 0, histories = 0;
    uint64_t n_hist = 0, n_other = 0;
    int rc = OMC_OK;
    int max_grid = 2048;               // history form: largest launch the particle stacks cover
    std::string msg;
    ~DijLane() {
        A.release();      // cached blocks: each release synchronises st, so all go before st does
        S.A.release();
        for (auto &g : gouts) g.release();
        if (fin) (void)hipEventDestroy(fin);
        if (own_stream && st) (void)hipStreamDestroy(st);
    }
};

struct DijGroupOut {
    std::vector<int64_t> collen;       // entries of each column of the group
    int64_t nnz = 0;
    int64_t *d_ir = nullptr;           // the group's columns in device memory
    double *d_val = nullptr, *d_var = nullptr;
};

// Batch tallies handed in by the caller (omc_dij_assemble) instead of
// transported: endep[(b * nbatch + k) * nreg + irl] for beamlet b, batch k.
struct DijGiven {
    const double *endep;
};

template <class R> struct DijRun {
    omc_ctx *c;
    PhantomDev *pd;
    const omc_run_opts *o;
    const DijGiven *given;
    int nbix, nbatch, G, NBJ, nblk, nchunk, ngroups;
    long nreg, nvox, ntile;
    bool event, count, timing, replay, want_var, on_device, fused;
    double rel;
    long long tgt;
    std::vector<long> nper, gfirst;
    SrcBeamlets<R> bsrc;
    double *d_carry, *d_tdose, *d_tvar;
    unsigned long long *d_cnt, *d_next;
    int *d_err;
    uint32_t *d_rn;
    long *d_rstart, *d_rend;
    std::vector<DijLane<R> *> lanes;
    std::vector<DijGroupOut> outs;
    std::mutex mu;
    std::condition_variable cv;
    int next_finish = 0;
    bool failed = false;
};

constexpr int RC_LANE_STOPPED = -101;     // a lane gave up because the other lane failed

template <class R> int dij_lane_init(DijRun<R> &D, DijLane<R> &l) {
    const int G = D.G;
    const size_t nslot = (size_t)G * D.NBJ;
    DevArena &A = l.A;
    l.d_e = A.alloc_n<double>(nslot * D.nreg);
    l.d_s = A.alloc_n<double>((size_t)G * D.nreg);
    l.d_s2 = A.alloc_n<double>((size_t)G * D.nreg);
    l.d_bmax = A.alloc_n<double>((size_t)G * D.nblk);
    l.d_thresh = A.alloc_n<double>(G);
    l.d_nper = A.alloc_n<double>(G);
    l.d_ebal = A.alloc_n<double>(2 * EB_N);
    l.d_any = A.alloc_n<uint8_t>((size_t)G * D.ntile);
    if (D.fused) l.d_touch = A.alloc_n<uint8_t>(nslot * D.ntile);
    l.d_counts = A.alloc_n<int>((size_t)G * D.nchunk);
    l.d_off = A.alloc_n<int64_t>((size_t)G * D.nchunk);
    l.d_slot_off = A.alloc_n<long>(nslot + 1);
    l.d_slot_bix = A.alloc_n<int>(nslot);
    l.d_slot_bt = A.alloc_n<int>(nslot);
    if (!A.ok)
        return omc_seterr(OMC_ERR_NOMEM, "Dij tallies for a group of %d beamlets x %d batches do not fit", G, D.NBJ);
    CK(hipEventCreateWithFlags(&l.fin, hipEventDisableTiming));
    CK(hipMemsetAsync(l.d_e, 0, sizeof(double) * nslot * D.nreg, l.st));
    CK(hipMemsetAsync(l.d_ebal, 0, sizeof(double) * 2 * EB_N, l.st));
    if (D.fused) CK(hipMemsetAsync(l.d_touch, 0, nslot * D.ntile, l.st));
    else CK(hipMemsetAsync(l.d_any, 1, (size_t)G * D.ntile, l.st));     // every tile counts as touched
    const omc_run_opts *o = D.o;
    if (D.given) return OMC_OK;
    if (D.event) {
        long long maxjob = 0;
        for (int g0 = 0; g0 < D.nbix; g0 += G) {
            long long t = 0;
            for (int b = g0; b < std::min(D.nbix, g0 + G); b++) t += D.nper[b];
            maxjob = std::max(maxjob, t * D.NBJ);
        }
        long long tgt = D.tgt;
        if (tgt > 4 * maxjob) tgt = std::max<long long>(4 * maxjob, 64);
        int rc = l.S.alloc(tgt);
        if (rc) return rc;
        fill_launch(D.c, D.pd, l.S, o, l.L);
        l.L.src.bix = D.bsrc;
        l.L.src.use_beamlets = 1;
        l.L.src.slot_off = l.d_slot_off;
        l.L.src.slot_beamlet = l.d_slot_bix;
        l.L.src.slot_batch = l.d_slot_bt;
        l.L.endep = l.d_e;
        l.L.slot_stride = D.nreg;
        l.L.touched = l.d_touch;
        l.L.touch_stride = D.ntile;
        l.L.ebal = l.d_ebal;
        l.L.counters = D.count ? D.d_cnt : nullptr;
    } else {
        HistArgs<R> &H = l.H;
        memset(&H, 0, sizeof H);
        H.t = tables_of<R>(D.c);
        H.g = geom_of<R>(D.pd);
        H.bix = D.bsrc;
        H.use_beamlets = 1;
        H.slot_off = l.d_slot_off;
        H.slot_beamlet = l.d_slot_bix;
        H.slot_stride = D.nreg;
        H.touch_stride = D.ntile;
        H.ebal = l.d_ebal;
        H.seed = o->seed;
        H.endep = l.d_e;
        H.counters = D.count ? D.d_cnt : nullptr;
        H.err = D.d_err;
        H.next = D.d_next;
        H.chunk = 1;
        H.max_steps = hist_max_steps(o);
        H.replay = D.d_rn;
        H.replay_off = D.d_rstart;
        H.replay_end = D.d_rend;
        H.nsplit = o->nsplit;
        // the largest launch is one batch of the largest group: the particle
        // stacks of nsplit > 8 cover that many blocks and no more
        const int hblock = o->block > 0 ? o->block : 256;
        long long maxlaunch = 1;
        for (int g0 = 0; g0 < D.nbix; g0 += G) {
            long long t = 0;
            for (int b = g0; b < std::min(D.nbix, g0 + G); b++) t += D.nper[b];
            maxlaunch = std::max(maxlaunch, t);
        }
        l.max_grid = o->grid > 0 ? o->grid : (int)std::min<long long>((maxlaunch + hblock - 1) / hblock, 2048);
        int rc = alloc_split_stacks<R>(A, H, o->nsplit, hblock, l.max_grid);
        if (rc) return rc;
    }
    return OMC_OK;
}

// Given tallies (omc_dij_assemble): batches bt .. bt+nb-1 of the group's
// beamlets into their slots, with the tile marks the transport would set.
template <class R> int dij_given_job(DijRun<R> &D, DijLane<R> &l, int g0, int ng, int bt, int nb) {
    const hipStream_t stream = l.st;
    const long nreg = D.nreg, ntile = D.ntile;
    if (D.fused) l.hmark.assign((size_t)nb * ng * ntile, 0);
    for (int k = 0; k < nb; k++)
        for (int sl = 0; sl < ng; sl++) {
            const double *src = D.given->endep + ((size_t)(g0 + sl) * D.nbatch + bt + k) * nreg;
            const size_t slot = (size_t)k * ng + sl;
            CK(hipMemcpyAsync(l.d_e + slot * nreg, src, sizeof(double) * nreg, hipMemcpyHostToDevice, stream));
            if (D.fused)
                for (long r = 0; r < nreg; r++)
                    if (src[r] != 0.0) l.hmark[slot * ntile + (r >> TALLY_TILE_LOG2)] = 1;
        }
    if (D.fused) CK(hipMemcpyAsync(l.d_touch, l.hmark.data(), l.hmark.size(), hipMemcpyHostToDevice, stream));
    CK(hipStreamSynchronize(stream));    // the host marks may go
    return OMC_OK;
}

// Transport (or given tallies), batch statistics, finish (in group order)
// and compaction of group gi on lane l; the group's columns land in
// D.outs[gi].
template <class R> int dij_lane_group(DijRun<R> &D, DijLane<R> &l, int gi) {
    const omc_run_opts *o = D.o;
    const hipStream_t stream = l.st;
    const long nreg = D.nreg;
    const int nchunk = D.nchunk, g0 = gi * D.G, ng = std::min(D.G, D.nbix - g0);
    std::vector<long> soff(ng + 1, 0);
    std::vector<int> sbix(ng);
    std::vector<double> snper(ng);
    for (int s = 0; s < ng; s++) {
        sbix[s] = g0 + s;
        snper[s] = (double)D.nper[g0 + s];
        soff[s + 1] = soff[s] + D.nper[g0 + s];
    }
    const long njob = soff[ng];
    if (!D.event && !D.given) {
        CK(hipMemcpyAsync(l.d_slot_off, soff.data(), sizeof(long) * (ng + 1), hipMemcpyHostToDevice, stream));
        CK(hipMemcpyAsync(l.d_slot_bix, sbix.data(), sizeof(int) * ng, hipMemcpyHostToDevice, stream));
    }
    CK(hipMemcpyAsync(l.d_nper, snper.data(), sizeof(double) * ng, hipMemcpyHostToDevice, stream));
    if (!D.fused) {
        CK(hipMemsetAsync(l.d_s, 0, sizeof(double) * ng * nreg, stream));
        CK(hipMemsetAsync(l.d_s2, 0, sizeof(double) * ng * nreg, stream));
    }
    Events ev(2);
    for (int bt = 0; bt < D.nbatch; bt += D.NBJ) {
        const int nb = std::min(D.NBJ, D.nbatch - bt);
        if (D.given) {
            int rc = dij_given_job(D, l, g0, ng, bt, nb);
            if (rc) return rc;
        } else if (D.event) {
            // one job for batches bt .. bt+nb-1 of the group: slot k * ng + s
            // is beamlet g0 + s in batch bt + k, with that batch's stream keys
            std::vector<long> voff((size_t)nb * ng + 1, 0);
            std::vector<int> vbix((size_t)nb * ng), vbt((size_t)nb * ng);
            for (int k = 0; k < nb; k++)
                for (int sl = 0; sl < ng; sl++) {
                    const size_t v = (size_t)k * ng + sl;
                    vbix[v] = g0 + sl;
                    vbt[v] = bt + k;
                    voff[v + 1] = voff[v] + D.nper[g0 + sl];
                }
            CK(hipMemcpyAsync(l.d_slot_off, voff.data(), sizeof(long) * voff.size(), hipMemcpyHostToDevice, stream));
            CK(hipMemcpyAsync(l.d_slot_bix, vbix.data(), sizeof(int) * vbix.size(), hipMemcpyHostToDevice, stream));
            CK(hipMemcpyAsync(l.d_slot_bt, vbt.data(), sizeof(int) * vbt.size(), hipMemcpyHostToDevice, stream));
            l.L.src.nslot = nb * ng;
            l.L.src.batch = bt;
            int rc = run_job_grow(stream, l.S, l.L, voff.back(), 0, D.count, D.timing, l.js, l.d_e,
                                  (size_t)nb * ng * nreg, l.d_ebal);
            if (rc) return rc;     // run_job synchronised the stream: the host vectors may go
        } else {
            // one launch per batch; batch k of the job tallies into slots k * ng + s
            HistArgs<R> &H = l.H;
            for (int k = 0; k < nb; k++) {
                const int b = bt + k;
                H.nslot = ng;
                H.batch = b;
                H.nhist = njob;
                H.hist_base = 0;
                H.endep = l.d_e + (size_t)k * ng * nreg;
                H.touched = D.fused ? l.d_touch + (size_t)k * ng * D.ntile : nullptr;
                if (D.replay) {
                    std::vector<long> rs(njob), re(njob);
                    for (int sl = 0; sl < ng; sl++)
                        for (long i = 0; i < D.nper[g0 + sl]; i++) {
                            long gh = D.gfirst[g0 + sl] + (long)b * D.nper[g0 + sl] + i;
                            rs[soff[sl] + i] = (long)o->replay_offsets[gh];
                            re[soff[sl] + i] = (long)o->replay_offsets[gh + 1];
                        }
                    CK(hipMemcpyAsync(D.d_rstart, rs.data(), sizeof(long) * njob, hipMemcpyHostToDevice, stream));
                    CK(hipMemcpyAsync(D.d_rend, re.data(), sizeof(long) * njob, hipMemcpyHostToDevice, stream));
                    CK(hipStreamSynchronize(stream));   // rs/re go out of scope
                }
                CK(hipMemsetAsync(D.d_next, 0, sizeof(unsigned long long), stream));
                int block = o->block > 0 ? o->block : 256;
                long want = (njob + block - 1) / block;
                int grid = o->grid > 0 ? o->grid : (int)std::min<long>(want, 2048);
                if (grid > l.max_grid) grid = l.max_grid;
                if (D.timing) CK(hipEventRecord(ev.ev[0], stream));
                CK(launch_history<R>(H, D.replay ? RNG_REPLAY : RNG_STREAMS, D.count, grid, block, stream));
                if (D.timing) {
                    CK(hipEventRecord(ev.ev[1], stream));
                    CK(hipEventSynchronize(ev.ev[1]));
                    l.t_hist += ev.ms(0, 1) * 1e-3;
                }
                l.n_hist++;
            }
        }
        l.histories += (double)njob * nb;
        // batch statistics in batch order, as batch-by-batch jobs take them
        if (!D.fused)
            for (int k = 0; k < nb; k++)
                CK(launch_accum_slots(l.d_e + (size_t)k * ng * nreg, l.d_s, l.d_s2, nreg, nreg, ng, l.d_nper,
                                      l.d_ebal + EB_DEPOSIT, stream));
    }
    if (D.timing) CK(hipEventRecord(ev.ev[0], stream));
    {
        std::unique_lock<std::mutex> lk(D.mu);
        D.cv.wait(lk, [&] { return D.failed || D.next_finish == gi; });
        if (D.failed) return RC_LANE_STOPPED;
    }
    if (gi > 0) CK(hipStreamWaitEvent(stream, D.lanes[(gi - 1) % D.lanes.size()]->fin, 0));
    if (D.fused)
        CK(launch_dij_fold_finish(l.d_e, l.d_touch, ng, D.nbatch, nreg, l.d_nper, l.d_s, l.d_s2, l.d_any, D.d_carry,
                                  D.pd->dmass, D.pd->drho, l.d_bmax, D.d_tdose, D.d_tvar, l.d_ebal + EB_DEPOSIT,
                                  stream));
    else
        CK(launch_dij_finish(l.d_s, l.d_s2, D.d_carry, ng, nreg, D.pd->dmass, D.pd->drho, D.nbatch, l.d_bmax,
                             D.d_tdose, D.d_tvar, stream));
    CK(hipEventRecord(l.fin, stream));
    {
        std::lock_guard<std::mutex> lk(D.mu);
        D.next_finish = gi + 1;
        D.cv.notify_all();
    }
    CK(launch_dij_select(l.d_s, l.d_any, nreg, ng, l.d_bmax, D.rel, l.d_thresh, l.d_counts, stream));
    std::vector<int> counts((size_t)ng * nchunk);
    std::vector<int64_t> off((size_t)ng * nchunk);
    CK(hipMemcpyAsync(counts.data(), l.d_counts, sizeof(int) * ng * nchunk, hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    int64_t tot = 0;
    for (size_t k = 0; k < counts.size(); k++) {
        off[k] = tot;
        tot += counts[k];
    }
    DijGroupOut &out = D.outs[gi];
    out.collen.resize(ng);
    for (int s = 0; s < ng; s++) {
        int64_t end = off[(size_t)(s + 1) * nchunk - 1] + counts[(size_t)(s + 1) * nchunk - 1];
        int64_t begin = s ? off[(size_t)s * nchunk - 1] + counts[(size_t)s * nchunk - 1] : 0;
        out.collen[s] = end - begin;
    }
    out.nnz = tot;
    {
        // compact straight into device buffers kept until the columns are
        // stitched (on the device, then downloaded once for host output)
        l.gouts.emplace_back();
        DevArena &GA = l.gouts.back();
        GA.use_cache(stream);
        out.d_ir = GA.alloc_n<int64_t>((size_t)std::max<int64_t>(tot, 1));
        out.d_val = GA.alloc_n<double>((size_t)std::max<int64_t>(tot, 1));
        out.d_var = D.want_var ? GA.alloc_n<double>((size_t)std::max<int64_t>(tot, 1)) : nullptr;
        if (!GA.ok) return omc_seterr(OMC_ERR_NOMEM, "Dij device output (%lld entries)", (long long)tot);
        if (tot > 0) {
            CK(hipMemcpyAsync(l.d_off, off.data(), sizeof(int64_t) * ng * nchunk, hipMemcpyHostToDevice, stream));
            CK(launch_dij_compact(l.d_s, l.d_s2, l.d_any, nreg, ng, l.d_thresh, l.d_off, out.d_ir, out.d_val,
                                  out.d_var, stream));
        }
    }
    if (D.timing) {
        CK(hipEventRecord(ev.ev[1], stream));
        CK(hipEventSynchronize(ev.ev[1]));
        l.t_other += ev.ms(0, 1) * 1e-3;
        l.n_other += 4;
    }
    CK(hipStreamSynchronize(stream));   // the host vectors counts/off go out of scope
    return OMC_OK;
}

// Groups gi = first, first + step, ... on lane l; on failure the error text is
// kept with the lane (error messages are per thread) and the other lane is
// released from its wait for the finish order.
template <class R> void dij_lane_run(DijRun<R> &D, DijLane<R> &l, int first, int step) {
    (void)hipSetDevice(D.c->device);      // the current device is per thread
    for (int gi = first; gi < D.ngroups; gi += step) {
        int rc = dij_lane_group(D, l, gi);
        if (rc) {
            if (rc != RC_LANE_STOPPED) {
                l.rc = rc;
                l.msg = omc_host_last_error();
            }
            std::lock_guard<std::mutex> lk(D.mu);
            D.failed = true;
            D.cv.notify_all();
            return;
        }
    }
}

template <class R>
int run_dij(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_spectrum_desc *spec, const omc_beamlets_desc *bix,
            int nbatch, double rel, const omc_run_opts *o, omc_csc *dij, omc_csc *var, double *plan_dose,
            double *plan_var, omc_run_stats *st, const DijGiven *given = nullptr) {
    auto wall0 = std::chrono::steady_clock::now();
    DijRun<R> D;
    D.c = c;
    D.pd = pd;
    D.o = o;
    D.given = given;
    D.nbix = bix->nbixels;
    D.nbatch = nbatch;
    D.nreg = pd->nreg;
    D.nvox = D.nreg - 1;
    D.event = o->form == OMC_FORM_EVENT;
    D.count = (o->count_events & 1) != 0;
    D.timing = o->count_events != 0;
    D.replay = o->rng == OMC_RNG_REPLAY;
    D.want_var = var != nullptr;
    D.on_device = o->csc_on_device != 0;
    D.rel = rel;
    const int nbix = D.nbix;
    const long nreg = D.nreg, nvox = D.nvox;
    D.nper = beamlet_nper(bix, nbatch);
    for (long n : D.nper)
        if (n >= (1L << 32)) return omc_seterr(OMC_ERR_LIMIT, "more than 2^32 histories per batch of a beamlet");
    D.tgt = o->event_target > 0 ? o->event_target : default_target(c->device, o->nsplit);
    // lanes: up to three for an event-form run of several groups (o->streams
    // pins the number), each planned on its share of the memory budget.  On
    // the 1960-beamlet plan at 96 GiB: 1 lane 7.2, 2 lanes 10.1, 3 lanes
    // 10.4, 4 lanes 7-9 M histories/s.
    size_t budget = o->hbm_budget > 0 ? (size_t)o->hbm_budget : dij_default_budget();
    DijPlan plan;
    dij_plan(o, nreg, nbix, nbatch, D.tgt, budget, plan);
    int NL = 1;
    if (D.event) NL = o->streams > 0 ? o->streams : std::min(3, (nbix + plan.G - 1) / plan.G);
    if (NL > 1) dij_plan(o, nreg, nbix, nbatch, D.tgt, budget / NL, plan);
    D.G = plan.G;
    D.NBJ = plan.nbj;
    // every batch of a group in one job: the batch statistics fold into the
    // finish, over the tiles the transport touched
    D.fused = D.NBJ == nbatch;
    D.ngroups = (nbix + D.G - 1) / D.G;
    NL = std::min(NL, D.ngroups);
    D.ntile = dij_ntile(nreg);
    D.nblk = dij_nblk(nreg);
    D.nchunk = dij_nchunk(nreg);
    DevArena A;
    ArenaGuard ag{A};
    A.use_cache(stream);
    int rc = given ? OMC_OK : upload_beamlets<R>(A, spec, bix, D.bsrc);
    if (rc) return rc;
    D.d_carry = A.alloc_n<double>(nreg);
    D.d_tdose = A.alloc_n<double>(nreg);
    D.d_tvar = A.alloc_n<double>(nreg);
    D.d_cnt = A.alloc_n<unsigned long long>(16);
    D.d_next = A.alloc_n<unsigned long long>(1);
    D.d_err = A.alloc_n<int>(4);
    D.d_rn = nullptr;
    D.d_rstart = D.d_rend = nullptr;
    if (!A.ok) return omc_seterr(OMC_ERR_NOMEM, "Dij run arrays");
    // Replayed draws (verification): histories are numbered beamlet by beamlet,
    // batch by batch, as the reference's serial loop consumes its stream.
    D.gfirst.assign(nbix + 1, 0);
    for (int b = 0; b < nbix; b++) D.gfirst[b + 1] = D.gfirst[b] + (long)nbatch * D.nper[b];
    if (D.replay) {
        if (o->replay_len < o->replay_offsets[D.gfirst[nbix]])
            return omc_seterr(OMC_ERR_ARG, "replay stream shorter than its offsets");
        D.d_rn = A.alloc_n<uint32_t>(o->replay_len > 0 ? o->replay_len : 1);
        long maxjob = 0;
        for (int g0 = 0; g0 < nbix; g0 += D.G) {
            long t = 0;
            for (int b = g0; b < std::min(nbix, g0 + D.G); b++) t += D.nper[b];
            maxjob = std::max(maxjob, t);
        }
        D.d_rstart = A.alloc_n<long>(maxjob);
        D.d_rend = A.alloc_n<long>(maxjob);
        if (!A.ok) return omc_seterr(OMC_ERR_NOMEM, "replay stream");
        if (o->replay_len > 0)
            CK(hipMemcpy(D.d_rn, o->replay, sizeof(uint32_t) * o->replay_len, hipMemcpyHostToDevice));
    }
    CK(hipMemsetAsync(D.d_carry, 0, sizeof(double) * nreg, stream));
    CK(hipMemsetAsync(D.d_tdose, 0, sizeof(double) * nreg, stream));
    CK(hipMemsetAsync(D.d_tvar, 0, sizeof(double) * nreg, stream));
    CK(hipMemsetAsync(D.d_cnt, 0, sizeof(unsigned long long) * 16, stream));
    CK(hipMemsetAsync(D.d_err, 0, sizeof(int) * 4, stream));
    std::vector<std::unique_ptr<DijLane<R>>> lanes;
    for (int k = 0; k < NL; k++) {
        lanes.emplace_back(new DijLane<R>());
        DijLane<R> &l = *lanes.back();
        if (k == 0) {
            l.st = stream;
        } else {
            CK(hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking));
            l.own_stream = true;
        }
        l.A.use_cache(l.st);
        l.S.A.use_cache(l.st);
        D.lanes.push_back(&l);
    }
    Events ev(2);
    CK(hipEventRecord(ev.ev[0], stream));
    for (int k = 1; k < NL; k++) CK(hipStreamWaitEvent(lanes[k]->st, ev.ev[0], 0));   // after the set-up above
    for (auto &l : lanes) l->js.origin = ev.ev[0];       // every lane's work follows it
    for (auto &l : lanes)
        if ((rc = dij_lane_init(D, *l))) return rc;
    D.outs.resize(D.ngroups);
    {
        std::vector<std::thread> th;
        for (int k = 1; k < NL; k++) th.emplace_back(dij_lane_run<R>, std::ref(D), std::ref(*lanes[k]), k, NL);
        dij_lane_run<R>(D, *lanes[0], 0, NL);
        for (auto &t : th) t.join();
    }
    for (auto &l : lanes)
        if (l->rc) return omc_seterr(l->rc, "%s", l->msg.c_str());
    for (auto &l : lanes) CK(hipStreamSynchronize(l->st));
    CK(hipEventRecord(ev.ev[1], stream));
    CK(hipStreamSynchronize(stream));
    int err[4];
    CK(hipMemcpy(err, D.d_err, sizeof err, hipMemcpyDeviceToHost));
    if (err[0]) return omc_seterr(OMC_ERR_STACK, "particle stack overflow on the device");
    if (err[1]) return omc_seterr(OMC_ERR_ARG, "replay stream underrun");
    if (err[2]) return omc_seterr(OMC_ERR_LIMIT, "history step guard reached");
    if (plan_dose) CK(hipMemcpy(plan_dose, D.d_tdose + 1, sizeof(double) * nvox, hipMemcpyDeviceToHost));
    if (plan_var) CK(hipMemcpy(plan_var, D.d_tvar + 1, sizeof(double) * nvox, hipMemcpyDeviceToHost));
    // columns in beamlet order, group after group
    int64_t nnz = 0;
    for (auto &g : D.outs) nnz += g.nnz;
    auto fill_csc = [&](omc_csc *m, bool variance) -> int {
        memset(m, 0, sizeof *m);
        m->nrows = nvox;
        m->ncols = nbix;
        m->nnz = nnz;
        m->device = D.on_device ? 1 : 0;
        m->jc = (int64_t *)malloc(sizeof(int64_t) * (nbix + 1));
        // columns stitched on the device; host output downloads each array
        // once from stitch buffers of the block cache (hipFree would
        // synchronise the whole device), device output keeps its own
        DevArena T;
        ArenaGuard tg{T};
        T.use_cache(stream);
        int64_t *d_ir = nullptr;
        double *d_val = nullptr;
        if (D.on_device) {
            if (hipMalloc(&d_ir, sizeof(int64_t) * (nnz ? nnz : 1)) != hipSuccess) d_ir = nullptr;
            if (hipMalloc(&d_val, sizeof(double) * (nnz ? nnz : 1)) != hipSuccess) d_val = nullptr;
            (void)hipGetLastError();
        } else {
            d_ir = T.alloc_n<int64_t>((size_t)(nnz ? nnz : 1));
            d_val = T.alloc_n<double>((size_t)(nnz ? nnz : 1));
        }
        m->ir = D.on_device ? d_ir : nullptr;
        m->val = D.on_device ? d_val : nullptr;
        m->device = D.on_device ? 1 : 0;
        if (!m->jc || !d_ir || !d_val) {
            omc_csc_free(m);
            return omc_seterr(OMC_ERR_NOMEM, "device CSC of %lld entries", (long long)nnz);
        }
        int64_t e = 0;
        int col = 0;
        m->jc[0] = 0;
        for (auto &g : D.outs) {
            for (int64_t len : g.collen) {
                m->jc[col + 1] = m->jc[col] + len;
                col++;
            }
            if (g.nnz == 0) continue;
            CK(hipMemcpyAsync(d_ir + e, g.d_ir, sizeof(int64_t) * g.nnz, hipMemcpyDeviceToDevice, stream));
            CK(hipMemcpyAsync(d_val + e, variance ? g.d_var : g.d_val, sizeof(double) * g.nnz,
                              hipMemcpyDeviceToDevice, stream));
            e += g.nnz;
        }
        CK(hipStreamSynchronize(stream));
        if (D.on_device) return OMC_OK;
        // host arrays: pinned blocks from the reuse pool (a DMA download into
        // mapped pages), else ordinary memory
        m->ir = (int64_t *)pinned_block_alloc(sizeof(int64_t) * (nnz ? nnz : 1));
        m->val = (double *)pinned_block_alloc(sizeof(double) * (nnz ? nnz : 1));
        m->device = 2;
        if (!m->ir || !m->val) {
            pinned_block_free(m->ir);
            pinned_block_free(m->val);
            m->ir = (int64_t *)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
            m->val = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
            m->device = 0;
        }
        int rc2 = OMC_OK;
        if (!m->ir || !m->val) {
            rc2 = omc_seterr(OMC_ERR_NOMEM, "host CSC of %lld entries", (long long)nnz);
        } else if (nnz > 0 && (hipMemcpy(m->ir, d_ir, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost) != hipSuccess ||
                               hipMemcpy(m->val, d_val, sizeof(double) * nnz, hipMemcpyDeviceToHost) != hipSuccess)) {
            rc2 = omc_seterr(OMC_ERR_DEVICE, "download of the CSC: %s", hipGetErrorString(hipGetLastError()));
        }
        if (rc2) omc_csc_free(m);
        return rc2;
    };
    if (dij && (rc = fill_csc(dij, false))) return rc;
    if (var && (rc = fill_csc(var, true))) {
        if (dij) omc_csc_free(dij);
        return rc;
    }
    if (st) {
        unsigned long long cnt[16];
        CK(hipMemcpy(cnt, D.d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
        st->gpu_seconds = ev.ms(0, 1) * 1e-3;
        stats_counters(st, cnt);
        double histories = 0;
        for (auto &l : lanes) {
            double eb[EB_N];
            CK(hipMemcpy(eb, l->d_ebal, sizeof eb, hipMemcpyDeviceToHost));
            stats_energy(st, eb);
            histories += l->histories;
            if (D.event) stats_from_job(st, l->js);
            st->kernel_seconds[OMC_K_HISTORY] += l->t_hist;
            st->kernel_launches[OMC_K_HISTORY] += l->n_hist;
            st->kernel_seconds[OMC_K_OTHER] += l->t_other;
            st->kernel_launches[OMC_K_OTHER] += l->n_other;
        }
        if (D.event && D.timing) {
            std::vector<const JobStats *> jobs;
            for (auto &l : lanes) jobs.push_back(&l->js);
            stats_busy(st, jobs);
        }
        st->histories = histories;
        st->groups = D.ngroups;
        st->group_size = D.G;
        st->batches_per_job = D.NBJ;
        st->streams = NL;
        st->nnz = nnz;
        st->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    }
    return OMC_OK;
}

}  // namespace

int omc_dij_run(omc_ctx *c, hipStream_t stream, PhantomDev *pd, const omc_spectrum_desc *spec,
                const omc_beamlets_desc *bix, int nbatch, double rel, const omc_run_opts *o, omc_csc *dij,
                omc_csc *var, double *plan_dose, double *plan_var, omc_run_stats *st) {
    if (!c || !pd || !spec || !bix || bix->nbixels < 1 || bix->nbeams < 1 || !bix->ibeam || !bix->src ||
        !bix->corner || !bix->side1 || !bix->side2 || nbatch < 1)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dij: bad arguments");
    if (nbatch > 255) return omc_seterr(OMC_ERR_LIMIT, "omc_run_dij: at most 255 batches");
    if (!bix->nhist && bix->nhist_each < 1) return omc_seterr(OMC_ERR_ARG, "omc_run_dij: no histories per beamlet");
    if (spec->spectrum && (!spec->cdfinv1 || !spec->cdfinv2))
        return omc_seterr(OMC_ERR_ARG, "spectrum source without inverse CDF");
    int rc = omc_check_opts(o);
    if (rc) return rc;
    if (o->rng == OMC_RNG_REPLAY && o->form != OMC_FORM_HISTORY)
        return omc_seterr(OMC_ERR_ARG, "omc_run_dij: replayed random numbers need the history form");
    (void)hipSetDevice(c->device);
    if (st) memset(st, 0, sizeof *st);
    if (dij) memset(dij, 0, sizeof *dij);
    if (var) memset(var, 0, sizeof *var);
    return o->precision == OMC_FP32
               ? run_dij<float>(c, stream, pd, spec, bix, nbatch, rel, o, dij, var, plan_dose, plan_var, st)
               : run_dij<double>(c, stream, pd, spec, bix, nbatch, rel, o, dij, var, plan_dose, plan_var, st);
}

extern "C" int omc_dij_assemble(omc_ctx *c, int phantom_id, int nbixels, int nbatch, const int64_t *nperbatch,
                                const double *endep, double rel_threshold, const omc_run_opts *opts, omc_csc *dij,
                                omc_csc *var, omc_run_stats *stats) {
    PhantomDev *pd = c ? omc_get_phantom(c, phantom_id) : nullptr;
    if (!pd) return omc_seterr(OMC_ERR_ARG, "no phantom %d", phantom_id);
    if (nbixels < 1 || nbatch < 1 || nbatch > 255 || !nperbatch || !endep || (!dij && !var))
        return omc_seterr(OMC_ERR_ARG, "omc_dij_assemble: bad arguments");
    omc_run_opts o;
    omc_run_opts_default(&o);
    if (opts) o = *opts;
    o.rng = OMC_RNG_STREAMS;
    o.count_events = 0;
    int rc = omc_check_opts(&o);
    if (rc) return rc;
    // the histories per batch as a beamlet set: the plan and the finish read
    // nothing else of it
    std::vector<int64_t> nh(nbixels);
    for (int b = 0; b < nbixels; b++) {
        if (nperbatch[b] < 1) return omc_seterr(OMC_ERR_ARG, "omc_dij_assemble: beamlet %d has %lld histories per batch",
                                                b, (long long)nperbatch[b]);
        nh[b] = nperbatch[b] * nbatch;
    }
    omc_beamlets_desc bix;
    memset(&bix, 0, sizeof bix);
    bix.nbixels = nbixels;
    bix.nbeams = 1;
    bix.nhist = nh.data();
    omc_spectrum_desc spec;
    memset(&spec, 0, sizeof spec);
    (void)hipSetDevice(c->device);
    if (stats) memset(stats, 0, sizeof *stats);
    if (dij) memset(dij, 0, sizeof *dij);
    if (var) memset(var, 0, sizeof *var);
    DijGiven given = {endep};
    omc_csc scratch;
    memset(&scratch, 0, sizeof scratch);
    rc = run_dij<float>(c, c->stream, pd, &spec, &bix, nbatch, rel_threshold, &o, dij ? dij : &scratch, var, nullptr,
                        nullptr, stats, &given);
    if (!dij) omc_csc_free(&scratch);
    return rc;
}

extern "C" int omc_run_dij_ex(omc_ctx *c, int phantom_id, const omc_spectrum_desc *spec, const omc_beamlets_desc *bix,
                              int nbatch, double rel_threshold, const omc_run_opts *opts, omc_csc *dij, omc_csc *var,
                              double *plan_dose, double *plan_var, omc_run_stats *stats) {
    PhantomDev *pd = c ? omc_get_phantom(c, phantom_id) : nullptr;
    if (!pd) return omc_seterr(OMC_ERR_ARG, "no phantom %d", phantom_id);
    return omc_dij_run(c, c->stream, pd, spec, bix, nbatch, rel_threshold, opts, dij, var, plan_dose, plan_var, stats);
}

extern "C" int omc_run_dij(omc_ctx *c, int phantom_id, const omc_spectrum_desc *spec, const omc_beamlets_desc *bix,
                           int nbatch, double rel_threshold, const omc_run_opts *opts, omc_csc *dij, omc_csc *var,
                           omc_run_stats *stats) {
    return omc_run_dij_ex(c, phantom_id, spec, bix, nbatch, rel_threshold, opts, dij, var, nullptr, nullptr, stats);
}

extern "C" void omc_csc_free(omc_csc *m) {
    if (!m) return;
    free(m->jc);
    if (m->device == 1) {
        if (m->ir) (void)hipFree(m->ir);
        if (m->val) (void)hipFree(m->val);
    } else if (m->device == 2) {
        pinned_block_free(m->ir);
        pinned_block_free(m->val);
    } else {
        free(m->ir);
        free(m->val);
    }
    m->jc = m->ir = nullptr;
    m->val = nullptr;
    m->nnz = 0;
    m->device = 0;
}

extern "C" int omc_csc_download(const omc_csc *m, int64_t *ir, double *val) {
    if (!m || m->device != 1 || (m->nnz > 0 && (!ir || !val)))
        return omc_seterr(OMC_ERR_ARG, "omc_csc_download: a device matrix and host arrays of nnz entries");
    if (m->nnz == 0) return OMC_OK;
    CK(hipMemcpy(ir, m->ir, sizeof(int64_t) * m->nnz, hipMemcpyDeviceToHost));
    CK(hipMemcpy(val, m->val, sizeof(double) * m->nnz, hipMemcpyDeviceToHost));
    return OMC_OK;
}

// ------------------------------------------------------------------ request server
// Continuous batching over HIP streams: a dispatcher admits queued requests
// in arrival order while their device memory fits the budget and a worker
// slot is free; each admitted request runs on its own stream (default
// priority), so kernels of concurrent requests overlap on the device and a
// request that arrives while others run starts as soon as memory and a slot
// free up.  Results do not depend on the mix: every history's random stream
// is keyed by (seed, beamlet, batch, history).
struct omc_server {
    omc_ctx *ctx = nullptr;
    int64_t budget = 0;
    int max_inflight = 1;
    std::mutex mu;
    std::condition_variable cv;
    struct Req {
        int64_t id;
        omc_request r;
        // owned copies of the caller's arrays
        std::vector<double> xb, yb, zb, rho, c1, c2, bsrc, corner, side1, side2;
        std::vector<int> med, ibeam;
        std::vector<int64_t> nhist;
        size_t bytes = 0;
        omc_result res;
    };
    std::deque<Req *> queue;
    std::map<int64_t, Req *> all;
    int64_t next_id = 1;
    int64_t in_use = 0;
    int running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::thread dispatcher;
    std::deque<Req *> ready;            // admitted, waiting for a worker
    std::chrono::steady_clock::time_point t0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

size_t request_bytes(omc_server *s, const omc_request &r) {
    const omc_phantom_desc &p = r.phantom;
    long nreg = (long)p.nx * p.ny * p.nz + 1;
    size_t b = omc_phantom_bytes(&p);
    long long tgt = default_target(s->ctx->device, r.nsplit);
    size_t ev = r.precision == OMC_FP64 ? event_bytes<double>(tgt) : event_bytes<float>(tgt);
    // dose: every batch in one job, each with its own tallies
    if (r.kind == OMC_REQ_DOSE) return b + (size_t)(r.nbatch + 2) * nreg * 8 + ev;
    // Dij: the plan with the fewest transport jobs that the request's share
    // of the budget holds (at least one beamlet and one batch per job)
    omc_run_opts o;
    omc_run_opts_default(&o);
    o.precision = r.precision == OMC_FP64 ? OMC_FP64 : OMC_FP32;
    DijPlan plan;
    size_t share = (size_t)s->budget / (size_t)s->max_inflight;
    return b + dij_plan(&o, nreg, r.bix.nbixels, r.nbatch, tgt, share > b ? share - b : 1, plan);
}

void run_request(omc_server *s, omc_server::Req *q) {
    omc_ctx *c = s->ctx;
    (void)hipSetDevice(c->device);
    hipStream_t st = nullptr;
    omc_result &res = q->res;
    int rc = OMC_OK;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        rc = omc_seterr(OMC_ERR_DEVICE, "stream creation failed");
    }
    PhantomDev *pd = nullptr;
    omc_phantom_desc pdsc = q->r.phantom;
    if (!rc) rc = omc_phantom_make(c, &pdsc, &pd, st);
    omc_run_opts o;
    omc_run_opts_default(&o);
    o.precision = q->r.precision == OMC_FP64 ? OMC_FP64 : OMC_FP32;
    o.form = OMC_FORM_EVENT;
    o.seed = q->r.seed;
    o.nsplit = q->r.nsplit;
    o.batches_per_job = q->r.kind == OMC_REQ_DOSE ? q->r.nbatch : 0;
    // concurrent requests already overlap one another's drains; extra group
    // streams per Dij request only contend for the hardware queues (served
    // mix at half load: p50 1.00 s with up to 3 streams per request, 0.81 s
    // with one; Dij requests 2.13 s against 1.01 s)
    if (s->max_inflight > 1) o.streams = 1;
    omc_run_stats stats;
    memset(&stats, 0, sizeof stats);
    if (!rc && q->r.kind == OMC_REQ_DOSE) {
        res.nreg = pd->nreg;
        // pooled pinned blocks (every entry is written by the run), else ordinary memory
        res.sum = (double *)pinned_block_alloc(sizeof(double) * pd->nreg);
        res.sum2 = (double *)pinned_block_alloc(sizeof(double) * pd->nreg);
        if (!res.sum) res.sum = (double *)calloc(pd->nreg, sizeof(double));
        if (!res.sum2) res.sum2 = (double *)calloc(pd->nreg, sizeof(double));
        if (!res.sum || !res.sum2) rc = omc_seterr(OMC_ERR_NOMEM, "result arrays");
        if (!rc)
            rc = omc_dosxyz_run(c, st, pd, &q->r.dose_src, q->r.nperbatch, q->r.nbatch, &o, res.sum, res.sum2, &stats);
    } else if (!rc) {
        size_t phb = omc_phantom_bytes(&pdsc);
        o.hbm_budget = (int64_t)(q->bytes > phb ? q->bytes - phb : 0);
        rc = omc_dij_run(c, st, pd, &q->r.spec, &q->r.bix, q->r.nbatch, q->r.rel_threshold, &o, &res.dij, nullptr,
                         nullptr, nullptr, &stats);
    }
    if (pd) omc_phantom_destroy(pd);
    if (st) (void)hipStreamDestroy(st);
    std::lock_guard<std::mutex> lk(s->mu);
    res.error = rc;
    if (rc) snprintf(res.message, sizeof res.message, "%s", omc_host_last_error());
    res.state = rc ? OMC_REQ_FAILED : OMC_REQ_DONE;
    res.histories = stats.histories;
    res.end_t = s->now();
    s->in_use -= (int64_t)q->bytes;
    s->running--;
    s->cv.notify_all();
}

void worker_loop(omc_server *s) {
    for (;;) {
        omc_server::Req *q = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->ready.empty(); });
            if (s->ready.empty()) return;
            q = s->ready.front();
            s->ready.pop_front();
            q->res.state = OMC_REQ_RUNNING;
            q->res.start_t = s->now();
        }
        run_request(s, q);
    }
}

void dispatch_loop(omc_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] {
            if (s->stop) return true;
            if (s->queue.empty() || s->running >= s->max_inflight) return false;
            omc_server::Req *q = s->queue.front();
            return s->in_use == 0 || s->in_use + (int64_t)q->bytes <= s->budget;
        });
        if (s->stop) return;
        omc_server::Req *q = s->queue.front();
        s->queue.pop_front();
        s->in_use += (int64_t)q->bytes;
        s->running++;
        s->ready.push_back(q);
        s->cv.notify_all();
    }
}

}  // namespace

extern "C" int omc_server_create(omc_ctx *ctx, int64_t hbm_budget_bytes, int max_inflight, omc_server **srv) {
    if (!ctx || !srv || max_inflight < 1 || max_inflight > 16) return omc_seterr(OMC_ERR_ARG, "omc_server_create: bad arguments");
    omc_server *s = new omc_server();
    s->ctx = ctx;
    if (hbm_budget_bytes <= 0) {
        // 3/4 of the free memory, at most 96 GiB: past ~100 GiB held by the
        // process, allocations take seconds to map and unmap
        size_t freeb = 0, totalb = 0;
        (void)hipSetDevice(ctx->device);
        if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) freeb = (size_t)8 << 30;
        freeb += devcache_held();
        hbm_budget_bytes = (int64_t)std::min(freeb * 3 / 4, (size_t)96 << 30);
    }
    s->budget = hbm_budget_bytes;
    s->max_inflight = max_inflight;
    s->t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < max_inflight; k++) s->workers.emplace_back(worker_loop, s);
    s->dispatcher = std::thread(dispatch_loop, s);
    *srv = s;
    return OMC_OK;
}

extern "C" int omc_server_submit(omc_server *s, const omc_request *r, int64_t *req_id) {
    if (!s || !r || !req_id) return omc_seterr(OMC_ERR_ARG, "omc_server_submit: bad arguments");
    if (r->kind != OMC_REQ_DOSE && r->kind != OMC_REQ_DIJ) return omc_seterr(OMC_ERR_ARG, "request kind %d", r->kind);
    const omc_phantom_desc &p = r->phantom;
    if (p.nx < 1 || p.ny < 1 || p.nz < 1 || !p.xb || !p.yb || !p.zb || !p.med || !p.rho || r->nbatch < 1)
        return omc_seterr(OMC_ERR_ARG, "omc_server_submit: bad phantom or batches");
    if (r->nsplit < 0 || r->nsplit > 64) return omc_seterr(OMC_ERR_ARG, "nsplit %d (1..64)", r->nsplit);
    auto *q = new omc_server::Req();
    q->r = *r;
    if (q->r.nsplit < 1) q->r.nsplit = 1;
    long nvox = (long)p.nx * p.ny * p.nz;
    q->xb.assign(p.xb, p.xb + p.nx + 1);
    q->yb.assign(p.yb, p.yb + p.ny + 1);
    q->zb.assign(p.zb, p.zb + p.nz + 1);
    q->med.assign(p.med, p.med + nvox);
    q->rho.assign(p.rho, p.rho + nvox);
    q->r.phantom.xb = q->xb.data();
    q->r.phantom.yb = q->yb.data();
    q->r.phantom.zb = q->zb.data();
    q->r.phantom.med = q->med.data();
    q->r.phantom.rho = q->rho.data();
    const omc_spectrum_desc &sp = r->kind == OMC_REQ_DOSE ? r->dose_src.spec : r->spec;
    if (sp.spectrum) {
        if (!sp.cdfinv1 || !sp.cdfinv2) { delete q; return omc_seterr(OMC_ERR_ARG, "spectrum without inverse CDF"); }
        q->c1.assign(sp.cdfinv1, sp.cdfinv1 + OMC_INVDIM);
        q->c2.assign(sp.cdfinv2, sp.cdfinv2 + OMC_INVDIM);
    }
    omc_spectrum_desc &dsp = r->kind == OMC_REQ_DOSE ? q->r.dose_src.spec : q->r.spec;
    dsp.cdfinv1 = sp.spectrum ? q->c1.data() : nullptr;
    dsp.cdfinv2 = sp.spectrum ? q->c2.data() : nullptr;
    if (r->kind == OMC_REQ_DIJ) {
        const omc_beamlets_desc &b = r->bix;
        if (b.nbixels < 1 || b.nbeams < 1 || !b.ibeam || !b.src || !b.corner || !b.side1 || !b.side2) {
            delete q;
            return omc_seterr(OMC_ERR_ARG, "omc_server_submit: bad beamlets");
        }
        q->ibeam.assign(b.ibeam, b.ibeam + b.nbixels);
        q->bsrc.assign(b.src, b.src + 3 * b.nbeams);
        q->corner.assign(b.corner, b.corner + 3 * b.nbixels);
        q->side1.assign(b.side1, b.side1 + 3 * b.nbixels);
        q->side2.assign(b.side2, b.side2 + 3 * b.nbixels);
        if (b.nhist) q->nhist.assign(b.nhist, b.nhist + b.nbixels);
        omc_beamlets_desc &d = q->r.bix;
        d.ibeam = q->ibeam.data();
        d.src = q->bsrc.data();
        d.corner = q->corner.data();
        d.side1 = q->side1.data();
        d.side2 = q->side2.data();
        d.nhist = b.nhist ? q->nhist.data() : nullptr;
    } else if (r->nperbatch < 1) {
        delete q;
        return omc_seterr(OMC_ERR_ARG, "omc_server_submit: no histories");
    }
    memset(&q->res, 0, sizeof q->res);
    q->bytes = request_bytes(s, q->r);
    std::lock_guard<std::mutex> lk(s->mu);
    if ((int64_t)q->bytes > s->budget) {
        delete q;
        return omc_seterr(OMC_ERR_BUDGET, "request needs %zu bytes, the server budget is %lld", request_bytes(s, *r),
                          (long long)s->budget);
    }
    q->id = s->next_id++;
    q->res.state = OMC_REQ_QUEUED;
    q->res.submit_t = s->now();
    s->all[q->id] = q;
    s->queue.push_back(q);
    *req_id = q->id;
    s->cv.notify_all();
    return OMC_OK;
}

extern "C" int omc_server_poll(omc_server *s, int64_t id, int *state) {
    if (!s || !state) return omc_seterr(OMC_ERR_ARG, "omc_server_poll: bad arguments");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return omc_seterr(OMC_ERR_ARG, "no request %lld", (long long)id);
    *state = it->second->res.state;
    return OMC_OK;
}

extern "C" int omc_server_wait(omc_server *s, int64_t id, omc_result *res) {
    if (!s || !res) return omc_seterr(OMC_ERR_ARG, "omc_server_wait: bad arguments");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return omc_seterr(OMC_ERR_ARG, "no request %lld", (long long)id);
    omc_server::Req *q = it->second;
    s->cv.wait(lk, [&] { return q->res.state == OMC_REQ_DONE || q->res.state == OMC_REQ_FAILED; });
    *res = q->res;
    s->all.erase(it);
    delete q;
    return res->error;
}

extern "C" void omc_result_free(omc_result *res) {
    if (!res) return;
    if (!pinned_block_free(res->sum)) free(res->sum);
    if (!pinned_block_free(res->sum2)) free(res->sum2);
    res->sum = res->sum2 = nullptr;
    omc_csc_free(&res->dij);
}

extern "C" void omc_server_destroy(omc_server *s) {
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        // finish what is admitted or queued
        s->cv.wait(lk, [&] { return s->queue.empty() && s->ready.empty() && s->running == 0; });
        s->stop = true;
        s->cv.notify_all();
    }
    s->dispatcher.join();
    for (auto &t : s->workers) t.join();
    for (auto &kv : s->all) {
        omc_result_free(&kv.second->res);
        delete kv.second;
    }
    delete s;
}
