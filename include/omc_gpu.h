/*
 * omc_gpu.h -- C ABI of the MI355X transport engine.
 *
 * The reference (e0404/ompMC) is driven by two user codes that each run the
 * whole simulation inside one process with global state: the CLI
 * ucodes/omc_dosxyz/omc_dosxyz.c (main, :1072-1307) and the MATLAB mex
 * ucodes/omc_matrad/omc_matrad.c (mexFunction, :1417-1789; legacy
 * matRad_ompInterface.c:41-652).  Between them and the engine sits the
 * per-step plugin API of src/ompmc.h:39-41,422 (ausgab / howfar / hownear /
 * initRegions over the thread-private particle stack), which a GPU cannot
 * call back into.  This header therefore replaces the engine + ucode pair at
 * the level those front ends need: build the tables once, upload a phantom,
 * run a dose or a dose-influence (Dij) simulation, read the results.  All
 * arguments are plain C pointers and sizes; every call returns 0 or an OMC_ERR_*
 * code (never exits) and omc_last_error() describes the failure.
 *
 * Threading: a context may be used from one thread at a time, except the
 * serving calls (omc_server_*), which are thread safe.
 */
#ifndef OMC_GPU_H
#define OMC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMC_API_VERSION 1

/* error codes (also returned by the host layer) */
#define OMC_OK 0
#define OMC_ERR_IO 1
#define OMC_ERR_FORMAT 2
#define OMC_ERR_MEDIUM 3
#define OMC_ERR_LIMIT 4
#define OMC_ERR_ARG 5
#define OMC_ERR_DEVICE 6
#define OMC_ERR_NOMEM 7
#define OMC_ERR_BUDGET 8
#define OMC_ERR_STACK 9
#define OMC_ERR_CANCELLED 10

/* transport precision: FP32 transport with FP64 tallies, or FP64 throughout */
#define OMC_FP32 32
#define OMC_FP64 64
/* execution form: one lane per history (plain), or event queues (batched) */
#define OMC_FORM_HISTORY 0
#define OMC_FORM_EVENT 1
/* random numbers: keyed streams -- in the history form one Philox4x32-10
 * stream per history, keyed by (seed, history); in the event form one
 * xoshiro128** stream per particle, seeded by splitmix64 from a key of
 * (seed, history) or of the parent's key and draw count -- or a replayed
 * 24-bit stream (RANMAR output) with per-history offsets */
#define OMC_RNG_STREAMS 0
#define OMC_RNG_REPLAY 1

typedef struct omc_ctx omc_ctx;

/* Last error message of the calling thread. */
const char *omc_last_error(void);
int omc_api_version(void);

/* Builds all physics tables on the host (reference initMediaData,
 * src/ompmc.c:5468-5505: PEGS4, XCOM, pgs4form, msnew, spinms) and uploads
 * FP32 and FP64 copies to device `device`.  media: nmed PEGS4 names
 * (reference media.med_names, filled from the egsphant header or mcGeo.material). */
int omc_ctx_create(omc_ctx **ctx, int nmed, const char *const *media, const char *pegs_file,
                   const char *pgs4form_file, const char *data_folder, int device);
void omc_ctx_destroy(omc_ctx *ctx);
/* Host FP64 view of a physics table by name (names of the reference's
 * arrays, e.g. "gmfp0", "esig1", "spin_rej"); for verification. */
int omc_ctx_table(omc_ctx *ctx, const char *name, const double **data, long *n);

/* Voxel phantom (reference struct Geom + initRegions, omc_dosxyz.c:48-59,
 * 890-962; omc_matrad.c:364-453, 1103-1175).  med is 1-based per voxel (0 =
 * vacuum), x fastest; ecut/pcut are the global cuts (total energies, MeV). */
typedef struct omc_phantom_desc {
    int nx, ny, nz;
    const double *xb, *yb, *zb;
    const int *med;
    const double *rho;
    double ecut, pcut;
} omc_phantom_desc;

int omc_phantom_upload(omc_ctx *ctx, const omc_phantom_desc *desc, int *phantom_id);
int omc_phantom_release(omc_ctx *ctx, int phantom_id);

/* Energy source: mono energy or the spectrum inverse CDF of the reference
 * (omc_dosxyz.c:396-510): cdfinv1/cdfinv2 hold OMC_INVDIM = 1000 entries. */
typedef struct omc_spectrum_desc {
    int spectrum;
    double energy;
    const double *cdfinv1, *cdfinv2;
} omc_spectrum_desc;

/* Spectrum file reader (omc_dosxyz.c:396-510; omc_matrad.c:627-769): fills
 * cdfinv1/cdfinv2 (OMC_INVDIM = 1000 entries each) from a spectrum file of
 * the reference's format (title line; nbins, emin, imode; upper edge and
 * weight per bin). */
int omc_spectrum_file(const char *path, double *cdfinv1, double *cdfinv2);

/* omc_dosxyz collimated point source (omc_dosxyz.c:521-617, 964-1068). */
typedef struct omc_dosxyz_source_desc {
    omc_spectrum_desc spec;
    int charge;
    double ssd;
    double collimator[4];   /* xl xu yl yu */
} omc_dosxyz_source_desc;

/* Options of a run (precision, form, random numbers, launch shape). */
typedef struct omc_run_opts {
    int precision;          /* OMC_FP32 | OMC_FP64 */
    int form;               /* OMC_FORM_HISTORY | OMC_FORM_EVENT */
    int rng;                /* OMC_RNG_STREAMS | OMC_RNG_REPLAY */
    uint64_t seed;          /* run seed: part of every stream's key */
    int nsplit;             /* photon splitting factor (vrt.nsplit, 1..64); both forms.  In the event
                               form split photons fly voxel by voxel (no delta tracking) */
    const uint32_t *replay;       /* OMC_RNG_REPLAY: draws (24-bit integers) */
    const int64_t *replay_offsets;/* first draw of each history, nhist+1 entries */
    int64_t replay_len;
    int grid, block;        /* 0 = automatic */
    int count_events;       /* 1: fill the event counters and per-kernel times of omc_run_stats;
                               2: per-kernel times only (events around every transport kernel) */
    int64_t event_target;   /* event form: particles in flight per round (0 = automatic: 4096 per compute
                               unit, four times that with nsplit > 1) */
    int electron_budget;    /* event form: condensed-history steps per charged-particle event (0 = automatic) */
    int group;              /* Dij: beamlets transported together (0 = as many as hbm_budget allows,
                               1 = one beamlet at a time, the reference's loop order) */
    int64_t hbm_budget;     /* device bytes a Dij run may use for its tallies (0 = half the free memory,
                               at most 96 GiB); dose runs: a quarter of it bounds the batch tallies */
    int voxel_walk;         /* event form: 1 = photons fly voxel by voxel as the reference's photon()
                               does; 0 = delta tracking with the phantom's majorant (same distribution) */
    int electron_turns;     /* event form: refills of electron_budget steps an electron may hold its lane
                               within one round before it is requeued (0 = automatic: 24, -1 = no limit) */
    int batches_per_job;    /* event form: batches transported together, each into tallies of its own
                               (0 = as many as the memory budget holds, 1 = one batch per job) */
    int streams;            /* event-form Dij: beamlet groups transported concurrently, each on its own
                               stream (0 = automatic: up to 3 when the run has several groups; 1..4) */
    int64_t max_particle_steps; /* event form: condensed-history steps after which an electron counts as
                               stuck, is dropped and the run fails with OMC_ERR_LIMIT (0 = 2^20); history
                               form: transport calls per history (photon flights, electron transports up
                               to a secondary) after which the history is cut and the run fails with
                               OMC_ERR_LIMIT (0 = 10^8) */
    int csc_on_device;      /* Dij: 1 = leave the matrices in device memory (omc_csc.device = 1) for a
                               consumer on the GPU; 0 = host arrays, as the mex interface needs */
    int graph;              /* event form: 0 = each job's chunks of rounds replay a HIP graph of its first
                               chunk (eager launches when count_events times the kernels); -1 = eager
                               launches */
    int lds_tables;         /* event form, FP32: 0 = the charged-particle kernel copies the electron
                               coefficients of each energy bin (up to 64 KiB) into LDS once per workgroup
                               and reads them there; -1 = reads them from the tables in device memory.
                               Same values either way */
} omc_run_opts;

/* Kernel classes of omc_run_stats.kernel_seconds / kernel_launches */
#define OMC_K_HISTORY 0     /* plain history kernel */
#define OMC_K_SOURCE 1      /* event form: primaries */
#define OMC_K_PHOTON 2      /* event form: photon events */
#define OMC_K_CHARGED 3     /* event form: electron/positron events */
#define OMC_K_OTHER 4       /* batch statistics, Dij finish and compaction */

typedef struct omc_run_stats {
    double gpu_seconds;             /* device time of the transport kernels */
    double wall_seconds;            /* whole call */
    double histories;
    uint64_t n_howfar, n_hownear, n_cross, n_dep, n_rn, n_photon, n_electron;
    int errors;                     /* bit 0 stack or pool overflow, 1 replay underrun, 2 step guard */
    /* count_events runs: device time and launches per kernel class, and the
     * howfar/hownear/crossing/deposit counts of the photon and charged-particle
     * event kernels ([0] photon, [1] charged) */
    double kernel_seconds[5];
    uint64_t kernel_launches[5];
    uint64_t ev_howfar[2], ev_hownear[2], ev_cross[2], ev_dep[2];
    /* event form: events processed, rounds, largest number of particles in flight */
    uint64_t events_photon, events_charged, rounds, peak_inflight;
    /* Dij: beamlet groups, beamlets per group, non-zeros */
    int64_t groups, group_size, nnz;
    /* event form: batches transported together in one job; Dij: concurrent streams */
    int64_t batches_per_job, streams;
    /* energy balance (FP64, MeV, weighted): energy of the primaries (photons E,
     * electrons E - m, positrons E + m), of photons that left the phantom (the
     * reference discards them untallied, src/ompmc.c:2008-2017), dropped
     * without a deposit (quirk Q6 of src/ompmc.c:5051-5056, iteration
     * guards), and deposited in all regions including region 0.  Without
     * photon splitting e_source = e_escape + e_lost + e_deposit up to
     * rounding; with it, in expectation. */
    double e_source, e_escape, e_lost, e_deposit;
    /* event-form runs with count_events: per kernel class, the time during
     * which at least one launch of the class ran (launches of concurrent
     * streams overlap, so kernel_seconds, their sum, can exceed it) */
    double kernel_busy_seconds[5];
    /* event form: chunks of rounds enqueued by replaying a HIP graph, and
     * jobs whose graph the runtime refused (they ran eager launches) */
    uint64_t graph_replays, graph_fallbacks;
    /* event form: bytes of electron coefficients the charged-particle kernel copied into LDS per
     * workgroup (0: it read them from device memory; see omc_run_opts.lds_tables) */
    uint64_t lds_table_bytes;
} omc_run_stats;

/* Defaults: FP32, event form, keyed streams, seed 1802, nsplit 1, automatic launch. */
void omc_run_opts_default(omc_run_opts *o);

/* Dose run with omc_dosxyz semantics (main, omc_dosxyz.c:1206-1263):
 * nbatch batches of nperbatch histories; after each batch the per-region
 * energy e is folded into sum += e, sum2 += e^2 (accumEndep, :696-717).
 * sum/sum2 are host arrays of nvox+1 doubles (region 0 = outside). */
int omc_run_dosxyz(omc_ctx *ctx, int phantom_id, const omc_dosxyz_source_desc *src, int64_t nperbatch,
                   int nbatch, const omc_run_opts *opts, double *sum, double *sum2, omc_run_stats *stats);

/* Dose and relative uncertainty from the batch sums (accumulateResults,
 * omc_dosxyz.c:719-799), and the .3ddose writer (:801-886). */
int omc_dosxyz_dose(omc_ctx *ctx, int phantom_id, const double *sum, const double *sum2, int64_t nperbatch,
                    int nbatch, double *dose, double *unc);
int omc_write_3ddose_file(omc_ctx *ctx, int phantom_id, const char *path, const double *dose, const double *unc);

/* matRad beamlet set (mcSrc of omc_matrad.c:780-829). */
typedef struct omc_beamlets_desc {
    int nbixels, nbeams;
    const int *ibeam;              /* 0-based beam of each bixel */
    const double *src;             /* [nbeams*3] source points */
    const double *corner, *side1, *side2;   /* [nbixels*3] */
    int gaussian;                  /* sourceGeometry 'gaussian' */
    double sigma;                  /* sourceGaussianWidth */
    const int64_t *nhist;          /* histories per bixel, or NULL for nhist_each */
    int64_t nhist_each;
} omc_beamlets_desc;

/* Compressed sparse column matrix (MATLAB layout): column j holds rows
 * ir[jc[j]..jc[j+1]-1] with values val[]; allocated by the library, released
 * with omc_csc_free.  device = 1 (runs with csc_on_device): ir and val are
 * device addresses on the run's device; 0 or 2: host memory (2: pinned
 * blocks the library reuses once released).  jc is always on the host. */
typedef struct omc_csc {
    int64_t nrows, ncols, nnz;
    int64_t *jc, *ir;
    double *val;
    int device;
} omc_csc;
void omc_csc_free(omc_csc *m);
/* No counterpart in the reference: copies the rows and values of a matrix
 * kept on the device (device = 1) into host arrays of nnz entries. */
int omc_csc_download(const omc_csc *m, int64_t *ir, double *val);

/* Dose-influence matrix with omc_matrad semantics (mexFunction,
 * omc_matrad.c:1576-1763): per bixel nbatch batches, dose per history in Gy,
 * entries strictly above rel_threshold times the bixel maximum, rows are
 * voxel indices (x fastest, 0-based).  var (optional) receives the variance
 * matrix of the same pattern, including the reference's carry-over of the
 * previous bixel's variance array. */
int omc_run_dij(omc_ctx *ctx, int phantom_id, const omc_spectrum_desc *spec, const omc_beamlets_desc *bix,
                int nbatch, double rel_threshold, const omc_run_opts *opts, omc_csc *dij, omc_csc *var,
                omc_run_stats *stats);
/* omc_run_dij with two optional outputs of nvox doubles each: the dose of
 * all beamlets summed with unit weights, and the variance of that sum (the
 * sum over beamlets of the batch estimate of the variance of the mean,
 * (sum x^2/n - mean^2)/(n-1), without the reference's carry-over); no
 * counterpart in the reference, for plan-level statistics. */
int omc_run_dij_ex(omc_ctx *ctx, int phantom_id, const omc_spectrum_desc *spec, const omc_beamlets_desc *bix,
                   int nbatch, double rel_threshold, const omc_run_opts *opts, omc_csc *dij, omc_csc *var,
                   double *plan_dose, double *plan_var, omc_run_stats *stats);

/* The scoring and assembly of omc_matrad's beamlet loop on given batch
 * tallies, by the device kernels of omc_run_dij: accumEndep per batch
 * (omc_matrad.c:897-918), accumulateResults with iout = 1 (:920-1012: dose in
 * Gy per history, the variance with the carry of the previous beamlet's
 * variance array, rho < 0.044 zeroed), the maximum, the strict threshold
 * rel_threshold * max and the CSC append (:1603-1705).  endep holds nbixels x
 * nbatch arrays of nvox + 1 doubles, endep[(b * nbatch + k) * (nvox + 1) + irl]
 * = beamlet b's energy (MeV) in batch k at region irl (0 = outside), and
 * nperbatch[b] the histories per batch it was divided by.  opts (NULL: the
 * defaults) chooses the grouping (group, batches_per_job, streams,
 * hbm_budget) as for omc_run_dij.  The reference has the arithmetic inside its
 * mexFunction only; this entry point has no counterpart of its own.  It feeds
 * the Dij kernels with exact inputs for verification and lets a caller with
 * tallies from elsewhere reuse the assembly. */
int omc_dij_assemble(omc_ctx *ctx, int phantom_id, int nbixels, int nbatch, const int64_t *nperbatch,
                     const double *endep, double rel_threshold, const omc_run_opts *opts, omc_csc *dij, omc_csc *var,
                     omc_run_stats *stats);

/* No counterpart in the reference: omc_run_dosxyz in the history form for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Every buffer the run
 * needs beyond its inputs and outputs (the batch tally, the source's inverse
 * CDF, counters, error flags) lies in one fixed array of
 * OMC_NOALLOC_SCRATCH_BYTES that the library holds on each device it runs on
 * until the process ends; the call allocates nothing.  sum and sum2 are
 * DEVICE arrays of nvox + 1 doubles on the context's device that receive the
 * batch sums (as omc_run_dosxyz's host arrays do).  Limits: nvox + 1 <=
 * OMC_NOALLOC_MAX_REGIONS, form OMC_FORM_HISTORY, keyed random streams,
 * nsplit <= 8; other requests fail with OMC_ERR_ARG.  Calls that overlap on
 * one device share the array and race; the phantom and the context's tables
 * are uploaded beforehand by the ordinary (allocating) calls. */
#define OMC_NOALLOC_MAX_REGIONS 4194304
#define OMC_NOALLOC_SCRATCH_BYTES 33574656
int omc_run_dosxyz_noalloc(omc_ctx *ctx, int phantom_id, const omc_dosxyz_source_desc *src, int64_t nperbatch,
                           int nbatch, const omc_run_opts *opts, double *sum, double *sum2, omc_run_stats *stats);

/* Injected-random-number sampler evaluation for verification: ncase cases of
 * the golden-vector layout (kinds 1..16: compton, pair, photo, rayleigh,
 * brems, moller, bhabha, annih, rannih, msdist, mscat, sscat, drange, eloss,
 * photon, electron), inputs in[ncase*in_stride], draws rn with offsets
 * rn_off[ncase+1]; outputs out[ncase*out_stride] (see the test suite for the
 * layout) and rn_used[ncase].  Uses the phantom for geometry lookups. */
int omc_sampler_eval(omc_ctx *ctx, int phantom_id, int precision, int ncase, const int *kind, const double *in,
                     int in_stride, const uint32_t *rn, const int64_t *rn_off, double *out, int out_stride,
                     int64_t *rn_used);

/* ----------------------------------------------------------------- measurement */
/* No counterpart in the reference: the FP64 atomic throughput of the device,
 * the secondary roofline of the tallies.  Adds 1.0 at nops addresses drawn
 * uniformly from a table of nslot doubles (a power of two, 2^10..2^30) with
 * the hardware FP64 atomic add the tallies use; *ops_per_second is the rate
 * of one timed launch after a warm-up launch. */
int omc_atomic_probe(int device, int64_t nslot, int64_t nops, double *ops_per_second);

/* ----------------------------------------------------------------- serving */
/* A resident request loop on the context's device: requests (dose or Dij
 * over their own phantom) are admitted while their device memory fits the
 * HBM budget, and the transport of admitted requests shares the device. */
typedef struct omc_server omc_server;

#define OMC_REQ_DOSE 0
#define OMC_REQ_DIJ 1
#define OMC_REQ_QUEUED 0
#define OMC_REQ_RUNNING 1
#define OMC_REQ_DONE 2
#define OMC_REQ_FAILED 3

typedef struct omc_request {
    int kind;                         /* OMC_REQ_DOSE | OMC_REQ_DIJ */
    omc_phantom_desc phantom;         /* copied at submit */
    omc_dosxyz_source_desc dose_src;  /* OMC_REQ_DOSE */
    int64_t nperbatch;                /* OMC_REQ_DOSE */
    omc_spectrum_desc spec;           /* OMC_REQ_DIJ */
    omc_beamlets_desc bix;            /* OMC_REQ_DIJ, copied at submit */
    double rel_threshold;             /* OMC_REQ_DIJ */
    int nbatch;
    uint64_t seed;
    int precision;
    int nsplit;                       /* photon splitting factor, 1..64 (0 = 1) */
} omc_request;

typedef struct omc_result {
    int state;
    int error;
    double *sum, *sum2;               /* OMC_REQ_DOSE: nvox+1 each (library-owned until omc_result_free) */
    int64_t nreg;
    omc_csc dij;                      /* OMC_REQ_DIJ */
    double submit_t, start_t, end_t;  /* seconds on the server clock */
    double histories;
    char message[256];                /* OMC_REQ_FAILED: what failed (the worker thread's omc_last_error) */
} omc_result;

/* hbm_budget_bytes 0: 3/4 of the free device memory, at most 96 GiB */
int omc_server_create(omc_ctx *ctx, int64_t hbm_budget_bytes, int max_inflight, omc_server **srv);
int omc_server_submit(omc_server *srv, const omc_request *req, int64_t *req_id);
int omc_server_poll(omc_server *srv, int64_t req_id, int *state);
int omc_server_wait(omc_server *srv, int64_t req_id, omc_result *res);
void omc_result_free(omc_result *res);
void omc_server_destroy(omc_server *srv);

#ifdef __cplusplus
}
#endif
#endif
