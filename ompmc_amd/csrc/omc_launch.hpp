// Launchers of the templated kernels, instantiated per precision in
// omc_k32.hip (float) and omc_k64.hip (double, built without FP contraction
// so that it rounds like the CPU reference).
#pragma once
#include <hip/hip_runtime.h>
#include "omc_event.hpp"
#include "omc_dij.hpp"

namespace omc {

template <class R> hipError_t launch_history(const HistArgs<R> &a, int rng, bool count, int grid, int block,
                                             hipStream_t st);
template <class R> hipError_t launch_sampler(const SampArgs<R> &a, hipStream_t st);
hipError_t launch_accumulate(double *e, double *sum, double *sum2, long n, double scale, double *dep,
                             hipStream_t st);
// FP64 atomic throughput probe (grid x 256 lanes, nper scattered adds each)
hipError_t launch_atomic_probe(double *t, unsigned long long mask, int grid, int nper, unsigned long long seed,
                               hipStream_t st);

// Everything the rounds of an event-queue job need.
template <class R> struct EvLaunch {
    const Tables<R> *t;
    const Geom<R> *g;
    uint64_t seed;
    Pool<R> pool[2][2];          // [round parity][photon, charged]
    EvCtl *ctl;
    long long cap;
    double *endep;
    long slot_stride;
    uint8_t *touched;            // optional tile marks of the tallies, touch_stride per slot
    long touch_stride;
    double *ebal;                // optional [EB_N] energy balance of the run
    unsigned long long *counters;
    int budget;
    int turns;
    int max_age;
    int delta;
    int nsplit;                  // photon splitting factor; > 1 transports photons with k_event_split
    int graph;                   // 1: untimed chunks of rounds replay a HIP graph
    const CBin<R> *ebin;         // the electron half of the charged bins (Tables::cbin)
    int eb_words;                // > 0: the charged kernel copies it to LDS (16-byte words)
    int tlist;                   // deposit-list entries per charged-kernel lane in LDS (0: direct atomics)
    SrcArgs<R> src;              // source of the job (ctl and out are set per round)
    int grid_event, grid_charged, grid_source;   // workgroups: photon kernels, charged kernel, source
};

// Enqueues rounds r0 .. r0+nround-1 of the job whose control block is L.ctl.
template <class R> hipError_t launch_rounds(const EvLaunch<R> &L, long r0, int nround, bool count, hipStream_t st);
// One part of round r: 0 source, 1 photon events, 2 charged events, 3 round end.
template <class R> hipError_t launch_round_part(const EvLaunch<R> &L, long r, int part, bool count, hipStream_t st);
// The same kernels as a graph, nodes added one by one in stream order.
template <class R> hipError_t build_rounds_graph(const EvLaunch<R> &L, long r0, int nround, bool count,
                                                 hipGraph_t *out);
hipError_t launch_job_start(EvCtl *c, long long nhist, long long target, long long cap, long long inj_max, int dest,
                            hipStream_t st);

// Dij post-processing (omc_dij.hpp), FP64
hipError_t launch_dij_fold_finish(double *e, uint8_t *touched, int ng, int nbatch, long nreg, const double *nper,
                                  double *dose, double *var, uint8_t *any, double *carry, const double *mass,
                                  const double *rho, double *blockmax, double *tot_dose, double *tot_var, double *dep,
                                  hipStream_t st);
hipError_t launch_accum_slots(double *e, double *sum, double *sum2, long n, long stride, int nslot, const double *nper,
                              double *dep, hipStream_t st);
hipError_t launch_dij_finish(double *sum, double *sum2, double *carry, int ng, long nreg, const double *mass,
                             const double *rho, int nbatch, double *blockmax, double *tot_dose, double *tot_var,
                             hipStream_t st);
hipError_t launch_dij_select(const double *dose, const uint8_t *any, long nreg, int ng, const double *blockmax,
                             double rel, double *thresh, int *counts, hipStream_t st);
hipError_t launch_dij_compact(const double *dose, const double *var, const uint8_t *any, long nreg, int ng,
                              const double *thresh, const int64_t *off, int64_t *ir, double *val, double *vval,
                              hipStream_t st);

}  // namespace omc
