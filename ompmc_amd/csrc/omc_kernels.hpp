// Kernel templates: primary sources, the plain history kernel, batch
// accumulation and the injected-random-number sampler kernel.  Instantiated
// for float in omc_k32.hip and for double in omc_k64.hip.
#pragma once
#include "omc_dev.hpp"

namespace omc {

// ------------------------------------------------------------------ sources
// Collimated point source of omc_dosxyz (omc_dosxyz.c:368-632, 964-1068).
template <class R> struct SrcDosxyz {
    int spectrum;
    R energy, deltak;
    const R *cdfinv1, *cdfinv2;
    int charge;
    R ssd, xinl, yinl, xsize, ysize;
    int ixinl, iyinl;
};

// Bixel source of omc_matrad (omc_matrad.c:1177-1413): a point (or gaussian
// focal spot) per beam shooting through a rectangle per bixel.
template <class R> struct SrcBeamlets {
    int spectrum;
    R energy, deltak;
    const R *cdfinv1, *cdfinv2;
    const int *ibeam;
    const R *src;                    // [nbeams*3]
    const R *corner, *side1, *side2; // [nbix*3]
    int gaussian;
    R sigma;
};

template <class R, class E> __device__ __forceinline__ R sample_energy(E &e, int spectrum, R energy, R deltak,
                                                                        const R *c1, const R *c2) {
    if (!spectrum) return energy;
    R r1 = e.rnd(), r2 = e.rnd();
    int k = (int)M<R>::fmin(deltak * r1, deltak - (R)1.0);
    return c1[k] + r2 * c2[k];
}

template <class R, class E> __device__ void init_history(E &e, const SrcDosxyz<R> &s, Part<R> &p) {
    const R RM = (R)0.5109989461;
    const Geom<R> &g = e.g;
    p.iq = s.charge;
    R ein = sample_energy<R>(e, s.spectrum, s.energy, s.deltak, s.cdfinv1, s.cdfinv2);
    p.e = p.iq != 0 ? ein + RM : ein;
    R rxyz;
    if (s.xsize == (R)0.0 || s.ysize == (R)0.0) {
        p.x = s.xinl;
        p.y = s.yinl;
        rxyz = M<R>::sqrt(s.ssd * s.ssd + p.x * p.x + p.y * p.y);
        p.w = s.ssd / rxyz;
    } else {
        R fw, r3;
        int trials = 0;
        do {
            r3 = e.rnd();
            p.x = r3 * s.xsize + s.xinl;
            r3 = e.rnd();
            p.y = r3 * s.ysize + s.yinl;
            r3 = e.rnd();
            rxyz = M<R>::sqrt(s.ssd * s.ssd + p.x * p.x + p.y * p.y);
            p.w = s.ssd / rxyz;
            fw = p.w * p.w * p.w;
        } while (r3 >= fw && !e.stuck(trials));
    }
    p.z = g.zb[0];
    p.u = p.x / rxyz;
    p.v = p.y / rxyz;
    int ix, iy;
    if (s.xsize == (R)0.0) {
        ix = s.ixinl;
    } else {
        ix = s.ixinl - 1;
        while (ix < g.nx - 1 && g.xb[ix + 1] < p.x) ix++;
    }
    if (s.ysize == (R)0.0) {
        iy = s.iyinl;
    } else {
        iy = s.iyinl - 1;
        while (iy < g.ny - 1 && g.yb[iy + 1] < p.y) iy++;
    }
    p.ix = ix; p.iy = iy; p.iz = 0;
    p.wt = (R)1.0;
}

// The reference's entry voxel along one axis (omc_matrad.c:1391-1406): a
// linear scan that stops at the first interior boundary b[i + 1] >= x, i.e.
// the number of interior boundaries b[1..n-1] below x.  Counted by
// bisection with the same comparison, so the index is the scan's, in
// log2(n) dependent loads instead of up to n.
template <class R> __device__ __forceinline__ int entry_index(const R *b, int n, R x) {
    int lo = 0, hi = n - 1;            // the answer lies in [lo, hi]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b[mid] < x) lo = mid;      // b[1..mid] are all below x (b increases)
        else hi = mid - 1;
    }
    return lo;
}

template <class R, class E> __device__ void init_history(E &e, const SrcBeamlets<R> &s, int ib, Part<R> &p) {
    const Geom<R> &g = e.g;
    const R tiny = sizeof(R) == 8 ? (R)(2.0 * 2.2250738585072014e-308) : (R)0;
    p.iq = 0;
    p.e = sample_energy<R>(e, s.spectrum, s.energy, s.deltak, s.cdfinv1, s.cdfinv2);
    R r1 = e.rnd(), r2 = e.rnd();
    const R *s1 = s.side1 + 3 * ib, *s2 = s.side2 + 3 * ib, *c = s.corner + 3 * ib;
    R xiso = r1 * s1[0] + r2 * s2[0] + c[0];
    R yiso = r1 * s1[1] + r2 * s2[1] + c[1];
    R ziso = r1 * s1[2] + r2 * s2[2] + c[2];
    int ibeam = s.ibeam[ib];
    R sp0 = s.src[3 * ibeam], sp1 = s.src[3 * ibeam + 1], sp2 = s.src[3 * ibeam + 2];
    if (s.gaussian) {
        R n1 = M<R>::sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
        R n2 = M<R>::sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
        R a = e.rnd(), b = e.rnd();
        R rr = M<R>::sqrt((R)-2.0 * M<R>::log(a));
        R th = (R)2.0 * (R)3.14159265358979323846 * b;
        R g0 = rr * M<R>::cos(th) * s.sigma, g1 = rr * M<R>::sin(th) * s.sigma;
        sp0 = sp0 + g0 * (s1[0] / n1) + g1 * (s2[0] / n2);
        sp1 = sp1 + g0 * (s1[1] / n1) + g1 * (s2[1] / n2);
        sp2 = sp2 + g0 * (s1[2] / n1) + g1 * (s2[2] / n2);
    }
    R xd = xiso - sp0, yd = yiso - sp1, zd = ziso - sp2;
    R vnorm = M<R>::sqrt(xd * xd + yd * yd + zd * zd);
    R u = -(xd) / vnorm, v = -(yd) / vnorm, w = -(zd) / vnorm;
    R ustep = sizeof(R) == 8 ? (R)1.7976931348623157e308 : (R)3.402823466e38f, d;
    if (u > (R)0) { d = (g.xb[g.nx] - xiso) / u; if (d < ustep) ustep = d; }
    if (u < (R)0) { d = -(xiso - g.xb[0]) / u; if (d < ustep) ustep = d; }
    if (v > (R)0) { d = (g.yb[g.ny] - yiso) / v; if (d < ustep) ustep = d; }
    if (v < (R)0) { d = -(yiso - g.yb[0]) / v; if (d < ustep) ustep = d; }
    if (w > (R)0) { d = (g.zb[g.nz] - ziso) / w; if (d < ustep) ustep = d; }
    if (w < (R)0) { d = -(ziso - g.zb[0]) / w; if (d < ustep) ustep = d; }
    p.x = xiso + ustep * u;
    p.y = yiso + ustep * v;
    p.z = ziso + ustep * w;
    p.u = -u; p.v = -v; p.w = -w;
    if (p.x < g.xb[0]) p.x = g.xb[0] + tiny;
    if (p.x > g.xb[g.nx]) p.x = g.xb[g.nx] - tiny;
    if (p.y < g.yb[0]) p.y = g.yb[0] + tiny;
    if (p.y > g.yb[g.ny]) p.y = g.yb[g.ny] - tiny;
    if (p.z < g.zb[0]) p.z = g.yb[0] + tiny;          // reference quirk Q12
    if (p.z > g.zb[g.nz]) p.z = g.zb[g.nz] - tiny;
    p.ix = entry_index(g.xb, g.nx, p.x);
    p.iy = entry_index(g.yb, g.ny, p.y);
    p.iz = entry_index(g.zb, g.nz, p.z);
    p.wt = (R)1.0;
}

// ------------------------------------------------------------------ history kernel
constexpr int HIST_STACK = 40;        // deepest LIFO depth below the two register slots

enum RngMode { RNG_STREAMS = 0, RNG_REPLAY = 1 };

template <class R> struct HistArgs {
    const Tables<R> *t;          // device copies (global memory), shared by all lanes
    const Geom<R> *g;
    SrcDosxyz<R> dos;
    SrcBeamlets<R> bix;
    int use_beamlets;            // 0: dosxyz source; 1: beamlet source
    const long *slot_off;        // beamlet source: [nslot+1] first history of each slot in this launch
    const int *slot_beamlet;     // beamlet of each slot
    int nslot;
    long slot_stride;            // tally stride between slots (doubles)
    int batch;                   // batch index (part of the beamlet random stream key)
    long nhist;                  // histories in this launch
    long hist_base;              // global index of history 0 (random stream id)
    uint64_t seed;
    const uint32_t *replay;      // RNG_REPLAY: 24-bit stream
    const long *replay_off;      // [nhist+1] first draw of each history
    const long *replay_end;      // optional [nhist] end of each history's draws (else replay_off[h+1])
    double *endep;               // [nreg] FP64 tally, region 0 outside
    uint8_t *touched;            // optional tile marks of the tallies (TallyAtomic), touch_stride per slot
    long touch_stride;
    double *ebal;                // optional [EB_N] energy balance (source, escape, lost)
    unsigned long long *counters;// [8] event counters (optional)
    int *err;                    // [4] 0 overflow, 1 replay underrun, 2 step guard
    unsigned long long *next;    // dynamic work counter (zeroed by the host)
    int chunk;
    long max_steps;              // transport calls per history after which it counts as stuck (step guard)
    int nsplit;                  // photon splitting factor; > 1 selects k_history_split
    Part<R> *stack;              // nsplit > HISTORY_SCRATCH_NSPLIT: per-lane particle stacks in device memory,
    int stack_depth;             // stack_depth entries for each lane of the launch (null: stacks in scratch)
};

// Random stream of history h_local of batch `batch` of a beamlet: the
// beamlet, batch and history index occupy disjoint bit fields, so every
// history of a Dij run has its own stream whatever the grouping of beamlets.
__device__ __forceinline__ uint64_t beamlet_stream(int beamlet, int batch, long h_local) {
    return ((uint64_t)(beamlet + 1) << 40) + ((uint64_t)batch << 32) + (uint64_t)h_local;
}

// Slot of history h: the largest s with off[s] <= h (off[0] = 0 <= h < off[n]).
__device__ __forceinline__ int find_slot(const long *off, int n, long h) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= h) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <class R, class G>
__device__ __forceinline__ void make_rng(G &g, const HistArgs<R> &a, long h, uint64_t stream);
template <> __device__ __forceinline__ void make_rng<float, Philox>(Philox &g, const HistArgs<float> &a, long,
                                                                   uint64_t stream) {
    g.init(a.seed, stream);
}
template <> __device__ __forceinline__ void make_rng<double, Philox>(Philox &g, const HistArgs<double> &a, long,
                                                                    uint64_t stream) {
    g.init(a.seed, stream);
}
template <> __device__ __forceinline__ void make_rng<float, Replay>(Replay &g, const HistArgs<float> &a, long h,
                                                                   uint64_t) {
    g.base = a.replay; g.pos = a.replay_off[h]; g.end = a.replay_end ? a.replay_end[h] : a.replay_off[h + 1];
    g.err = a.err + 1;
}
template <> __device__ __forceinline__ void make_rng<double, Replay>(Replay &g, const HistArgs<double> &a, long h,
                                                                    uint64_t) {
    g.base = a.replay; g.pos = a.replay_off[h]; g.end = a.replay_end ? a.replay_end[h] : a.replay_off[h + 1];
    g.err = a.err + 1;
}

// Plain history-by-history form: one lane runs whole showers, LIFO order,
// one random stream per history.  Lanes fetch chunks of histories from a
// global counter so long showers do not idle the rest of the grid.
// energy a primary brings into the balance (EB_SOURCE)
template <class R> __device__ __forceinline__ double source_energy(const Part<R> &p) {
    const double RMD = 0.5109989461;
    return (double)p.wt * (p.iq == 0 ? (double)p.e : (double)p.e + (p.iq > 0 ? RMD : -RMD));
}

template <class R, class G, bool COUNT>
__global__ void __launch_bounds__(256) k_history(HistArgs<R> a) {
    Part<R> deep[HIST_STACK];
    Counters tot = {0, 0, 0, 0, 0, 0, 0};
    double eb_src = 0.0, eb_esc = 0.0, eb_lost = 0.0;
    for (;;) {
        long h0 = 0;
        if (threadIdx.x % warpSize == 0) h0 = (long)atomicAdd(a.next, (unsigned long long)a.chunk * warpSize);
        h0 = __shfl(h0, 0) + (long)(threadIdx.x % warpSize) * a.chunk;
        if (__shfl(h0, 0) >= a.nhist) break;
        for (long h = h0; h < h0 + a.chunk && h < a.nhist; h++) {
            int slot = 0, beamlet = 0;
            uint64_t stream = (uint64_t)(a.hist_base + h);
            if (a.use_beamlets) {
                slot = find_slot(a.slot_off, a.nslot, h);
                beamlet = a.slot_beamlet[slot];
                stream = beamlet_stream(beamlet, a.batch, h - a.slot_off[slot]);
            }
            G rng;
            make_rng<R, G>(rng, a, h, stream);
            TallyAtomic tally;
            tally.init(a.endep + (size_t)slot * a.slot_stride,
                       a.touched ? a.touched + (size_t)slot * a.touch_stride : nullptr);
            Engine<R, G, TallyAtomic> e{*a.t, *a.g, rng, tally, {0, 0, 0, 0, 0, 0, 0}, COUNT};
            Part<R> p, q;
            if (a.use_beamlets) init_history(e, a.bix, beamlet, p);
            else init_history(e, a.dos, p);
            eb_src += source_energy(p);
            int depth = 0;
            long guard = 0;
            for (;;) {
                int d = e.step(p, q);
                if (e.aborted) { atomicOr(a.err + 2, 1); break; }
                if (d < 0) {
                    if (depth == 0) break;
                    p = deep[--depth];
                } else if (d > 0) {
                    if (depth >= HIST_STACK) { atomicOr(a.err, 1); break; }
                    deep[depth++] = p;
                    p = q;
                }
                if (++guard > a.max_steps) { atomicOr(a.err + 2, 1); break; }
            }
            tally.flush();
            eb_esc += e.esc;
            eb_lost += e.lost;
            if (COUNT) {
                tot.nhowfar += e.cnt.nhowfar; tot.nhownear += e.cnt.nhownear; tot.ncross += e.cnt.ncross;
                tot.ndep += e.cnt.ndep; tot.nrn += e.cnt.nrn; tot.nphoton += e.cnt.nphoton;
                tot.nelectron += e.cnt.nelectron;
            }
        }
    }
    if (a.ebal) {
        wave_add(a.ebal + EB_SOURCE, eb_src);
        wave_add(a.ebal + EB_ESCAPE, eb_esc);
        wave_add(a.ebal + EB_LOST, eb_lost);
    }
    if (COUNT && a.counters) {
        atomicAdd(a.counters + 0, tot.nhowfar); atomicAdd(a.counters + 1, tot.nhownear);
        atomicAdd(a.counters + 2, tot.ncross); atomicAdd(a.counters + 3, tot.ndep);
        atomicAdd(a.counters + 4, tot.nrn); atomicAdd(a.counters + 5, tot.nphoton);
        atomicAdd(a.counters + 6, tot.nelectron);
    }
}

// History form with photon splitting: the whole LIFO stack of a history
// lives in an array, indexed like the reference's particle stack, because
// the splitting photon routine pushes sub-photons and compacts the stack in
// place.  Up to nsplit HISTORY_SCRATCH_NSPLIT the array is in scratch
// (SPLIT_STACK entries); the reference's algorithm fills ~70 entries at
// nsplit 8, ~220 at 20 and ~790 at 64 (6 MV on a 38 cm slab), so larger
// factors (DEEP) use a stack of a.stack_depth entries per lane in device
// memory, which the host allocates for every lane of the launch.
constexpr int SPLIT_STACK = 96;
constexpr int HISTORY_SCRATCH_NSPLIT = 8;

// depth of the device-memory stacks for a splitting factor
inline int split_stack_depth(int nsplit) { return 32 * nsplit + 64; }

template <class R, class G, bool COUNT, bool DEEP>
__global__ void __launch_bounds__(256) k_history_split(HistArgs<R> a) {
    Part<R> local[DEEP ? 1 : SPLIT_STACK];
    Part<R> *st = DEEP ? a.stack + ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * (size_t)a.stack_depth : local;
    const int cap = DEEP ? a.stack_depth : SPLIT_STACK;
    Counters tot = {0, 0, 0, 0, 0, 0, 0};
    double eb_src = 0.0, eb_esc = 0.0, eb_lost = 0.0;
    for (;;) {
        long h0 = 0;
        if (threadIdx.x % warpSize == 0) h0 = (long)atomicAdd(a.next, (unsigned long long)a.chunk * warpSize);
        h0 = __shfl(h0, 0) + (long)(threadIdx.x % warpSize) * a.chunk;
        if (__shfl(h0, 0) >= a.nhist) break;
        for (long h = h0; h < h0 + a.chunk && h < a.nhist; h++) {
            int slot = 0, beamlet = 0;
            uint64_t stream = (uint64_t)(a.hist_base + h);
            if (a.use_beamlets) {
                slot = find_slot(a.slot_off, a.nslot, h);
                beamlet = a.slot_beamlet[slot];
                stream = beamlet_stream(beamlet, a.batch, h - a.slot_off[slot]);
            }
            G rng;
            make_rng<R, G>(rng, a, h, stream);
            TallyAtomic tally;
            tally.init(a.endep + (size_t)slot * a.slot_stride,
                       a.touched ? a.touched + (size_t)slot * a.touch_stride : nullptr);
            Engine<R, G, TallyAtomic> e{*a.t, *a.g, rng, tally, {0, 0, 0, 0, 0, 0, 0}, COUNT};
            e.nsplit = a.nsplit;
            if (a.use_beamlets) init_history(e, a.bix, beamlet, st[0]);
            else init_history(e, a.dos, st[0]);
            eb_src += source_energy(st[0]);
            int np = 0, err = 0;
            long guard = 0;
            while (np >= 0) {
                if (st[np].iq == 0) {
                    np = e.photon_split(st, np, cap, &err);
                } else {
                    Part<R> q;
                    int d = e.electron(st[np], q);
                    if (d < 0) {
                        np -= 1;
                    } else if (d > 0) {
                        if (np + 1 >= cap) { err = 1; break; }
                        st[++np] = q;
                    }
                }
                if (err) break;
                if (e.aborted) { atomicOr(a.err + 2, 1); break; }
                if (++guard > a.max_steps) { atomicOr(a.err + 2, 1); break; }
            }
            if (err) atomicOr(a.err, 1);
            tally.flush();
            eb_esc += e.esc;
            eb_lost += e.lost;
            if (COUNT) {
                tot.nhowfar += e.cnt.nhowfar; tot.nhownear += e.cnt.nhownear; tot.ncross += e.cnt.ncross;
                tot.ndep += e.cnt.ndep; tot.nrn += e.cnt.nrn; tot.nphoton += e.cnt.nphoton;
                tot.nelectron += e.cnt.nelectron;
            }
        }
    }
    if (a.ebal) {
        wave_add(a.ebal + EB_SOURCE, eb_src);
        wave_add(a.ebal + EB_ESCAPE, eb_esc);
        wave_add(a.ebal + EB_LOST, eb_lost);
    }
    if (COUNT && a.counters) {
        atomicAdd(a.counters + 0, tot.nhowfar); atomicAdd(a.counters + 1, tot.nhownear);
        atomicAdd(a.counters + 2, tot.ncross); atomicAdd(a.counters + 3, tot.ndep);
        atomicAdd(a.counters + 4, tot.nrn); atomicAdd(a.counters + 5, tot.nphoton);
        atomicAdd(a.counters + 6, tot.nelectron);
    }
}

// Batch statistics (omc_dosxyz.c:696-717; omc_matrad.c:897-918): after each
// batch, sum += e/scale, sum2 += (e/scale)^2, e = 0; the raw energies are
// also summed into *dep (the EB_DEPOSIT term of the energy balance) when
// dep is set.
__global__ void k_accumulate(double *e, double *sum, double *sum2, long n, double scale, double *dep);

// ------------------------------------------------------------------ sampler parity kernel
// One thread per case: the same inputs and the same injected random numbers
// as a golden case of the reference harness; outputs in the harness layout.
enum SampKind {
    S_COMPTON = 1, S_PAIR, S_PHOTO, S_RAYLEIGH, S_BREMS, S_MOLLER, S_BHABHA, S_ANNIH, S_RANNIH,
    S_MSDIST, S_MSCAT, S_SSCAT, S_DRANGE, S_ELOSS, S_PHOTON, S_ELECTRON
};

template <class R> struct SampArgs {
    const Tables<R> *t;          // device copies (global memory), shared by all lanes
    const Geom<R> *g;
    int ncase;
    const int *kind;
    const double *in;        // [ncase*in_stride]
    int in_stride;
    const uint32_t *rn;      // concatenated 24-bit draws
    const long *rn_off;      // [ncase+1]
    double *out;             // [ncase*out_stride]
    int out_stride;
    long *rn_used;           // [ncase]
    int *err;
};

template <class R> __device__ void load_part(const double *v, Part<R> &p, const Geom<R> &g) {
    p.iq = (int)v[0];
    int ir = (int)v[1];
    if (ir <= 0) { p.ix = -1; p.iy = 0; p.iz = 0; }
    else { int r = ir - 1; p.ix = r % g.nx; p.iy = (r / g.nx) % g.ny; p.iz = r / (g.nx * g.ny); }
    p.e = (R)v[2]; p.x = (R)v[3]; p.y = (R)v[4]; p.z = (R)v[5];
    p.u = (R)v[6]; p.v = (R)v[7]; p.w = (R)v[8]; p.wt = (R)v[9];
}

template <class R> __device__ void store_part(double *v, const Part<R> &p, const Geom<R> &g) {
    v[0] = p.iq;
    v[1] = p.ix < 0 ? 0 : 1 + p.ix + g.nx * (p.iy + g.ny * p.iz);
    v[2] = p.e; v[3] = p.x; v[4] = p.y; v[5] = p.z; v[6] = p.u; v[7] = p.v; v[8] = p.w; v[9] = p.wt;
    v[10] = 0.0;
}

// Logging tally used by the single photon()/electron() parity cases.
struct TallyRec {
    double *buf;        // 2 doubles per call: region, wt*edep
    int cap, n;
    __device__ __forceinline__ void add(int region, double v) {
        if (n < cap) { buf[2 * n] = region; buf[2 * n + 1] = v; }
        n++;
    }
    __device__ __forceinline__ void flush() {}
};

template <class R>
__global__ void __launch_bounds__(64) k_sampler(SampArgs<R> a) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncase) return;
    const double *in = a.in + (size_t)c * a.in_stride;
    double *out = a.out + (size_t)c * a.out_stride;
    Replay rng;
    rng.base = a.rn; rng.pos = a.rn_off[c]; rng.end = a.rn_off[c + 1]; rng.err = a.err;
    TallyRec tally;
    tally.buf = out + 64; tally.cap = (a.out_stride - 64) / 2; tally.n = 0;
    Engine<R, Replay, TallyRec> e{*a.t, *a.g, rng, tally, {0, 0, 0, 0, 0, 0, 0}, true};
    Part<R> p, q;
    int kind = a.kind[c];
    int d = 0;
    out[0] = 0;
    switch (kind) {
    case S_COMPTON: load_part(in, p, *a.g); e.compton(p, q); d = 1; break;
    case S_PAIR: load_part(in, p, *a.g); e.pair((int)in[11], p, q); d = 1; break;
    case S_PHOTO: load_part(in, p, *a.g); e.photo(p); d = 0; break;
    case S_RAYLEIGH: load_part(in, p, *a.g); e.rayleigh(p, (int)in[11], (R)in[2], (R)in[12], (int)in[13]); d = 0; break;
    case S_BREMS: load_part(in, p, *a.g); e.brems(p, q); d = 1; break;
    case S_MOLLER: load_part(in, p, *a.g); d = e.moller(p, q); break;
    case S_BHABHA: load_part(in, p, *a.g); e.bhabha(p, q); d = 1; break;
    case S_ANNIH: load_part(in, p, *a.g); e.annih(p, q); d = 1; break;
    case S_RANNIH: load_part(in, p, *a.g); e.rannih(p, q); d = 1; break;
    case S_PHOTON:
        // the reference's photon() keeps transporting the same photon after a
        // Rayleigh scatter within one call
        load_part(in, p, *a.g);
        do { d = e.photon(p, q); } while (d == 0 && p.iq == 0);
        break;
    case S_ELECTRON: load_part(in, p, *a.g); d = e.electron(p, q); break;
    case S_MSDIST: {
        load_part(in + 6, p, *a.g);
        Part<R> f;
        R us = e.msdist(p, (int)in[0], (int)in[1], (R)in[2], (R)in[3], (R)in[4], (R)in[5], f);
        out[1] = us; out[2] = f.x; out[3] = f.y; out[4] = f.z; out[5] = f.u; out[6] = f.v; out[7] = f.w;
        d = 2;
        break;
    }
    case S_MSCAT: {
        bool si = true, fi = true;
        typename Engine<R, Replay, TallyRec>::Mscats ms = {0, 0, (R)0};
        typename Engine<R, Replay, TallyRec>::Spinr sr = {0, 0};
        R cost = (R)7, sint = (R)7;
        e.mscat((int)in[0], (int)in[1], si, fi, (R)in[2], (R)in[3], (R)in[4], (R)in[5], (R)in[6], cost, sint, ms, sr);
        out[1] = cost; out[2] = sint;
        d = 2;
        break;
    }
    case S_SSCAT: {
        R cost, sint;
        e.sscat((int)in[0], (int)in[1], (R)in[2], (R)in[3], (R)in[4], cost, sint);
        out[1] = cost; out[2] = sint;
        d = 2;
        break;
    }
    case S_DRANGE:
        out[1] = e.drange((int)in[0], (int)in[1], (int)in[2], (R)in[3], (R)in[4], (R)in[5], (R)in[6]);
        d = 2;
        break;
    case S_ELOSS:
        out[1] = e.eloss((int)in[0], (int)in[1], (R)in[3], (R)in[4], (R)in[5], (R)in[6], (R)in[7], (int)in[8]);
        d = 2;
        break;
    default:
        d = 3;
    }
    // out[0]: stack depth after the call (-1, 0, 1) for particle cases, 2 for
    // scalar cases; particles at out[1..11] and out[12..22]
    out[0] = d;
    if (d <= 1) {
        if (d >= 0) store_part(out + 1, p, *a.g);
        if (d == 1) store_part(out + 12, q, *a.g);
    }
    out[23] = tally.n;
    out[24] = (double)e.cnt.nhowfar;
    out[25] = (double)e.cnt.nhownear;
    a.rn_used[c] = rng.pos - a.rn_off[c];
}

}  // namespace omc
