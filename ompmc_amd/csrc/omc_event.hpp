// Event-queue ("batched") transport form.
//
// Particles live in structure-of-arrays pools in HBM, one pool per type
// (photons, charged particles), double-buffered by round parity.  A round is
// four launches on one stream:
//   k_source     samples this round's injected primaries into the input pool,
//   k_event<1>   transports the queued photons (persistent lanes, below;
//                k_event_split when photons are split),
//   k_event<0>   transports the queued electrons/positrons, `budget`
//                condensed-history steps (electron_run) between refills,
//   k_round_end  retires the consumed input pools and plans the next
//                injection so that the number of particles in flight stays
//                near a target.
// Queue lengths live in device memory (EvCtl), so rounds are enqueued
// without host round trips; the host reads EvCtl now and then to see whether
// the job has drained.  Surviving particles and the secondaries they produced
// are appended to the output pools with one atomic per wavefront (ballot +
// popcount for the lane offsets).  Every particle carries its own random
// stream: primaries are keyed like the plain history form, secondaries by a
// hash of the parent's stream and draw position, so a result does not
// depend on scheduling, grid size or the target.
#pragma once
#include "omc_kernels.hpp"

namespace omc {

template <class R> struct Pool {
    R *e, *x, *y, *z, *u, *v, *w, *wt;
    R *demfp, *sig0;             // resumable electron state (charged pool)
    int *ix, *iy, *iz, *iq;
    int *phase;                  // electron_run phase (bits 0-7) and the particle's requeue count (bits 8-31)
    int *slot;                   // tally slot (beamlet of a Dij group); 0 for dose runs
    uint32_t *ctr;               // draws taken from the particle's stream
    uint64_t *stream;            // the stream's key
    uint4 *rs;                   // the stream's generator state, where ctr > 0
};

// A particle's random stream: xoshiro128** keyed by the stream's 64-bit key,
// with the count of draws taken from it (the position that keys the
// streams of the particle's secondaries).  A particle that has drawn travels
// through the queues with its generator state; one that has not (a new
// secondary) starts from its key.
struct StreamAt {
    Xoshiro g;
    uint32_t n;
    __device__ __forceinline__ void at(uint64_t seed, uint64_t stream, uint32_t pos, const uint4 *saved) {
        n = pos;
        if (pos == 0) {
            g.init(seed, stream);
        } else {
            const uint4 s = *saved;
            g.s0 = s.x; g.s1 = s.y; g.s2 = s.z; g.s3 = s.w;
        }
    }
    __device__ __forceinline__ uint4 state() const { return make_uint4(g.s0, g.s1, g.s2, g.s3); }
    template <class R> __device__ __forceinline__ R next() {
        n++;
        return g.template next<R>();
    }
};

__device__ __forceinline__ uint64_t child_stream(uint64_t s, uint32_t n) {
    // splitmix64 finaliser of (parent stream, draw position)
    uint64_t z = s ^ (0x9E3779B97F4A7C15ull * ((uint64_t)n + 1ull)) ^ 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Device-resident control block of one job (one batch of histories).
struct EvCtl {
    unsigned long long cnt[2][2];     // [round parity][photon, charged] queued particles (charged: two bands, see emit)
    long long nhist, next_hist;       // histories of the job / injected so far
    long long inj_h0, inj_n, inj_at;  // injection of the coming round
    long long target, cap;
    long long inj_max;                // primaries injected per round at most (0: no limit)
    long long rounds, done_round;     // rounds retired; round at which the job drained (-1 while running)
    unsigned long long events[2];     // photon / charged particles taken from the queues
    unsigned long long peak;          // largest number of particles in flight
    unsigned long long claim[2][2];   // [round parity][photon, charged] queue entries claimed by lanes
    int inj_dest;                     // pool of the primaries: 0 photons, 1 charged
    int err;                          // bit 0: pool overflow; bit 1: a particle exceeded an iteration guard
    // the first particle dropped by a guard: why (Engine::aborted, 5 lane
    // trips, 6 requeues), the source line of the guard, and its state
    int drop_why, drop_site, drop_iq, drop_ix, drop_iy, drop_iz, drop_phase, drop_age;
    float drop_e;
};

// Records the first particle a guard dropped (one lane of the whole job wins).
template <class R>
__device__ __forceinline__ void note_drop(EvCtl *c, int why, int site, const Part<R> &p, int phase, int age) {
    if (atomicCAS(&c->drop_why, 0, why) == 0) {
        c->drop_site = site;
        c->drop_iq = p.iq;
        c->drop_ix = p.ix;
        c->drop_iy = p.iy;
        c->drop_iz = p.iz;
        c->drop_phase = phase;
        c->drop_age = age;
        c->drop_e = (float)p.e;
    }
}

template <class R> struct EvArgs {
    const Tables<R> *t;
    const Geom<R> *g;
    uint64_t seed;
    Pool<R> in;                       // input pool of this kernel's type
    const unsigned long long *n_in;   // its length
    Pool<R> outp, oute;               // photon and charged output pools
    unsigned long long *n_out;        // [2] their lengths
    unsigned long long *claim;        // entries of the input queue claimed so far (zero at launch)
    long long cap;                    // capacity of each pool
    double *endep;                    // tally base
    long slot_stride;                 // tally stride per slot
    uint8_t *touched;                 // optional tile marks of the tallies (TallyAtomic)
    long touch_stride;                // tile marks per slot
    double *ebal;                     // optional [EB_N] energy balance (escape, lost)
    unsigned long long *counters;     // [8] event counters (optional)
    int *err;                         // EvCtl::err
    int budget;                       // electron ustep iterations between refills of a wave
    int turns;                        // electrons: refills a particle may hold its lane before it is requeued (0: no limit)
    int max_age;                      // electrons: requeues after which a particle counts as stuck (dropped, err bit 1)
    int delta;                        // photons: 1 delta tracking, 0 voxel-by-voxel flight
    int nsplit;                       // photon splitting factor (vrt.nsplit); > 1 runs k_event_split for photons
    EvCtl *ctl;                       // the job's control block (guard reports)
    const CBin<R> *ebin;              // charged kernel: the electron half of Tables::cbin
    int eb_words;                     // its size in 16-byte words when copied to LDS (0: read in place)
    int tlist;                        // charged kernel: deposit-list entries per lane in LDS (TallyList)
};

template <class R> __device__ __forceinline__ void load(const Pool<R> &P, long i, Part<R> &p) {
    p.e = P.e[i]; p.x = P.x[i]; p.y = P.y[i]; p.z = P.z[i];
    p.u = P.u[i]; p.v = P.v[i]; p.w = P.w[i]; p.wt = P.wt[i];
    p.ix = P.ix[i]; p.iy = P.iy[i]; p.iz = P.iz[i]; p.iq = P.iq[i];
}

template <class R>
__device__ __forceinline__ void store(const Pool<R> &P, long i, const Part<R> &p, R demfp, R sig0, int phase,
                                      uint64_t stream, uint32_t ctr, uint4 rs, int slot) {
    P.e[i] = p.e; P.x[i] = p.x; P.y[i] = p.y; P.z[i] = p.z;
    P.u[i] = p.u; P.v[i] = p.v; P.w[i] = p.w; P.wt[i] = p.wt;
    P.ix[i] = p.ix; P.iy[i] = p.iy; P.iz[i] = p.iz; P.iq[i] = p.iq;
    P.demfp[i] = demfp; P.sig0[i] = sig0; P.phase[i] = phase;
    P.stream[i] = stream; P.ctr[i] = ctr; P.slot[i] = slot;
    if (ctr != 0u) P.rs[i] = rs;                 // a new stream (ctr 0) starts from its key
}

// Reserves an output slot for every lane with `want`, one atomic per wave.
// Reached by all lanes of the wavefront together.
__device__ __forceinline__ long long wave_reserve(unsigned long long *counter, bool want) {
    unsigned long long mask = __ballot(want);
    if (mask == 0) return -1;
    int lane = threadIdx.x & 63;
    int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    unsigned long long below = mask & ((1ull << lane) - 1ull);
    return want ? (long long)(base + __popcll(below)) : -1;
}

// Total energy (MeV) above which a charged particle joins the high band of
// the charged queue.  The queue holds two bands in one pool: particles below
// fill it from index 0 up, those above from the last index down, and the
// charged kernel takes the low band first.  A wave's lanes then hold
// particles of similar energy, whose steps branch alike: on the c3 plan
// 22.8 M histories/s against 21.6 M with one unsorted queue (thresholds 1.2
// and 1.5 alike, 2.5 slightly slower).  Where a particle sits in the queue
// does not change its track.
constexpr float CHARGED_BAND_E = 1.5f;

// Appends a particle to the photon or the charged output pool.  A photon of
// zero weight (one that lost the roulette of photon splitting) would deposit
// nothing and is not queued.  The charged pool's counter packs the low band's
// length (bits 0-31) and the high band's (bits 32-63); one atomic per wave
// reserves both, and a reservation that would make the bands meet writes
// nothing and reports an overflow (the job is rerun with larger pools).
template <class R>
__device__ __forceinline__ void emit(const EvArgs<R> &a, bool want, const Part<R> &p, R demfp, R sig0, int phase,
                                     uint64_t stream, uint32_t ctr, uint4 rs, int slot) {
    bool ph = want && p.iq == 0 && p.wt != (R)0, ch = want && p.iq != 0;
    long long ip = wave_reserve(a.n_out + 0, ph);
    if (ph) {
        if (ip < a.cap) store(a.outp, ip, p, demfp, sig0, phase, stream, ctr, rs, slot);
        else atomicOr(a.err, 1);
    }
    const bool hi = p.e > (R)CHARGED_BAND_E;
    const unsigned long long mlo = __ballot(ch && !hi), mhi = __ballot(ch && hi);
    if (mlo | mhi) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)(mlo | mhi)) - 1;
        const unsigned long long nlo = (unsigned long long)__popcll(mlo), nhi = (unsigned long long)__popcll(mhi);
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.n_out + 1, nlo | (nhi << 32));
        base = __shfl(base, leader);
        const unsigned long long lo0 = base & 0xffffffffull, hi0 = base >> 32;
        const unsigned long long below = (1ull << lane) - 1ull;
        if (ch) {
            if ((long long)(lo0 + nlo + hi0 + nhi) <= a.cap) {
                const long long ie = hi ? a.cap - 1 - (long long)(hi0 + __popcll(mhi & below))
                                        : (long long)(lo0 + __popcll(mlo & below));
                store(a.oute, ie, p, demfp, sig0, phase, stream, ctr, rs, slot);
            } else {
                atomicOr(a.err, 1);
            }
        }
    }
}

// Charged kernel's tally.  Like TallyAtomic it sums a lane's consecutive
// deposits to one region in a register, but a finished sum goes to a list of
// (slot, region, energy) entries the lane holds in LDS, and the wave adds the
// lists to device memory with FP64 atomics once per pass of its loop, where
// it meets to refill.  On gfx950 a load's wait for its data also waits for
// every earlier store and atomic of the wave, so a tally atomic inside the
// step loop stalled the following steps' loads.  A full list falls back to a
// direct atomic.  The sums are the same; only their order changes.  Measured
// on the 1960-beamlet plan in FP32, two runs each on one box: 22.55 and 22.82
// against 22.51 and 22.40 M histories/s (the 10,044-beamlet plan unchanged).
// The atomics still wait where the wave meets, so the gain stays well below
// the 4.4% a copy that issued no tally atomics in the step loop ran faster.
struct TallyList {
    double *endep;       // the particle's slot (direct atomics of a full list)
    uint8_t *touched;
    int slot, cur, k, cap;
    double acc;
    int2 *key;           // this lane's entries, strided by the workgroup size
    double *val;
    int stride;
    __device__ __forceinline__ void init(double *p, uint8_t *tf, int s) {
        endep = p;
        touched = tf;
        slot = s;
        cur = -1;
        acc = 0.0;
    }
    __device__ __forceinline__ void push(int r, double v) {
        if (r < 0 || v == 0.0) return;
        if (k < cap) {
            key[k * stride] = make_int2(slot, r);
            val[k * stride] = v;
            k++;
        } else {
            atomicAdd(endep + r, v);
            if (touched) touched[r >> TALLY_TILE_LOG2] = 1;
        }
    }
    __device__ __forceinline__ void add(int region, double v) {
        if (region != cur) {
            push(cur, acc);
            cur = region;
            acc = v;
        } else {
            acc += v;
        }
    }
    __device__ __forceinline__ void flush() {
        push(cur, acc);
        cur = -1;
        acc = 0.0;
    }
    // adds the lane's list to the tallies and empties it
    __device__ __forceinline__ void drain(double *base, long slot_stride, uint8_t *tbase, long touch_stride) {
        for (int j = 0; j < k; j++) {
            const int2 kk = key[j * stride];
            atomicAdd(base + (size_t)kk.x * slot_stride + kk.y, val[j * stride]);
            if (tbase) tbase[(size_t)kk.x * touch_stride + (kk.y >> TALLY_TILE_LOG2)] = 1;
        }
        k = 0;
    }
};

template <bool COUNT>
__device__ __forceinline__ void flush_counters(unsigned long long *dst, const Counters &c) {
    if (!COUNT || !dst) return;
    unsigned long long v[7] = {c.nhowfar, c.nhownear, c.ncross, c.ndep, c.nrn, c.nphoton, c.nelectron};
#pragma unroll
    for (int k = 0; k < 7; k++) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(dst + k, x);
    }
}

// The photon kernel takes the registers it needs (149 VGPRs, 3 waves per
// SIMD).  Capped at 4 waves (128 VGPRs, 12 bytes of spills) its busy time on
// the c3 plan fell by a fifth with the plan's rate unchanged, but single
// dose runs, where it has a stream to itself, ran 1% slower (c2 16.8 against
// 17.0 M histories/s).
// Occupancy hint of the charged-particle kernel (waves per SIMD).  With the
// whole engine inlined the kernel keeps its state in registers; capping it at
// 4 or 3 waves spills that state to scratch, and on gfx950 2 waves measured
// fastest (C3 subset: 6.0 M histories/s at 2, 5.9 M at 3, 4.9 M at 4).
#ifndef OMC_CHARGED_WAVES_PER_EU
#define OMC_CHARGED_WAVES_PER_EU 2
#endif

// Persistent transport of one particle type.  Every lane holds one particle
// in registers and advances it -- one photon() call, or up to a.budget
// condensed-history steps of electron_run() -- then the wave meets: lanes
// whose particle left, or turned into the other type, hand it and any
// secondary to the output queues and claim the next queued particle (one
// atomic per wave).  A particle of this kernel's type stays in its lane until
// it is gone, so the queues only carry new secondaries and type changes, and
// no lane idles while the queue still holds work.  The wave exits when its
// lanes are empty and the queue is exhausted.
//
// Two things keep a launch short.  A wave reads the claim counter before it
// bumps it, so once the queue is exhausted idle waves leave without
// serialising on that one address (8192 waves queueing on it cost ~100 us
// per launch).  And an electron that has held its lane for a.turns refills
// is handed back to the queue with its resumable state (position in its
// random stream, distance to the next discrete interaction, step phase): an
// electron in single-scattering mode along a voxel face can take thousands
// of steps, and without the hand-back it alone would hold the round -- and
// every particle queued behind the round -- open for tens of milliseconds.
// Where a particle is requeued does not change its track.  The requeue count
// travels with the particle (high bits of its phase word): a particle that
// would make no progress is dropped and reported after a.max_age requeues,
// instead of cycling through rounds until the job's round guard ends it.
template <class R, bool PHOTON, bool COUNT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PHOTON ? 1 : OMC_CHARGED_WAVES_PER_EU)))
k_event(EvArgs<R> a) {
    // the charged queue: low band [0, n_lo), then the high band from cap - 1 down
    const unsigned long long pk = *a.n_in;
    // Charged kernel: the electron bins of every medium (nmed x (MEKE + 1)
    // records of 80 bytes, at most 64 KiB; launched with that much dynamic
    // LDS) are copied to LDS once per workgroup, and the step loop reads its
    // per-bin coefficients there instead of from device memory.
    extern __shared__ __align__(16) unsigned char omc_lds[];
    const CBin<R> *eb = a.ebin;
    if (!PHOTON && a.eb_words > 0) {
        uint4 *dst = reinterpret_cast<uint4 *>(omc_lds);
        const uint4 *src = reinterpret_cast<const uint4 *>(a.ebin);
        for (int i = threadIdx.x; i < a.eb_words; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
        eb = reinterpret_cast<const CBin<R> *>(omc_lds);
    }
    // FP32 charged kernel: each lane's deposit list follows the bins in LDS,
    // keys [tlist][blockDim.x] then energies [tlist][blockDim.x].  (The FP64
    // kernel, which spills registers, ran 1% slower with the lists.)
    constexpr bool LIST = !PHOTON && sizeof(R) == 4;
    using Tally = typename std::conditional<LIST, TallyList, TallyAtomic>::type;
    Tally tally;
    if constexpr (LIST) {
        unsigned char *lst = omc_lds + (size_t)a.eb_words * 16u;
        tally.key = reinterpret_cast<int2 *>(lst) + threadIdx.x;
        tally.val = reinterpret_cast<double *>(lst + (size_t)a.tlist * blockDim.x * sizeof(int2)) + threadIdx.x;
        tally.stride = (int)blockDim.x;
        tally.cap = a.tlist;
        tally.k = 0;
        tally.init(a.endep, a.touched, 0);
    } else {
        tally.init(a.endep);
    }
    const long long n_lo = PHOTON ? (long long)pk : (long long)(pk & 0xffffffffull);
    const long long n = PHOTON ? (long long)pk : n_lo + (long long)(pk >> 32);
    Counters tot = {0, 0, 0, 0, 0, 0, 0};
    double eb_esc = 0.0;             // photons that left the phantom (photon kernel only)
    bool have = false;
    Part<R> p, q;
    uint64_t stream = 0;
    int slot = 0, phase = 0;
    R demfp = (R)0, sig0 = (R)0;
    StreamAt rng;
    int iters = 0, turns = 0, age = 0;
    for (;;) {
        if constexpr (LIST) tally.drain(a.endep, a.slot_stride, a.touched, a.touch_stride);
        long long idx = -1;
        if (__any(!have)) {
            const unsigned long long seen = __shfl(*(volatile const unsigned long long *)a.claim, 0);
            if ((long long)seen < n) idx = wave_reserve(a.claim, !have);
        }
        if (!have && idx >= 0 && idx < n) {
            iters = 0;
            turns = 0;
            if (!PHOTON && idx >= n_lo) idx = a.cap - 1 - (idx - n_lo);
            load(a.in, idx, p);
            stream = a.in.stream[idx];
            slot = a.in.slot[idx];
            demfp = a.in.demfp[idx];
            sig0 = a.in.sig0[idx];
            phase = a.in.phase[idx];
            age = phase >> 8;
            phase &= 0xff;
            rng.at(a.seed, stream, a.in.ctr[idx], a.in.rs + idx);
            if constexpr (LIST)
                tally.init(a.endep + (size_t)slot * a.slot_stride,
                           a.touched ? a.touched + (size_t)slot * a.touch_stride : nullptr, slot);
            else
                tally.init(a.endep + (size_t)slot * a.slot_stride,
                           a.touched ? a.touched + (size_t)slot * a.touch_stride : nullptr);
            have = true;
        }
        if (!__any(have)) break;
        int d = -1;
        bool done = false;
        uint32_t qpos = 0;
        if (have) {
            Engine<R, StreamAt, Tally, !PHOTON && sizeof(R) == 4> e{*a.t, *a.g, rng, tally,
                                                                          {0, 0, 0, 0, 0, 0, 0}, COUNT};
            e.nsplit = a.nsplit;          // roulette of the photons that charged particles produce
            e.eb = eb;
            if (PHOTON) {
                d = a.delta ? e.photon_delta(p, q) : e.photon(p, q);
            } else {
                typename decltype(e)::EState es;
                es.demfp = demfp;
                es.sig0 = sig0;
                es.phase = phase;
                d = e.electron_run(p, q, es, a.budget);
                demfp = es.demfp;
                sig0 = es.sig0;
                phase = es.phase;
            }
            qpos = rng.n;
            // a particle that made no progress within its guards is dropped
            // and reported (EvCtl::err bit 1), never transported forever
            if (e.aborted || ++iters > (1 << 22)) {
                atomicOr(a.err, 2);
                note_drop(a.ctl, e.aborted ? e.aborted : 5, e.abort_site, p, phase, age);
                d = -1;
            }
            // d: -1 particle gone, 0 particle continues, +1 continues with a secondary
            done = d < 0 || ((p.iq == 0) != PHOTON);
            if (!PHOTON && !done && a.turns > 0 && ++turns >= a.turns) {            // requeue, see above
                done = true;
                if (++age > a.max_age) {
                    atomicOr(a.err, 2);
                    note_drop(a.ctl, 6, 0, p, phase, age);
                    d = -1;
                }
            }
            if (done) tally.flush();
            if (PHOTON) eb_esc += e.esc;
            // dropped energy is rare (quirk Q6, guards): added where it occurs
            if (e.lost != 0.0 && a.ebal) atomicAdd(a.ebal + EB_LOST, e.lost);
            if (COUNT) {
                tot.nhowfar += e.cnt.nhowfar; tot.nhownear += e.cnt.nhownear; tot.ncross += e.cnt.ncross;
                tot.ndep += e.cnt.ndep; tot.nrn += e.cnt.nrn; tot.nphoton += e.cnt.nphoton;
                tot.nelectron += e.cnt.nelectron;
            }
        }
        emit(a, have && done && d >= 0, p, demfp, sig0, phase | (age << 8), stream, qpos, rng.state(), slot);
        emit(a, have && d == 1, q, (R)0, (R)0, 0, child_stream(stream, qpos), 0u, uint4{}, slot);
        if (done) have = false;
    }
    if (PHOTON && a.ebal) wave_add(a.ebal + EB_ESCAPE, eb_esc);
    flush_counters<COUNT>(a.counters, tot);
}

// Persistent transport of photons with splitting (src/ompmc.c:1867-2139,
// nsplit > 1).  A lane holds one photon and splits it: nsplit sub-photons of
// weight wt/nsplit fly stratified numbers of mean free paths along its ray,
// each from the interaction site of the one before; a sub-photon that leaves
// the geometry ends the split.  Charged products keep the sub-photon's
// weight and are queued.  Of the scattered photons only those of one
// randomly chosen sub-photon survive, with the full weight; there is at most
// one, and it goes back to the photon queue, with its place in its random
// stream, to be split in a later round.  (Kept in the lane, a photon that
// scatters many times held the whole launch open; requeued, it follows the
// same track.)  The reference keeps the same particles on its stack, in
// another order.
//
// The loop over sub-photons has the same trip count on every lane, so a
// wave queues the products of each sub-photon together (one atomic per
// wave).  Each product gets a random stream of its own, keyed by the
// photon's stream and its draw position after the interaction; the second
// product of a pair by the complement of the stream.
template <class R, bool COUNT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1)))
k_event_split(EvArgs<R> a) {
    const long long n = (long long)*a.n_in;
    const int ns = a.nsplit;
    Counters tot = {0, 0, 0, 0, 0, 0, 0};
    double eb_esc = 0.0, eb_lost = 0.0;
    for (;;) {
        // every lane is free here: the wave claims up to 64 photons
        long long idx = -1;
        const unsigned long long seen = __shfl(*(volatile const unsigned long long *)a.claim, 0);
        if ((long long)seen < n) idx = wave_reserve(a.claim, true);
        const bool have = idx >= 0 && idx < n;
        if (!__any(have)) break;
        Part<R> p;
        uint64_t stream = 0;
        int slot = 0;
        StreamAt rng;
        TallyAtomic tally;
        tally.init(a.endep);
        if (have) {
            load(a.in, idx, p);
            stream = a.in.stream[idx];
            slot = a.in.slot[idx];
            rng.at(a.seed, stream, a.in.ctr[idx], a.in.rs + idx);
            tally.init(a.endep + (size_t)slot * a.slot_stride,
                       a.touched ? a.touched + (size_t)slot * a.touch_stride : nullptr);
        }
        Engine<R, StreamAt, TallyAtomic> e{*a.t, *a.g, rng, tally, {0, 0, 0, 0, 0, 0, 0}, COUNT};
        e.nsplit = ns;
        // set-up of the split (src/ompmc.c:1900-1940)
        bool live = have, surv = false;
        R eig = (R)0, gle = (R)0, eta = (R)0, d_eta = (R)0, dpmfp_old = (R)0, wtsave = (R)0;
        int i_survive = -1, imed = -1, lgle = 0;
        R cohfac = (R)0, gmfp = (R)0;
        Part<R> save = p;
        typename Engine<R, StreamAt, TallyAtomic>::SplitXs xs;
        if (live) {
            if (COUNT) e.cnt.nphoton++;
            imed = e.medium(p);
            eig = p.e;
            if (eig <= e.pcut_of(imed) || p.wt == (R)0) {
                e.ausgab(p, eig);
                live = false;
            } else {
                gle = M<R>::log(eig);
                e.split_xs(gle, xs);
                R rnno = e.rnd();
                rnno /= (R)ns;
                d_eta = (R)1.0 / (R)ns;
                eta = (R)1.0 - rnno + d_eta;
                wtsave = p.wt / (R)ns;
                rnno = e.rnd();
                i_survive = (int)(rnno * (R)ns);
            }
        }
        // the sub-photons (src/ompmc.c:1942-2111)
        for (int isplit = 0; isplit < ns; isplit++) {
            if (!__any(live)) break;
            Part<R> c0, c1;
            bool w0 = false, w1 = false;
            uint32_t cpos = 0;
            if (live) {
                eta -= d_eta;
                if (eta <= (R)0.0) {
                    live = false;
                } else {
                    R dpmfp = -M<R>::log(eta) - dpmfp_old;
                    dpmfp_old += dpmfp;
                    Part<R> sp = save;
                    sp.wt = wtsave;
                    sp.iq = 0;
                    if (!e.split_fly(sp, dpmfp, xs, imed, lgle, cohfac, gmfp, ns - isplit)) {
                        live = false;
                    } else {
                        save.x = sp.x; save.y = sp.y; save.z = sp.z;
                        save.ix = sp.ix; save.iy = sp.iy; save.iz = sp.iz;
                        const bool mine = isplit == i_survive;
                        R r = e.rnd();
                        if (r <= (R)1.0 - cohfac) {
                            if (mine) {
                                sp.wt *= (R)ns;
                                e.rayleigh(sp, imed, eig, gle, lgle);
                                p = sp;
                                surv = true;
                            }
                        } else {
                            Part<R> q;
                            int k = e.interact_pcp(sp, q, imed, eig, gle, lgle);
                            if (sp.iq != 0) {
                                c0 = sp;
                                w0 = true;
                            } else if (mine) {
                                sp.wt *= (R)ns;
                                p = sp;
                                surv = true;
                            }
                            if (k == 1) {
                                if (q.iq != 0) {
                                    c1 = q;
                                    w1 = true;
                                } else if (mine) {
                                    q.wt *= (R)ns;
                                    p = q;
                                    surv = true;
                                }
                            }
                        }
                        cpos = rng.n;
                    }
                }
            }
            emit(a, w0, c0, (R)0, (R)0, 0, child_stream(stream, cpos), 0u, uint4{}, slot);
            emit(a, w1, c1, (R)0, (R)0, 0, child_stream(~stream, cpos), 0u, uint4{}, slot);
        }
        if (have && e.aborted) {
            atomicOr(a.err, 2);
            note_drop(a.ctl, e.aborted, e.abort_site, p, 0, 0);
            if (surv) e.lost += (double)p.wt * (double)p.e;
            surv = false;
        }
        tally.flush();
        eb_esc += e.esc;
        eb_lost += e.lost;
        // the survivor continues from its place in the photon's stream
        emit(a, have && surv, p, (R)0, (R)0, 0, stream, rng.n, rng.state(), slot);
        if (COUNT && have) {
            tot.nhowfar += e.cnt.nhowfar; tot.nhownear += e.cnt.nhownear; tot.ncross += e.cnt.ncross;
            tot.ndep += e.cnt.ndep; tot.nrn += e.cnt.nrn; tot.nphoton += e.cnt.nphoton;
            tot.nelectron += e.cnt.nelectron;
        }
    }
    if (a.ebal) {
        wave_add(a.ebal + EB_ESCAPE, eb_esc);
        wave_add(a.ebal + EB_LOST, eb_lost);
    }
    flush_counters<COUNT>(a.counters, tot);
}

// Primaries of the coming round: histories [inj_h0, inj_h0 + inj_n) of the
// job, written to the input pool of their type from offset inj_at.
template <class R> struct SrcArgs {
    const Tables<R> *t;
    const Geom<R> *g;
    SrcDosxyz<R> dos;
    SrcBeamlets<R> bix;
    int use_beamlets;
    const long *slot_off;        // [nslot+1] first history of each tally slot in the job (nslot 0: one slot)
    const int *slot_beamlet;     // beamlet source: beamlet of each slot
    const int *slot_batch;       // beamlet source: batch of each slot (null: every slot is of `batch`)
    int nslot;
    int batch;
    long hist_base;              // dosxyz source: stream of history 0 of the job
    uint64_t seed;
    const EvCtl *ctl;
    Pool<R> out;                 // destination pool (photons or charged, by source charge)
    double *ebal;                // optional [EB_N] energy balance: the primaries' energy
};

template <class R>
__global__ void __launch_bounds__(256) k_source(SrcArgs<R> a) {
    const long n = (long)a.ctl->inj_n, h0 = (long)a.ctl->inj_h0, at = (long)a.ctl->inj_at;
    double eb_src = 0.0;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        long h = h0 + k;
        int slot = 0, beamlet = 0;
        uint64_t stream = (uint64_t)(a.hist_base + h);
        if (a.nslot > 0) slot = find_slot(a.slot_off, a.nslot, h);
        if (a.use_beamlets) {
            beamlet = a.slot_beamlet[slot];
            stream = beamlet_stream(beamlet, a.slot_batch ? a.slot_batch[slot] : a.batch, h - a.slot_off[slot]);
        }
        StreamAt rng;
        rng.at(a.seed, stream, 0, nullptr);
        TallyNone tn;
        Engine<R, StreamAt, TallyNone> e{*a.t, *a.g, rng, tn, {0, 0, 0, 0, 0, 0, 0}, false};
        Part<R> p;
        if (a.use_beamlets) init_history(e, a.bix, beamlet, p);
        else init_history(e, a.dos, p);
        store(a.out, at + k, p, (R)0, (R)0, 0, stream, rng.n, rng.state(), slot);
        eb_src += source_energy(p);
    }
    if (a.ebal) wave_add(a.ebal + EB_SOURCE, eb_src);
}

// Retires the round that read the pools of parity `par` and plans the
// injection of the next round (pools of parity 1 - par).  One thread.
__device__ __forceinline__ void round_end(EvCtl *c, int par) {
    c->events[0] += c->cnt[par][0];
    c->events[1] += (c->cnt[par][1] & 0xffffffffull) + (c->cnt[par][1] >> 32);
    c->cnt[par][0] = 0;
    c->cnt[par][1] = 0;
    c->claim[par][0] = 0;
    c->claim[par][1] = 0;
    int nx = 1 - par;
    if ((long long)c->cnt[nx][0] > c->cap) {
        c->err |= 1;
        c->cnt[nx][0] = (unsigned long long)c->cap;
    }
    {
        // an overflowing charged queue holds unwritten entries in both bands:
        // it is dropped (the job is rerun with larger pools)
        const unsigned long long pkc = c->cnt[nx][1];
        if ((long long)((pkc & 0xffffffffull) + (pkc >> 32)) > c->cap) {
            c->err |= 1;
            c->cnt[nx][1] = 0;
        }
    }
    const long long nch = (long long)((c->cnt[nx][1] & 0xffffffffull) + (c->cnt[nx][1] >> 32));
    long long inflight = (long long)c->cnt[nx][0] + nch;
    long long room = c->target - inflight;
    long long left = c->nhist - c->next_hist;
    long long space = c->cap - (c->inj_dest ? nch : (long long)c->cnt[nx][0]);
    long long n = left < room ? left : room;
    if (n > space) n = space;
    if (c->inj_max > 0 && n > c->inj_max) n = c->inj_max;
    if (n < 0) n = 0;
    c->inj_h0 = c->next_hist;
    c->inj_n = n;
    c->inj_at = c->inj_dest ? (long long)(c->cnt[nx][1] & 0xffffffffull) : (long long)c->cnt[nx][0];
    c->cnt[nx][c->inj_dest] += (unsigned long long)n;
    c->next_hist += n;
    if ((unsigned long long)(inflight + n) > c->peak) c->peak = (unsigned long long)(inflight + n);
    c->rounds++;
    if (inflight + n == 0 && c->done_round < 0) c->done_round = c->rounds;
}

static __global__ void k_round_end(EvCtl *c, int par) {
    if (blockIdx.x == 0 && threadIdx.x == 0) round_end(c, par);
}

// Starts a job: clears the queues and plans the first injection into the
// pools of parity 0.
static __global__ void k_job_start(EvCtl *c, long long nhist, long long target, long long cap, long long inj_max,
                                   int dest) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    c->cnt[0][0] = c->cnt[0][1] = c->cnt[1][0] = c->cnt[1][1] = 0;
    c->claim[0][0] = c->claim[0][1] = c->claim[1][0] = c->claim[1][1] = 0;
    c->nhist = nhist;
    c->next_hist = 0;
    c->target = target;
    c->cap = cap;
    c->inj_max = inj_max;
    c->rounds = -1;
    c->done_round = -1;
    c->events[0] = c->events[1] = 0;
    c->peak = 0;
    c->inj_dest = dest;
    c->err = 0;
    c->drop_why = c->drop_site = 0;
    round_end(c, 1);
}

}  // namespace omc
