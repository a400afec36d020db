// Device-side transport physics for gfx950, shared by every kernel form.
//
// Everything is templated on the transport precision R (float: production,
// double: the FP64 GPU oracle run) and on the random-number source G, so the
// plain history kernel, the event-queue kernels and the sampler-parity kernels
// all execute the same physics.  The control flow of each function follows
// the reference routine it names (src/ompmc.c, ucodes/omc_dosxyz/omc_dosxyz.c,
// ucodes/omc_matrad/omc_matrad.c) so that, fed the same random numbers, the
// FP64 instantiation consumes them in the same order and reproduces the
// reference's histories.
//
// Differences in representation (not in results):
//  * the voxel region is carried as (ix, iy, iz) instead of a linear region
//    number, so howfar/hownear need no integer division; ix < 0 means
//    "outside the phantom" (the reference's region 0);
//  * the perpendicular boundary distance `dnear` is written but never read by
//    the reference, so it is not carried;
//  * photon() returns to the caller after each interaction instead of looping
//    over the next photon on the stack (equivalent for nsplit = 1, the only
//    splitting factor the GPU paths accept);
//  * the particle being transported and the one secondary an interaction
//    creates live in registers: every routine works on (p, q) and returns the
//    change of stack depth, -1 (p finished), 0 (p continues) or +1 (q is the
//    new top, p lies below it), which is exactly the reference's LIFO order.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace omc {

constexpr int MXGE = 2000;
constexpr int MXEKE = 500;
constexpr int MXRAYFF = 100;
constexpr int MXE_SPIN = 15;
constexpr int MXE_SPIN1 = 2 * MXE_SPIN + 1;
constexpr int MXQ_SPIN = 15;
constexpr int MXU_SPIN = 31;
constexpr int NSPIN_MED = 2 * (MXE_SPIN1 + 1) * (MXQ_SPIN + 1) * (MXU_SPIN + 1);
constexpr int MXL_MS = 63;
constexpr int MXQ_MS = 7;
constexpr int MXU_MS = 31;
constexpr int MXMED = 9;
constexpr uint8_t VACUUM = 255;

// Engine methods are inlined into the kernels: an out-of-line call takes the
// particles, the random-number state and the tally by reference, which keeps
// them in scratch memory (flat loads and stores) for the whole kernel.
#ifndef OMC_INL
#define OMC_INL __forceinline__
#endif

// ------------------------------------------------------------------ math
template <class R> struct M;
template <> struct M<double> {
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double sin(double x) { return ::sin(x); }
    static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
    static __device__ __forceinline__ double fmax(double a, double b) { return ::fmax(a, b); }
    static __device__ __forceinline__ double fmin(double a, double b) { return ::fmin(a, b); }
    static __device__ __forceinline__ double fabs(double a) { return ::fabs(a); }
    // x^3 and x^4 through pow(): repeated products round differently in a
    // large share of cases, and the FP64 path must round as the reference does
    static __device__ __forceinline__ double pow3(double x) { return ::pow(x, 3.0); }
    static __device__ __forceinline__ double pow4(double x) { return ::pow(x, 4.0); }
    // remaining path (cm at the medium's nominal density) below which an
    // electron step counts as having reached its interaction site: the
    // reference's 1e-9 (src/ompmc.c:5355)
    static __device__ __forceinline__ double reach_tol(double) { return 1.0E-9; }
};
template <> struct M<float> {
    static __device__ __forceinline__ float log(float x) { return ::logf(x); }
    static __device__ __forceinline__ float exp(float x) { return ::expf(x); }
    static __device__ __forceinline__ float sqrt(float x) { return ::sqrtf(x); }
    static __device__ __forceinline__ float sin(float x) { return ::sinf(x); }
    static __device__ __forceinline__ float cos(float x) { return ::cosf(x); }
    static __device__ __forceinline__ float fmax(float a, float b) { return ::fmaxf(a, b); }
    static __device__ __forceinline__ float fmin(float a, float b) { return ::fminf(a, b); }
    static __device__ __forceinline__ float fabs(float a) { return ::fabsf(a); }
    static __device__ __forceinline__ float pow3(float x) { return x * x * x; }
    static __device__ __forceinline__ float pow4(float x) { float y = x * x; return y * y; }
    // In float the path to the interaction, t, goes through t/rhof*rhof on a
    // step that reaches the site and keeps a residual of up to two roundings
    // (about 1.2e-7 t), far above the reference's absolute 1e-9 whenever the
    // voxel density differs from the nominal one; below 4 times that residual
    // the site counts as reached, as it does exactly in FP64.
    static __device__ __forceinline__ float reach_tol(float t) { return ::fmaxf(1.0E-9f, 5.0E-7f * t); }
};

// ------------------------------------------------------------------ tables
// Piecewise-linear function of a log energy: value = lv * c1 + c0, one
// interleaved pair per (medium, interval) so one 8/16-byte load serves a lookup.
template <class R> struct PW { R c1, c0; };

// One bin of a row of the q^(2+) surface (msnew tables): alias probability,
// alias bin within the row, and the bin's (a, u) pair, interleaved so that a
// sampling trial reads one 16-byte entry (a second one when it takes the
// alias) instead of a chain of loads from four arrays.
template <class R> struct alignas(16) MsBin {
    R w;
    int i;
    R f, u;
};
static_assert(sizeof(MsBin<float>) == 16 && sizeof(MsBin<double>) == 32 && offsetof(MsBin<double>, f) == 16,
              "the bin loads below assume this layout");

// Loads through pointers into device tables and phantom arrays, which live
// in global memory: addressed as such they compile to global loads instead
// of flat accesses.
template <class T> __device__ __forceinline__ T gld(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
typedef float omc_f4 __attribute__((ext_vector_type(4)));
typedef double omc_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ MsBin<float> gld(const MsBin<float> *p) {
    const omc_f4 v = *(const __attribute__((address_space(1))) omc_f4 *)p;      // one 16-byte load
    return MsBin<float>{v.x, __float_as_int(v.y), v.z, v.w};
}
__device__ __forceinline__ MsBin<double> gld(const MsBin<double> *p) {
    const __attribute__((address_space(1))) omc_d2 *q = (const __attribute__((address_space(1))) omc_d2 *)p;
    const omc_d2 a = q[0], b = q[1];                                             // two 16-byte loads
    return MsBin<double>{a.x, (int)(__double_as_longlong(a.y) & 0xffffffffLL), b.x, b.y};
}
template <class R> __device__ __forceinline__ PW<R> gld(const PW<R> *p) {
    const __attribute__((address_space(1))) R *q = (const __attribute__((address_space(1))) R *)p;
    PW<R> r;
    r.c1 = q[0];
    r.c0 = q[1];
    return r;
}

// The coefficients of one energy bin of one charged species in one medium,
// interleaved (80 bytes in FP32): what a condensed-history step, its
// multiple scattering and its discrete interaction read at the particle's
// own bin, so a step touches one or two cache lines instead of eight tables.
// Tables::cbin holds them as [2][nmed][cb_stride] (electrons, positrons;
// cb_stride = the largest MEKE + 1); the per-quantity arrays stay for reads
// at other bins (the range walk of eloss, the end of a step's interval).
template <class R> struct CBin {
    PW<R> dedx, sig, tmxs, eta, blcc, q1c, q2c, br1, br2;  // br1/br2: ebr1 (electrons), pbr1/pbr2 (positrons)
    R e_lo, range;                                          // e_array and range_ep at the bin
};
static_assert(sizeof(CBin<float>) % 16 == 0 && sizeof(CBin<double>) % 16 == 0, "copied in 16-byte words");

template <class R> struct Med {
    R ge0, ge1, eke0, eke1;
    R rho, te, thmoll, ap, delcm, zbrang, bpar0, bpar1;
    R blcc, xcc, esig_e, psig_e;
    R e_array0;                    // first energy of the electron grid
    R dl1[8], dl2[8], dl3[8], dl4[8], dl5[8], dl6[8];
    int sig_mono[2];
    int meke;
    int pad;
};

template <class R> struct Tables {
    int nmed;
    const Med<R> *med;
    const PW<R> *gmfp, *gbr1, *gbr2, *cohe, *pmax;              // [nmed*MXGE]
    const R *ray_xgrid, *ray_fcum, *ray_b, *ray_c;              // medium 0 only (reference quirk Q1)
    const int *ray_iarr;
    const PW<R> *esig, *psig, *ededx, *pdedx, *ebr1, *pbr1, *pbr2, *tmxs;
    const PW<R> *blcce, *etae, *etap, *q1ce, *q1cp, *q2ce, *q2cp; // [nmed*MXEKE]
    const R *range_ep;                                          // [2][nmed][MXEKE]
    const R *e_array;                                           // [nmed][MXEKE]
    const R *spin_rej;                                          // [nmed][NSPIN_MED]
    const MsBin<R> *ms;                                         // [NMS] q^(2+) surface bins
    R b2spin_min, dbeta2i, espml, dleneri, dqq1i, dllambi, dqmsi;
    const CBin<R> *cbin;                                        // [2][nmed][cb_stride] charged bins
    int cb_stride;
};

// Rectilinear voxel phantom resident in HBM: boundaries plus one byte of
// medium and one word of density ratio per voxel (x fastest).
template <class R> struct Geom {
    int nx, ny, nz;
    const R *xb, *yb, *zb;
    const uint8_t *med;
    const R *rhof;
    R pcut[MXMED], ecut[MXMED];    // region cuts per medium: global cut, or the PEGS AP/AE when higher
    R rhof_max[MXMED];             // largest density ratio of each medium in the phantom (0: absent)
    int uniform;                   // 1: equal spacing per axis, voxel lookup by multiplication
    R inv_dx, inv_dy, inv_dz;      // 1 / spacing when uniform
};

// ------------------------------------------------------------------ particles
template <class R> struct Part {
    R e, x, y, z, u, v, w, wt;
    int iq;
    int ix, iy, iz;                // ix < 0: outside the phantom
};



// ------------------------------------------------------------------ random numbers
// All sources return k * 2^-24 with k a 24-bit integer, the lattice of the
// reference's RANMAR generator (0 possible, 1 excluded).
// The plain history form's streams: one per history, keyed by (seed, history).
struct Philox {
    // Philox4x32-10 (Salmon et al., SC'11): counter-based, keyed per stream.
    uint32_t k0, k1, c0, c1, c2, c3;
    // the outputs of the last round not yet drawn, the next one in b0: each
    // draw shifts them down, so no access needs a run-time index (an indexed
    // buffer can be moved to scratch memory when registers are short)
    uint32_t b0, b1, b2, b3;
    int pos;                        // outputs of the last round drawn (4: none left)
    __device__ __forceinline__ void init(uint64_t seed, uint64_t stream) {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        c0 = (uint32_t)stream; c1 = (uint32_t)(stream >> 32); c2 = 0; c3 = 0;
        pos = 4;
    }
    __device__ __forceinline__ void round4() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3, a = k0, b = k1;
#pragma unroll
        for (int r = 0; r < 10; r++) {
            uint32_t hi0 = __umulhi(0xD2511F53u, x0), lo0 = 0xD2511F53u * x0;
            uint32_t hi1 = __umulhi(0xCD9E8D57u, x2), lo1 = 0xCD9E8D57u * x2;
            uint32_t y0 = hi1 ^ x1 ^ a, y1 = lo1, y2 = hi0 ^ x3 ^ b, y3 = lo0;
            x0 = y0; x1 = y1; x2 = y2; x3 = y3;
            a += 0x9E3779B9u; b += 0xBB67AE85u;
        }
        b0 = x0; b1 = x1; b2 = x2; b3 = x3;
        if (++c2 == 0) ++c3;
        pos = 0;
    }
    __device__ __forceinline__ uint32_t take() {
        uint32_t v = b0;
        b0 = b1; b1 = b2; b2 = b3;
        pos++;
        return v;
    }
    // drops the first k (0..3) outputs of a fresh round
    __device__ __forceinline__ void skip(int k) {
        if (k > 0) take();
        if (k > 1) take();
        if (k > 2) take();
    }
    template <class R> __device__ __forceinline__ R next() {
        if (pos >= 4) round4();
        return (R)(take() >> 8) * (R)(1.0 / 16777216.0);
    }
};

// The event form's streams, one per particle: xoshiro128** (Blackman and
// Vigna, "Scrambled linear pseudorandom number generators", ACM TOMS 47(4),
// 2021), 128 bits of state per stream, seeded from the stream's 64-bit key
// by the splitmix64 sequence (Steele, Lea and Flood, OOPSLA 2014), the
// seeding its authors recommend.  A draw costs a few 32-bit shifts, rotates
// and xors and two multiplications by small constants, where Philox4x32-10
// needs forty 32-bit multiplications, which issue at a quarter of the rate,
// per four draws: in the event form this generator made the charged-particle
// kernel 20% faster and the c3 plan 19% faster.
struct Xoshiro {
    uint32_t s0, s1, s2, s3;
    static __device__ __forceinline__ uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    static __device__ __forceinline__ uint32_t rotl(uint32_t v, int k) { return (v << k) | (v >> (32 - k)); }
    // the stream of key `stream` under the run's seed: distinct keys start
    // the splitmix64 sequence at distinct points
    __device__ __forceinline__ void init(uint64_t seed, uint64_t stream) {
        uint64_t x = seed;
        x = stream ^ splitmix(x);
        const uint64_t a = splitmix(x), b = splitmix(x);
        s0 = (uint32_t)a; s1 = (uint32_t)(a >> 32); s2 = (uint32_t)b; s3 = (uint32_t)(b >> 32);
        if ((s0 | s1 | s2 | s3) == 0u) s0 = 1u;      // the one state the generator never leaves
    }
    __device__ __forceinline__ uint32_t next32() {
        const uint32_t r = rotl(s1 * 5u, 7) * 9u;
        const uint32_t t = s1 << 9;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl(s3, 11);
        return r;
    }
    // the top 24 bits, on the reference's lattice k * 2^-24
    template <class R> __device__ __forceinline__ R next() { return (R)(next32() >> 8) * (R)(1.0 / 16777216.0); }
};

// Replays a precomputed stream of 24-bit integers (RANMAR output, or any
// injected list).  Running past the end sets *err and returns 0.5.
struct Replay {
    const uint32_t *base;
    long pos, end;
    int *err;
    template <class R> __device__ __forceinline__ R next() {
        if (pos >= end) { if (err) *err = 1; return (R)0.5; }
        return (R)base[pos++] * (R)(1.0 / 16777216.0);
    }
};

// ------------------------------------------------------------------ tallies
// Energy deposition: wt*edep is summed in a register while consecutive
// deposits hit the same region, and flushed with one FP64 atomic on change.
struct TallyNone {
    __device__ __forceinline__ void add(int, double) {}
    __device__ __forceinline__ void flush() {}
};

// With `touched` set (the tallies of a dose-influence run), every flush also
// marks the tile of TALLY_TILE regions that received it, so the batch
// statistics afterwards read only tiles that hold energy.  The mark is a
// plain byte store of 1 (idempotent, no atomic).
constexpr int TALLY_TILE_LOG2 = 6;
constexpr int TALLY_TILE = 1 << TALLY_TILE_LOG2;    // one wavefront of regions

struct TallyAtomic {
    double *endep;       // [nreg] with region 0 = outside
    uint8_t *touched;    // optional [ceil(nreg / TALLY_TILE)] tile marks
    int cur;
    double acc;
    __device__ __forceinline__ void init(double *p, uint8_t *tf = nullptr) {
        endep = p;
        touched = tf;
        cur = -1;
        acc = 0.0;
    }
    __device__ __forceinline__ void put() {
        if (cur >= 0 && acc != 0.0) {
            atomicAdd(endep + cur, acc);
            if (touched) touched[cur >> TALLY_TILE_LOG2] = 1;
        }
    }
    __device__ __forceinline__ void add(int region, double v) {
        if (region != cur) {
            put();
            cur = region;
            acc = v;
        } else {
            acc += v;
        }
    }
    __device__ __forceinline__ void flush() {
        put();
        cur = -1;
        acc = 0.0;
    }
};

// Energy balance of a run (FP64, MeV): [0] energy brought in by the
// primaries (photons E, electrons E - m, positrons E + m: the rest mass an
// annihilation returns), [1] energy of photons that left the phantom (the
// reference discards them without a tally, src/ompmc.c:2008-2017), [2]
// energy dropped without a deposit (the reference's lambda anomaly, quirk
// Q6, and particles ended by an iteration guard), [3] energy deposited in
// all regions including region 0 (summed by the batch statistics kernels).
// Without photon splitting [0] = [1] + [2] + [3] up to rounding.
enum { EB_SOURCE = 0, EB_ESCAPE = 1, EB_LOST = 2, EB_DEPOSIT = 3, EB_N = 4 };

// Sum over the wavefront, added once per wave (lane 0) to *dst when non-zero.
__device__ __forceinline__ void wave_add(double *dst, double v) {
    if (!dst) return;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(dst, v);
}

// Per-lane event counters feeding the algorithmic-bytes roofline.
struct Counters {
    unsigned long long nhowfar, nhownear, ncross, ndep, nrn, nphoton, nelectron;
};

// ------------------------------------------------------------------ transport engine
// LDS: the charged-particle event kernel's engine, whose electron bins are
// read through `eb` -- the kernel's LDS copy of them, or their place in
// device memory when they are not staged -- with generic loads; otherwise
// every bin comes from Tables::cbin with global loads.
template <class R, class G, class T, bool LDS = false> struct Engine {
    const Tables<R> &t;
    const Geom<R> &g;
    G &rng;
    T &tally;
    Counters cnt;
    bool count;
    int nsplit = 1;                 // photon splitting factor (vrt.nsplit)
    // set when a particle exceeded an iteration guard (it is dropped): 1
    // photon flight, 2 electron pass, 3 electron step, 4 rejection loop (the
    // source line of the loop in abort_site)
    int aborted = 0;
    int abort_site = 0;
    const CBin<R> *eb = nullptr;    // LDS engines: electron bins (LDS or device memory, same layout)

    // the interleaved coefficients of bin l of species qel (0 electrons,
    // 1 positrons) in medium imed, and loads of their fields
    __device__ __forceinline__ const CBin<R> *cbin(int qel, int imed, int l) const {
        // bins past MEKE are never reached below the tables' upper energy;
        // the clamp keeps a stray index inside the array
        l = min(max(l, 0), t.cb_stride - 1);
        if (LDS && qel == 0) return eb + (imed * t.cb_stride + l);
        return t.cbin + (((size_t)qel * t.nmed + imed) * t.cb_stride + l);
    }
    __device__ __forceinline__ PW<R> cld(const PW<R> *p) const {
        if (LDS) return *p;
        return gld(p);
    }
    __device__ __forceinline__ R cld(const R *p) const {
        if (LDS) return *p;
        return gld(p);
    }
    __device__ __forceinline__ R cpw(const PW<R> *p, R lv) const { return pwlf(cld(p), lv); }
    double esc = 0.0;               // wt * E of photons that left the phantom (energy balance)
    double lost = 0.0;              // energy dropped without a deposit (energy balance)

    // energy a charged particle carries into the balance: kinetic, plus the
    // two rest masses an annihilation returns for a positron
    static __device__ __forceinline__ double charged_energy(const Part<R> &p) {
        const double RMD = 0.5109989461;
        return (double)p.wt * ((double)p.e + (p.iq > 0 ? RMD : -RMD));
    }

    // Every rejection loop is bounded: past 2^20 trials of one loop (every
    // sampler here accepts a trial with a probability of a few per cent at
    // least) the particle counts as stuck; `aborted` is set, every loop then
    // ends at its next test, and the caller drops the particle and reports it
    // like the step guards.  No kernel can spin forever on a sampler.
    static constexpr int MAX_TRIALS = 1 << 20;
    __device__ __forceinline__ bool stuck(int &n, int site = __builtin_LINE()) {
        if (++n > MAX_TRIALS && !aborted) {
            aborted = 4;
            abort_site = site;
        }
        return aborted != 0;
    }

    __device__ __forceinline__ R rnd() {
        if (count) cnt.nrn++;
        return rng.template next<R>();
    }

    static __device__ __forceinline__ int pwlf_i(R lv, R c1, R c0) { return (int)(lv * c1 + c0); }
    // Photon table interval of a log energy, at least the first: a photon
    // just above PCUT = AP sits at the lower edge of the tables, where
    // float rounding of lv * c1 + c0 can give 0.99999994 instead of 1 and
    // the reference's formula an interval of -1 -- a read before the table
    // (measured: one photon in ~6e7 FP32 histories, whose garbage Rayleigh
    // fraction then kept it scattering until a guard dropped it).  In FP64
    // the formula stays at 0 there, so the clamp changes nothing.
    static __device__ __forceinline__ int photon_interval(R gle, const Med<R> &m) {
        const int i = pwlf_i(gle, m.ge1, m.ge0) - 1;
        return i < 0 ? 0 : i;
    }
    static __device__ __forceinline__ R pwlf(const PW<R> &q, R lv) { return lv * q.c1 + q.c0; }
    static __device__ __forceinline__ R pwlf(const PW<R> *tab, R lv) {
        PW<R> q = gld(tab);
        return lv * q.c1 + q.c0;
    }

    __device__ __forceinline__ int region(const Part<R> &p) const {
        return p.ix < 0 ? 0 : 1 + p.ix + g.nx * (p.iy + g.ny * p.iz);
    }
    __device__ __forceinline__ int medium(const Part<R> &p) const {
        if (p.ix < 0) return -1;
        uint8_t m = gld(g.med + (p.ix + g.nx * (p.iy + g.ny * p.iz)));
        return m == VACUUM ? -1 : (int)m;
    }
    __device__ __forceinline__ R rhof_of(const Part<R> &p) const {
        return p.ix < 0 ? (R)0 : gld(g.rhof + (p.ix + g.nx * (p.iy + g.ny * p.iz)));
    }
    __device__ __forceinline__ R pcut_of(int imed) const { return imed < 0 ? (R)0 : g.pcut[imed]; }
    __device__ __forceinline__ R ecut_of(int imed) const { return imed < 0 ? (R)0 : g.ecut[imed]; }

    // ausgab (omc_dosxyz.c:683-694)
    __device__ __forceinline__ void ausgab(const Part<R> &p, R edep) {
        if (count) cnt.ndep++;
        tally.add(region(p), (double)(p.wt * edep));
    }

    // howfar (omc_dosxyz.c:187-297): distance to the next voxel plane along z,
    // then x, then y with strict '<' (ties go to the first axis tested).
    // Returns the move: 0 none, +-1 x, +-2 y, +-3 z; *leaves when the new
    // region is outside.  The caller has checked p.ix >= 0.
    __device__ __forceinline__ int howfar(const Part<R> &p, R &ustep, bool &leaves) {
        if (count) cnt.nhowfar++;
        int mv = 0;
        R dist;
        leaves = false;
        if (p.w > (R)0) {
            dist = (gld(g.zb + (p.iz + 1)) - p.z) / p.w;
            if (dist < ustep) { ustep = dist; mv = 3; leaves = p.iz == g.nz - 1; }
        } else if (p.w < (R)0) {
            dist = -(p.z - gld(g.zb + (p.iz))) / p.w;
            if (dist < ustep) { ustep = dist; mv = -3; leaves = p.iz == 0; }
        }
        if (p.u > (R)0) {
            dist = (gld(g.xb + (p.ix + 1)) - p.x) / p.u;
            if (dist < ustep) { ustep = dist; mv = 1; leaves = p.ix == g.nx - 1; }
        } else if (p.u < (R)0) {
            dist = -(p.x - gld(g.xb + (p.ix))) / p.u;
            if (dist < ustep) { ustep = dist; mv = -1; leaves = p.ix == 0; }
        }
        if (p.v > (R)0) {
            dist = (gld(g.yb + (p.iy + 1)) - p.y) / p.v;
            if (dist < ustep) { ustep = dist; mv = 2; leaves = p.iy == g.ny - 1; }
        } else if (p.v < (R)0) {
            dist = -(p.y - gld(g.yb + (p.iy))) / p.v;
            if (dist < ustep) { ustep = dist; mv = -2; leaves = p.iy == 0; }
        }
        return mv;
    }

    __device__ __forceinline__ void apply_move(Part<R> &p, int mv, bool leaves) {
        if (count) cnt.ncross++;
        if (leaves) { p.ix = -1; return; }
        if (mv == 1) p.ix++; else if (mv == -1) p.ix--;
        else if (mv == 2) p.iy++; else if (mv == -2) p.iy--;
        else if (mv == 3) p.iz++; else if (mv == -3) p.iz--;
    }

    // hownear (omc_dosxyz.c:299-334)
    __device__ __forceinline__ R hownear(const Part<R> &p) {
        if (count) cnt.nhownear++;
        if (p.ix < 0) return (R)0;
        R tp = (R)1.0E10;
        tp = M<R>::fmin(tp, gld(g.xb + (p.ix + 1)) - p.x);
        tp = M<R>::fmin(tp, p.x - gld(g.xb + (p.ix)));
        tp = M<R>::fmin(tp, gld(g.yb + (p.iy + 1)) - p.y);
        tp = M<R>::fmin(tp, p.y - gld(g.yb + (p.iy)));
        tp = M<R>::fmin(tp, gld(g.zb + (p.iz + 1)) - p.z);
        tp = M<R>::fmin(tp, p.z - gld(g.zb + (p.iz)));
        return tp;
    }

    // ---------------------------------------------------------------- kinematics
    struct Uphi { R A, B, C, cosphi, sinphi; };

    // box-rejection azimuth (src/ompmc.c:102-122)
    __device__ __forceinline__ void azimuth(R &c, R &sn) {
        R xphi, xphi2, yphi, yphi2, r2;
        int trials = 0;
        do {
            xphi = rnd();
            xphi = (R)2.0 * xphi - (R)1.0;
            xphi2 = xphi * xphi;
            yphi = rnd();
            yphi2 = yphi * yphi;
            r2 = xphi2 + yphi2;
        } while ((r2 > (R)1.0) && !stuck(trials));
        r2 = (R)1 / r2;
        c = (xphi2 - yphi2) * r2;
        sn = (R)2.0 * xphi * yphi * r2;
    }

    __device__ __forceinline__ static void rotate(Part<R> &p, const Uphi &q, R costhe, R sinthe) {
        R sinps2 = q.A * q.A + q.B * q.B;
        if (sinps2 < (R)1.0E-20) {
            p.u = sinthe * q.cosphi;
            p.v = sinthe * q.sinphi;
            p.w = q.C * costhe;
        } else {
            R sinpsi = M<R>::sqrt(sinps2);
            R us = sinthe * q.cosphi, vs = sinthe * q.sinphi;
            R sindel = q.B / sinpsi, cosdel = q.A / sinpsi;
            p.u = q.C * cosdel * us - sindel * vs + q.A * costhe;
            p.v = q.C * sindel * us + cosdel * vs + q.B * costhe;
            p.w = -sinpsi * us + q.C * costhe;
        }
    }

    // uphi21 (src/ompmc.c:128-165)
    __device__ __forceinline__ void uphi21(Part<R> &p, Uphi &q, R costhe, R sinthe) {
        azimuth(q.cosphi, q.sinphi);
        q.A = p.u; q.B = p.v; q.C = p.w;
        rotate(p, q, costhe, sinthe);
    }

    // uphi32 (src/ompmc.c:167-202): the new particle inherits position,
    // region and weight of the one below it
    __device__ __forceinline__ static void uphi32(Part<R> &p, const Part<R> &below, const Uphi &q, R costhe, R sinthe) {
        p.x = below.x; p.y = below.y; p.z = below.z;
        p.u = below.u; p.v = below.v; p.w = below.w;
        p.ix = below.ix; p.iy = below.iy; p.iz = below.iz; p.wt = below.wt;
        rotate(p, q, costhe, sinthe);
    }

    __device__ __forceinline__ static void transfer(Part<R> &p, const Part<R> &from) {
        p.x = from.x; p.y = from.y; p.z = from.z;
        p.u = from.u; p.v = from.v; p.w = from.w;
        p.ix = from.ix; p.iy = from.iy; p.iz = from.iz; p.wt = from.wt;
    }

    // ---------------------------------------------------------------- photon interactions
    // Rayleigh (src/ompmc.c:1120-1163): form-factor CDF of medium 0 with bin
    // index 0, as the reference reads them (quirk Q1)
    __device__ OMC_INL void rayleigh(Part<R> &p, int imed, R eig, R gle, int lgle) {
        R pmax = pwlf(t.pmax + (imed * MXGE + lgle), gle);
        R xmax = (R)80.65506856998 * eig;
        R xv, costhe, csqthe, r0, r1;
        int trials = 0;
        do {
            r1 = rnd();
            int trials_in = 0;
            do {
                r0 = rnd();
                r0 *= pmax;
                int ib = t.ray_iarr[0] - 1;
                if ((t.ray_iarr[1] - 1) > ib)
                    while (r0 >= t.ray_fcum[ib + 1]) ib++;
                r0 = (r0 - t.ray_fcum[ib]) * t.ray_c[ib];
                xv = t.ray_xgrid[ib] * M<R>::exp(M<R>::log((R)1.0 + r0) * t.ray_b[ib]);
            } while ((xv >= xmax) && !stuck(trials_in));
            xv /= eig;
            costhe = (R)1.0 - (R)0.000307444456 * (xv * xv);
            csqthe = costhe * costhe;
        } while (((R)2.0 * r1 >= ((R)1.0 + csqthe)) && !stuck(trials));
        R sinthe = M<R>::sqrt((R)1.0 - csqthe);
        Uphi q;
        uphi21(p, q, costhe, sinthe);
    }

    __device__ __forceinline__ R pair_rej(const Med<R> &m, R xi, R esedei, R eseder, R tteig) {
        R a = (R)1.0 + eseder, b = (R)1.0 + esedei;
        R c = a * b / ((R)2.0 * tteig);
        R xm = xi - (R)0.5;
        return (R)2.0 + (R)3.0 * (esedei + eseder) -
               (R)4.0 * (esedei + eseder + (R)1.0 - (R)4.0 * (xm * xm)) *
                   ((R)1.0 + (R)0.25 * M<R>::log(c * c + m.zbrang * (xi * xi)));
    }

    // pair production (src/ompmc.c:1447-1685)
    __device__ OMC_INL void pair(int imed, Part<R> &p0, Part<R> &p1) {
        const Med<R> &m = t.med[imed];
        const R RM = (R)0.5109989461;
        R eig = p0.e, ese1, ese2;
        int iq1, iq2, l, l1;
        if (eig <= (R)2.1) {
            R r0 = rnd(), r1 = rnd();
            ese2 = RM + (R)0.5 * r0 * (eig - (R)2.0 * RM);
            ese1 = eig - ese2;
            if (r1 < (R)0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
        } else {
            R Amax, Bmax, delta, aux;
            if (eig < (R)50.0) {
                l = 4; l1 = 5;
                delta = (R)4.0 * m.delcm / eig;
                if (delta < (R)1.0) {
                    Amax = m.dl1[l] + delta * (m.dl2[l] + delta * m.dl3[l]);
                    Bmax = m.dl1[l1] + delta * (m.dl2[l1] + delta * m.dl3[l1]);
                } else {
                    aux = M<R>::log(delta + m.dl6[l]);
                    Amax = m.dl4[l] + m.dl5[l] * aux;
                    Bmax = m.dl4[l1] + m.dl5[l1] * aux;
                }
                aux = (R)1.0 - (R)2.0 * RM / eig;
                aux = aux * aux;
                aux *= Amax / (R)3.0;
                aux /= (Bmax + aux);
            } else {
                l = 6;
                Amax = m.dl1[l];
                Bmax = m.dl1[l + 1];
                aux = m.bpar1 * ((R)1.0 - m.bpar0 * RM / eig);
            }
            R eavail = eig - (R)2.0 * RM, rejf, rejmax, r4;
            int trials = 0;
            do {
                R br, r0 = rnd(), r1 = rnd();
                r4 = rnd();
                if (r0 > aux) {
                    br = (R)0.5 * r1;
                    rejmax = Bmax;
                    l1 = l + 1;
                } else {
                    R r2 = rnd(), r3 = rnd();
                    br = (R)0.5 * ((R)1.0 - M<R>::fmax(M<R>::fmax(r1, r2), r3));
                    rejmax = Amax;
                    l1 = l;
                }
                ese2 = br * eavail + RM;
                ese1 = eig - ese2;
                delta = (eig * m.delcm) / (ese2 * ese1);
                if (delta < (R)1.0) rejf = m.dl1[l1] + delta * (m.dl2[l1] + delta * m.dl3[l1]);
                else rejf = m.dl4[l1] + m.dl5[l1] * M<R>::log(delta + m.dl6[l1]);
            } while ((r4 * rejmax > rejf) && !stuck(trials));
            ese1 = eig - ese2;
            r4 = rnd();
            if (r4 < (R)0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
        }
        p0.e = ese1;
        p1.e = ese2;
        Uphi q;
        for (int j = 0; j < 2; j++) {
            R ese = j == 0 ? ese1 : ese2;
            R tteig = eig / RM, ttese = ese / RM;
            R esedei = ttese / (tteig - ttese), eseder = (R)1.0 / esedei;
            R pt = (R)3.14159265358979323846 * ttese;
            R ximin = (R)1.0 / ((R)1.0 + pt * pt);
            R rejmin = pair_rej(m, ximin, esedei, eseder, tteig);
            R ya = ((R)2.0 / tteig) * ((R)2.0 / tteig);
            R xitry = M<R>::fmax((R)0.01, M<R>::fmax(ximin, M<R>::fmin((R)0.5, M<R>::sqrt(ya / m.zbrang))));
            R galpha = (R)1.0 + (R)0.25 * M<R>::log(ya + m.zbrang * xitry * xitry);
            R gbeta = (R)0.5 * m.zbrang * xitry / (ya + m.zbrang * xitry * xitry);
            galpha -= gbeta * (xitry - (R)0.5);
            R ximid = galpha / ((R)3.0 * gbeta);
            if (galpha >= (R)0.0) ximid = (R)0.5 - ximid + M<R>::sqrt(ximid * ximid + (R)0.25);
            else ximid = (R)0.5 - ximid - M<R>::sqrt(ximid * ximid + (R)0.25);
            ximid = M<R>::fmax((R)0.01, M<R>::fmax(ximin, M<R>::fmin((R)0.5, ximid)));
            R rejmid = pair_rej(m, ximid, esedei, eseder, tteig);
            R rejtop = (R)1.0 * M<R>::fmax(rejmin, rejmid);
            R theta, rtest, rejfactor;
            // accepted when either test passes (reference quirk Q2)
            int trials = 0;
            do {
                R xitst = rnd();
                R rejtst = pair_rej(m, xitst, esedei, eseder, tteig);
                rtest = rnd();
                theta = M<R>::sqrt((R)1.0 / xitst - (R)1.0) / ttese;
                rejfactor = rejtst / rejtop;
            } while (((rtest > rejfactor) && (theta >= (R)3.14159265358979323846)) && !stuck(trials));
            R sinthe = M<R>::sin(theta), costhe = M<R>::cos(theta);
            if (j == 0) {
                uphi21(p0, q, costhe, sinthe);
            } else {
                uphi32(p1, p0, q, costhe, -sinthe);
            }
        }
        p1.iq = iq2;
        p0.iq = iq1;
    }

    // Klein-Nishina Compton (src/ompmc.c:1688-1801)
    __device__ OMC_INL void compton(Part<R> &p0, Part<R> &p1) {
        const R RM = (R)0.5109989461;
        R eig = p0.e, ko = p0.e / RM;
        R broi = (R)1.0 + (R)2.0 * ko, bro = (R)1.0 / broi;
        bool first = true;
        R sinthe = (R)0, costhe, br, aux, rejf3, temp = (R)0;
        R alph1 = (R)0, alph2 = (R)0, alpha = (R)0, rejmax = (R)0;
        int trials = 0;
        do {
            if (ko > (R)2.0) {
                if (first) {
                    alph1 = M<R>::log(broi);
                    alph2 = ko * (broi + (R)1.0) * (bro * bro);
                    alpha = alph1 + alph2;
                }
                R r3;
                int trials_in = 0;
                do {
                    R r1 = rnd(), r2 = rnd();
                    if (r1 * alpha < alph1) br = M<R>::exp(alph1 * r2) * bro;
                    else br = M<R>::sqrt(r2 * (broi * broi) + ((R)1.0 - r2)) * bro;
                    temp = ((R)1.0 - br) / (ko * br);
                    sinthe = M<R>::fmax((R)0.0, temp * ((R)2.0 - temp));
                    aux = (R)1.0 + br * br;
                    rejf3 = aux - br * sinthe;
                    r3 = rnd();
                } while ((r3 * aux > rejf3) && !stuck(trials_in));
            } else {
                if (first) rejmax = broi + bro;
                R r2;
                int trials_in = 0;
                do {
                    R r1 = rnd();
                    r2 = rnd();
                    br = bro + ((R)1.0 - bro) * r1;
                    temp = ((R)1.0 - br) / (ko * br);
                    sinthe = M<R>::fmax((R)0.0, temp * ((R)2.0 - temp));
                    rejf3 = (R)1.0 + br * br - br * sinthe;
                } while ((r2 * br * rejmax > rejf3) && !stuck(trials_in));
            }
            first = false;
        } while (((br < bro) || (br > (R)1)) && !stuck(trials));
        costhe = (R)1.0 - temp;
        sinthe = M<R>::sqrt(sinthe);
        R esg = br * eig, ese = eig - esg + RM;
        p0.e = esg;
        Uphi q;
        uphi21(p0, q, costhe, sinthe);
        aux = (R)1.0 + br * br - (R)2.0 * br * costhe;
        if (aux > (R)1.0E-8) {
            costhe = ((R)1.0 - br * costhe) / M<R>::sqrt(aux);
            sinthe = ((R)1.0 - costhe) * ((R)1.0 + costhe);
            sinthe = sinthe > (R)0.0 ? -M<R>::sqrt(sinthe) : (R)0.0;
        } else {
            costhe = (R)0.0;
            sinthe = (R)-1.0;
        }
        uphi32(p1, p0, q, costhe, sinthe);
        p1.e = ese;
        p1.iq = -1;
    }

    // photo-electric (src/ompmc.c:1804-1864); no binding energy (quirk Q3)
    __device__ OMC_INL void photo(Part<R> &p) {
        const R RM = (R)0.5109989461;
        p.e += RM;
        p.iq = -1;
        R eelec = p.e;
        if (eelec > ecut_of(medium(p))) {
            R beta = M<R>::sqrt((eelec - RM) * (eelec + RM)) / eelec;
            R gamma = eelec / RM;
            R alpha = (R)0.5 * gamma - (R)0.5 + (R)1.0 / gamma;
            R ratio = beta / alpha, costhe, sinth2, xi, r2;
            int trials = 0;
            do {
                R rn = rnd();
                rn = (R)2.0 * rn - (R)1.0;
                if (ratio <= (R)0.2) {
                    R fk = rn + (R)0.5 * ratio * ((R)1.0 - rn) * ((R)1.0 + rn);
                    if (gamma < (R)100.0) {
                        costhe = (beta + fk) / ((R)1.0 + beta * fk);
                    } else {
                        if (fk > (R)0.0) {
                            costhe = (R)1.0 - ((R)1.0 - fk) * (gamma - (R)3.0) / ((R)2.0 * ((R)1.0 + fk) * M<R>::pow3(gamma - (R)1.0));
                        } else {
                            costhe = (beta + fk) / ((R)1.0 + beta * fk);
                        }
                    }
                    xi = ((R)1.0 + beta * fk) * (gamma * gamma);
                } else {
                    xi = (gamma * gamma) * ((R)1.0 + alpha * (M<R>::sqrt((R)1.0 + ratio * ((R)2.0 * rn + ratio)) - (R)1.0));
                    costhe = ((R)1.0 - (R)1.0 / xi) / beta;
                }
                sinth2 = M<R>::fmax(((R)1.0 - costhe) * ((R)1.0 + costhe), (R)0.0);
                r2 = rnd();
            } while ((r2 > (R)0.5 * ((R)1.0 + gamma) * sinth2 * xi / gamma) && !stuck(trials));
            Uphi q;
            uphi21(p, q, costhe, M<R>::sqrt(sinth2));
        }
    }

    // photon step (src/ompmc.c:1867-2139) for nsplit = 1: transport to the
    // next interaction (or out of the phantom / below PCUT), interact, return.
    __device__ OMC_INL int photon(Part<R> &p, Part<R> &q) {
        if (count) cnt.nphoton++;
        int imed = medium(p);
        R eig = p.e;
        if (eig <= pcut_of(imed) || p.wt == (R)0) {
            ausgab(p, eig);
            return -1;
        }
        R gle = M<R>::log(eig);
        // splitting bookkeeping of the reference with nsplit = 1: two draws,
        // eta' = (1 - r + 1) - 1, survivor index 0
        R rn = rnd();
        R eta = (R)1.0 - rn + (R)1.0;
        (void)rnd();
        eta -= (R)1.0;
        R dpmfp = -M<R>::log(eta);
        int lgle = 0;
        R cohfac = (R)0, gmfp = (R)0;
        for (;;) {
            R tstep;
            if (imed != -1) {
                const Med<R> &m = t.med[imed];
                lgle = photon_interval(gle, m);
                R gmfpr0 = pwlf(t.gmfp + (imed * MXGE + lgle), gle);
                gmfp = gmfpr0 / rhof_of(p);
                cohfac = pwlf(t.cohe + (imed * MXGE + lgle), gle);
                gmfp *= cohfac;
                tstep = gmfp * dpmfp;
            } else {
                tstep = (R)1.0E8;
            }
            R ustep = tstep;
            if (p.ix < 0) {           // howfar: outside the geometry -> discard
                if (count) cnt.nhowfar++;
                p.x += ustep * p.u; p.y += ustep * p.v; p.z += ustep * p.w;
                esc += (double)p.wt * (double)eig;
                return -1;
            }
            bool leaves;
            int mv = howfar(p, ustep, leaves);
            p.x += ustep * p.u;
            p.y += ustep * p.v;
            p.z += ustep * p.w;
            if (imed != -1) dpmfp = M<R>::fmax((R)0.0, dpmfp - ustep / gmfp);
            if (mv != 0) {
                apply_move(p, mv, leaves);
                imed = medium(p);
            }
            if (imed != -1 && dpmfp <= (R)1.0E-05) break;
        }
        return interact(p, q, imed, eig, gle, lgle, cohfac);
    }

    // interaction at the end of a photon flight (src/ompmc.c:2044-2086):
    // Rayleigh with probability 1 - cohfac, else pair / Compton / photo by the
    // branching ratios of the medium at the site
    __device__ __forceinline__ int interact(Part<R> &p, Part<R> &q, int imed, R eig, R gle, int lgle, R cohfac) {
        R r = rnd();
        if (r <= (R)1.0 - cohfac) {
            rayleigh(p, imed, eig, gle, lgle);
            return 0;
        }
        return interact_pcp(p, q, imed, eig, gle, lgle);
    }

    // pair production, Compton or photo-electric absorption by the branching
    // ratios of the medium at the site (src/ompmc.c:2060-2086); returns 1
    // when q holds a second particle
    __device__ __forceinline__ int interact_pcp(Part<R> &p, Part<R> &q, int imed, R eig, R gle, int lgle) {
        const R RM = (R)0.5109989461;
        R r = rnd();
        R gbr1 = pwlf(t.gbr1 + (imed * MXGE + lgle), gle);
        if (r <= gbr1 && eig > (R)2.0 * RM) {
            pair(imed, p, q);
            return 1;
        }
        R gbr2 = pwlf(t.gbr2 + (imed * MXGE + lgle), gle);
        if (r < gbr2) {
            compton(p, q);
            return 1;
        }
        photo(p);
        return 0;
    }

    // voxel index along one axis: direct for equal spacing, else bisection
    __device__ __forceinline__ int locate(const R *b, int n, R x, R inv) const {
        int i;
        if (g.uniform) {
            i = (int)((x - b[0]) * inv);
        } else {
            int lo = 0, hi = n;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (b[mid] <= x) lo = mid;
                else hi = mid;
            }
            i = lo;
        }
        return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    }

    // Photon flight by delta (Woodcock) tracking, for the event form.  With
    // the photon energy fixed, the attenuation coefficient of a voxel is
    // rhof / (gmfp(E) cohfac(E)) of its medium, so the largest over the
    // phantom is the largest over media of rhof_max / (gmfp cohfac): an exact
    // majorant.  Flights are sampled with it and a tentative site is real
    // with probability mu(site) / mu_max; the interaction then follows the
    // reference exactly.  Same distribution of sites as the voxel-by-voxel
    // flight of photon() (not the same draws), without a dependent load per
    // voxel crossed.  The attenuation per unit density ratio of every medium,
    // 1 / (gmfp cohfac), is evaluated once per flight (the energy does not
    // change until the photon interacts) and kept in registers: a tentative
    // site then costs one round of loads (medium and density of its voxel),
    // and the tables are read again only for a real interaction.  The
    // majorant is formed from the same products, so mu(site) <= mu_max holds
    // exactly in floating point.
    __device__ OMC_INL int photon_delta(Part<R> &p, Part<R> &q) {
        if (count) cnt.nphoton++;
        int imed = medium(p);
        R eig = p.e;
        if (eig <= pcut_of(imed) || p.wt == (R)0) {
            ausgab(p, eig);
            return -1;
        }
        R gle = M<R>::log(eig);
        R mu_max = (R)0;
        R mu_rho[MXMED];
        // Unconditional loads (a medium beyond nmed reads medium 0's tables and
        // is then ignored): with no branch between them, every medium's loads
        // are in flight at once instead of one medium after another.
#pragma unroll
        for (int m = 0; m < MXMED; m++) {
            const int mc = m < t.nmed ? m : 0;
            const Med<R> &md = t.med[mc];
            const int lg = photon_interval(gle, md);
            const R mr = (R)1.0 / (pwlf(t.gmfp + (mc * MXGE + lg), gle) * pwlf(t.cohe + (mc * MXGE + lg), gle));
            const R rm = g.rhof_max[m];
            const bool use = m < t.nmed && rm > (R)0;
            mu_rho[m] = use ? mr : (R)0;
            mu_max = use ? M<R>::fmax(mu_max, mr * rm) : mu_max;
        }
        if (!(mu_max > (R)0)) {          // nothing to interact with: leaves
            p.ix = -1;
            esc += (double)p.wt * (double)eig;
            return -1;
        }
        const R xlo = gld(g.xb + (0)), xhi = gld(g.xb + (g.nx)), ylo = gld(g.yb + (0)), yhi = gld(g.yb + (g.ny)), zlo = gld(g.zb + (0)), zhi = gld(g.zb + (g.nz));
        for (int iters = 0;; iters++) {
            if (iters > (1 << 24)) {     // no progress: dropped, reported through `aborted`
                aborted = 1;
                abort_site = __LINE__;
                p.ix = -1;
                lost += (double)p.wt * (double)eig;
                return -1;
            }
            R rn = rnd();
            R s = -M<R>::log((R)1.0 - rn) / mu_max;
            if (count) cnt.nhowfar++;
            p.x += s * p.u;
            p.y += s * p.v;
            p.z += s * p.w;
            if (!(p.x >= xlo && p.x <= xhi && p.y >= ylo && p.y <= yhi && p.z >= zlo && p.z <= zhi)) {
                p.ix = -1;               // left the phantom (howfar discards it)
                esc += (double)p.wt * (double)eig;
                return -1;
            }
            p.ix = locate(g.xb, g.nx, p.x, g.inv_dx);
            p.iy = locate(g.yb, g.ny, p.y, g.inv_dy);
            p.iz = locate(g.zb, g.nz, p.z, g.inv_dz);
            if (count) cnt.ncross++;
            size_t v = p.ix + (size_t)g.nx * (p.iy + (size_t)g.ny * p.iz);
            const uint8_t mm = gld(g.med + (v));
            const R rho = gld(g.rhof + (v));
            R r2 = rnd();
            if (mm == VACUUM) continue;
            R mur = mu_rho[0];
#pragma unroll
            for (int m = 1; m < MXMED; m++) mur = (int)mm == m ? mu_rho[m] : mur;   // no dynamic register index
            if (r2 * mu_max <= mur * rho) {
                imed = (int)mm;
                break;
            }
        }
        // The interaction follows the loop rather than sitting in it: lanes
        // reach their real sites after different numbers of tentative
        // collisions, and out here every lane of the wave whose photon
        // interacts does so together, once per flight (inside the loop the
        // wave ran the interaction code once per trial in which any lane
        // accepted).  Each lane's operations and draws are the same.
        const Med<R> &md = t.med[imed];
        int lgle = photon_interval(gle, md);
        R cohfac = pwlf(t.cohe + (imed * MXGE + lgle), gle);
        return interact(p, q, imed, eig, gle, lgle, cohfac);
    }

    // Photon data of every medium at one energy: the sub-photons of a split
    // share the photon's energy, so their walks select these per voxel
    // instead of loading them through the voxel's medium.  Unconditional
    // loads, as in photon_delta (a medium beyond nmed reads medium 0's).
    struct SplitXs { R gmfp0[MXMED], coh[MXMED]; int lgle[MXMED]; };
    __device__ __forceinline__ void split_xs(R gle, SplitXs &xs) {
#pragma unroll
        for (int m = 0; m < MXMED; m++) {
            const int mc = m < t.nmed ? m : 0;
            const Med<R> &md = t.med[mc];
            const int lg = photon_interval(gle, md);
            xs.lgle[m] = lg;
            xs.gmfp0[m] = pwlf(t.gmfp + (mc * MXGE + lg), gle);
            xs.coh[m] = pwlf(t.cohe + (mc * MXGE + lg), gle);
        }
    }

    // medium (-1: vacuum or outside) and density ratio of the particle's
    // voxel, loaded together
    __device__ __forceinline__ void voxel_of(const Part<R> &p, int &imed, R &rho) const {
        if (p.ix < 0) {
            imed = -1;
            rho = (R)0;
            return;
        }
        const int v = p.ix + g.nx * (p.iy + g.ny * p.iz);
        const uint8_t m = gld(g.med + v);
        rho = gld(g.rhof + v);
        imed = m == VACUUM ? -1 : (int)m;
    }

    // Flight of one sub-photon of a split (src/ompmc.c:1960-2032): from its
    // start, `dpmfp` mean free paths along its direction, voxel by voxel.
    // Returns false when it leaves the geometry (the reference discards it
    // and ends the split; the `left` sub-photons from this one on, which
    // would follow it along the same ray, count as escaped in the energy
    // balance), true at its interaction site, with imed the
    // medium there.  lgle, cohfac and gmfp are the caller's: like the
    // reference's function-scope variables they keep the values of the last
    // voxel with a medium, which the interaction then uses.  The step
    // arithmetic is the reference's, in its order, on values from `xs`.
    __device__ __forceinline__ bool split_fly(Part<R> &p, R dpmfp, const SplitXs &xs, int &imed, int &lgle,
                                              R &cohfac, R &gmfp, int left) {
        R rho;
        voxel_of(p, imed, rho);
        for (;;) {
            R tstep;
            if (imed != -1) {
                R g0 = xs.gmfp0[0], ch = xs.coh[0];
                int lg = xs.lgle[0];
#pragma unroll
                for (int m = 1; m < MXMED; m++) {       // no dynamic register index
                    const bool s = imed == m;
                    g0 = s ? xs.gmfp0[m] : g0;
                    ch = s ? xs.coh[m] : ch;
                    lg = s ? xs.lgle[m] : lg;
                }
                lgle = lg;
                gmfp = g0 / rho;
                cohfac = ch;
                gmfp *= cohfac;
                tstep = gmfp * dpmfp;
            } else {
                tstep = (R)1.0E8;
            }
            R ustep = tstep;
            if (p.ix < 0) {           // howfar: outside the geometry -> discard
                if (count) cnt.nhowfar++;
                p.x += ustep * p.u; p.y += ustep * p.v; p.z += ustep * p.w;
                esc += (double)left * (double)p.wt * (double)p.e;
                return false;
            }
            bool leaves;
            int mv = howfar(p, ustep, leaves);
            p.x += ustep * p.u;
            p.y += ustep * p.v;
            p.z += ustep * p.w;
            if (imed != -1) dpmfp = M<R>::fmax((R)0.0, dpmfp - ustep / gmfp);
            if (mv != 0) {
                apply_move(p, mv, leaves);
                voxel_of(p, imed, rho);
            }
            if (imed != -1 && dpmfp <= (R)1.0E-05) return true;
        }
    }

    // photon() with splitting (src/ompmc.c:1867-2139) on a stack array: the
    // photon on top of st[0..np] is split into nsplit sub-photons with
    // stratified numbers of mean free paths along its ray; each interacts at
    // its site, and of the scattered photons only those of one randomly chosen
    // sub-photon survive, with the full weight.  Returns the new np (the
    // reference's stack order, including its in-place compaction, which
    // copies energy, charge, direction and weight but not the position).
    __device__ int photon_split(Part<R> *st, int np, int cap, int *err) {
        const R RM = (R)0.5109989461;
        if (count) cnt.nphoton++;
        int imed = medium(st[np]);
        R eig = st[np].e;
        if (eig <= pcut_of(imed) || st[np].wt == (R)0) {
            ausgab(st[np], eig);
            return np - 1;
        }
        R gle = M<R>::log(eig);
        int lgle = 0;
        R cohfac = (R)0, gmfp = (R)0;
        const int ns = nsplit;
        SplitXs xs;
        for (;;) {                                  // start of the splitting of the top photon
            split_xs(gle, xs);
            R rnno = rnd();
            rnno /= (R)ns;
            R d_eta = (R)1.0 / (R)ns;
            R eta_prime = (R)1.0 - rnno + d_eta;
            Part<R> save = st[np];
            R wtsave = save.wt / (R)ns;
            np -= 1;
            rnno = rnd();
            R a_survive = rnno * (R)ns;
            int i_survive = (int)a_survive;
            R dpmfp_old = (R)0.0;
            for (int isplit = 0; isplit < ns; isplit++) {
                eta_prime -= d_eta;
                if (eta_prime <= (R)0.0) break;
                R dpmfp = -M<R>::log(eta_prime) - dpmfp_old;
                dpmfp_old += dpmfp;
                np += 1;
                if (np >= cap - 1) {
                    *err = 1;
                    return -1;
                }
                Part<R> &p = st[np];
                p.x = save.x; p.y = save.y; p.z = save.z;
                p.u = save.u; p.v = save.v; p.w = save.w;
                p.e = save.e; p.wt = wtsave;
                p.ix = save.ix; p.iy = save.iy; p.iz = save.iz; p.iq = 0;
                if (!split_fly(p, dpmfp, xs, imed, lgle, cohfac, gmfp, ns - isplit)) {
                    np -= 1;
                    if (np < 0) return np;
                    break;                          // end of the splitting loop
                }
                save.x = p.x; save.y = p.y; save.z = p.z;
                save.ix = p.ix; save.iy = p.iy; save.iz = p.iz;
                R r = rnd();
                if (r <= (R)1.0 - cohfac) {
                    if (isplit != i_survive) {
                        np -= 1;
                    } else {
                        st[np].wt *= (R)ns;
                        rayleigh(st[np], imed, eig, gle, lgle);
                    }
                    continue;
                }
                r = rnd();
                int npold = np;
                R gbr1 = pwlf(t.gbr1 + (imed * MXGE + lgle), gle);
                if (r <= gbr1 && eig > (R)2.0 * RM) {
                    pair(imed, st[np], st[np + 1]);
                    np += 1;
                } else {
                    R gbr2 = pwlf(t.gbr2 + (imed * MXGE + lgle), gle);
                    if (r < gbr2) {
                        compton(st[np], st[np + 1]);
                        np += 1;
                    } else {
                        photo(st[np]);
                    }
                }
                int ip = npold;
                do {
                    if (st[ip].iq == 0) {
                        if (isplit != i_survive) {
                            if (ip < np) {
                                st[ip].e = st[np].e; st[ip].iq = st[np].iq;
                                st[ip].u = st[np].u; st[ip].v = st[np].v;
                                st[ip].w = st[np].w; st[ip].wt = st[np].wt;
                            }
                            np -= 1;
                        } else {
                            st[ip].wt *= (R)ns;
                            ip += 1;
                        }
                    } else {
                        ip += 1;
                    }
                } while (ip <= np);
            }
            if (np < 0) return np;
            if (st[np].iq != 0) return np;
            eig = st[np].e;
            imed = medium(st[np]);
            if (eig <= pcut_of(imed)) {
                ausgab(st[np], eig);
                return np - 1;
            }
            gle = M<R>::log(eig);
        }
    }

    // ---------------------------------------------------------------- electron physics
    struct Spinr { int i, j; };
    struct Mscats { int i, j; R omega2; };

    // spin rejection (src/ompmc.c:3115-3186)
    __device__ OMC_INL R spin_rejection(int imed, int qel, R elke, R beta2, R q1, R cost, bool &spin_index, bool is_single,
                                Spinr &sr) {
        if (spin_index) {
            spin_index = false;
            R ai;
            if (beta2 >= t.b2spin_min) {
                ai = (beta2 - t.b2spin_min) * t.dbeta2i;
                sr.i = (int)ai;
                ai -= (R)sr.i;
                sr.i += MXE_SPIN + 1;
            } else if (elke > t.espml) {
                ai = (elke - t.espml) * t.dleneri;
                sr.i = (int)ai;
                ai -= (R)sr.i;
            } else {
                sr.i = 0;
                ai = (R)-1.0;
            }
            R rn = rnd();
            if (rn < ai) sr.i += 1;
            if (is_single) {
                sr.j = 0;
            } else {
                R qq1 = (R)2.0 * q1;
                qq1 = qq1 / ((R)1.0 + qq1);
                R aj = qq1 * t.dqq1i;
                sr.j = (int)aj;
                if (sr.j >= MXQ_SPIN) {
                    sr.j = MXQ_SPIN;
                } else {
                    aj -= (R)sr.j;
                    rn = rnd();
                    if (rn < aj) sr.j += 1;
                }
            }
        }
        R xi = M<R>::sqrt((R)0.5 * ((R)1.0 - cost));
        R ak = xi * (R)MXU_SPIN;
        int k = (int)ak;
        ak -= (R)k;
        const R *row = t.spin_rej + (size_t)imed * NSPIN_MED +
                       ((qel * (MXE_SPIN1 + 1) + sr.i) * (MXQ_SPIN + 1) + sr.j) * (MXU_SPIN + 1);
        return ((R)1.0 - ak) * gld(row + (k)) + ak * gld(row + (k + 1));
    }

    // single elastic scattering (src/ompmc.c:3188-3217)
    __device__ OMC_INL void sscat(int imed, int qel, R chia2, R elke, R beta2, R &cost, R &sint) {
        bool spin_index = true;
        Spinr sr = {0, 0};
        R xi, rejf, rn;
        int trials = 0;
        do {
            xi = rnd();
            xi = (R)2.0 * chia2 * xi / ((R)1.0 - xi + chia2);
            cost = (R)1.0 - xi;
            rejf = spin_rejection(imed, qel, elke, beta2, (R)0, cost, spin_index, true, sr);
            rn = rnd();
        } while ((rn > rejf) && !stuck(trials));
        sint = M<R>::sqrt(xi * ((R)2.0 - xi));
    }

    // multiple-scattering angle (src/ompmc.c:3624-3803).  Within a bin of the
    // q^(2+) surface the reference's interpolation step is du = 0 (it
    // subtracts the bin value from itself), so u is the tabulated value.
    __device__ OMC_INL void mscat(int imed, int qel, bool &spin_index, bool &find_index, R elke, R beta2, R q1, R lambda,
                          R chia2, R &cost, R &sint, Mscats &ms, Spinr &sr) {
        R xi, rejf, rn;
        R explambda = M<R>::exp(-lambda);
        if (lambda <= (R)13.8) {
            R sprob = rnd();
            if (sprob < explambda) { cost = (R)1.0; sint = (R)0.0; return; }
            R wsum = ((R)1.0 + lambda) * explambda;
            if (sprob < wsum) {
                int trials = 0;
                do {
                    xi = rnd();
                    xi = (R)2.0 * chia2 * xi / ((R)1.0 - xi + chia2);
                    cost = (R)1.0 - xi;
                    rejf = spin_rejection(imed, qel, elke, beta2, q1, cost, spin_index, false, sr);
                    rn = rnd();
                } while ((rn > rejf) && !stuck(trials));
                sint = M<R>::sqrt(xi * ((R)2.0 - xi));
                return;
            }
            if (lambda <= (R)1) {
                int icount = 0;
                R wprob = explambda, sinz, cosz;
                wsum = explambda;
                cost = (R)1.0;
                sint = (R)0.0;
                int trials = 0;
                do {
                    icount += 1;
                    if (icount > 20) break;
                    wprob = wprob * lambda / (R)icount;
                    wsum = wsum + wprob;
                    int trials_in = 0;
                    do {
                        xi = rnd();
                        xi = (R)2.0 * chia2 * xi / ((R)1.0 - xi + chia2);
                        cosz = (R)1.0 - xi;
                        rejf = spin_rejection(imed, qel, elke, beta2, q1, cosz, spin_index, false, sr);
                        rn = rnd();
                    } while ((rn > rejf) && !stuck(trials_in));
                    sinz = xi * ((R)2.0 - xi);
                    if (sinz > (R)1.0E-20) {
                        sinz = M<R>::sqrt(sinz);
                        xi = rnd();
                        R phi = xi * (R)6.2831853;
                        cost = cost * cosz - sint * sinz * M<R>::cos(phi);
                        sint = M<R>::sqrt(M<R>::fmax((R)0.0, ((R)1.0 - cost) * ((R)1.0 + cost)));
                    }
                } while ((wsum <= sprob) && !stuck(trials));
                return;
            }
        }
        if (lambda <= (R)1.0E5) {
            R llmbda = M<R>::log(lambda);
            if (find_index) {
                R ai = llmbda * t.dllambi;
                ms.i = (int)ai;
                ai -= (R)ms.i;
                xi = rnd();
                if (xi < ai) ms.i += 1;
                if (q1 < (R)1.0E-3) {
                    ms.j = 0;
                } else if (q1 < (R)0.5) {
                    R aj = q1 * t.dqmsi;
                    ms.j = (int)aj;
                    aj -= (R)ms.j;
                    xi = rnd();
                    if (xi < aj) ms.j += 1;
                } else {
                    ms.j = MXQ_MS;
                }
                if (llmbda < (R)2.2299)
                    ms.omega2 = chia2 * (lambda + (R)4.0) *
                                ((R)1.347006 + llmbda * ((R)0.209364 - llmbda * ((R)0.45525 - llmbda * ((R)0.50142 - (R)0.081234 * llmbda))));
                else
                    ms.omega2 = chia2 * (lambda + (R)4.0) *
                                ((R)-2.77164 + llmbda * ((R)2.94874 - llmbda * ((R)0.1535754 - llmbda * (R)0.00552888)));
                find_index = false;
            }
            const MsBin<R> *row = t.ms + (ms.i * (MXQ_MS + 1) + ms.j) * (MXU_MS + 1);
            int trials = 0;
            do {
                xi = rnd();
                R ak = xi * (R)MXU_MS;
                int k = (int)ak;
                ak -= (R)k;
                MsBin<R> bin = gld(row + k);
                if (ak > bin.w) bin = gld(row + bin.i);
                R a = bin.f;
                R u = bin.u;
                R du = (R)0;
                xi = rnd();
                if (M<R>::fabs(a) < (R)0.2) {
                    R x1 = (R)0.5 * ((R)1.0 - xi) * a;
                    u += xi * du * ((R)1.0 + x1 * ((R)1.0 - xi * a));
                } else {
                    u -= du / a * ((R)1.0 - M<R>::sqrt((R)1.0 + xi * a * ((R)2.0 + a)));
                }
                xi = ms.omega2 * u / ((R)1.0 + (R)0.5 * ms.omega2 - u);
                if (xi > (R)1.99999) xi = (R)1.99999;
                cost = (R)1.0 - xi;
                rejf = spin_rejection(imed, qel, elke, beta2, q1, cost, spin_index, false, sr);
                rn = rnd();
            } while ((rn > rejf) && !stuck(trials));
            sint = M<R>::sqrt(xi * ((R)2.0 - xi));
        }
    }

    // two-substep multiple scattering, PRESTA-II style (src/ompmc.c:3805-3994)
    __device__ OMC_INL R msdist(const Part<R> &p, int imed, int iq, R rhof, R de, R tustep, R eke, Part<R> &out) {
        const R RM = (R)0.5109989461;
        const Med<R> &m = t.med[imed];
        int qel = (1 + iq) / 2;
        Mscats ms = {0, 0, (R)0};
        Spinr sr = {0, 0};
        R blccc = m.blcc, xcccc = m.xcc;
        R e = eke - (R)0.5 * de;
        R tau = e / RM, tau2 = tau * tau;
        R epsilon = de / eke, epsilonp = de / e;
        e *= ((R)1.0 - (epsilonp * epsilonp) * ((R)6.0 + (R)10.0 * tau + (R)5.0 * tau2) / ((R)24.0 * tau2 + (R)72.0 * tau + (R)48.0));
        R p2 = e * (e + (R)2.0 * RM);
        R beta2 = p2 / (p2 + RM * RM);
        R chia2 = xcccc / ((R)4.0 * p2 * blccc);
        R lambda = (R)0.5 * tustep * rhof * blccc / beta2;
        R tq = epsilonp / ((tau + (R)1.0) * (tau + (R)2.0));
        R temp2 = (R)0.166666 * ((R)4.0 + tau * ((R)6.0 + tau * ((R)7.0 + tau * ((R)4.0 + tau)))) * tq * tq;
        lambda *= ((R)1.0 - temp2);
        R elke = M<R>::log(e);
        int lelke = pwlf_i(elke, m.eke1, m.eke0) - 1;
        if (lelke < 0) {
            lelke = 0;
            elke = ((R)1.0 - m.eke0) / m.eke1;
        }
        const CBin<R> *cb = cbin(qel, imed, lelke);
        R etap = cpw(&cb->eta, elke);
        R xi_corr = cpw(&cb->q1c, elke);
        R gamma = cpw(&cb->q2c, elke);
        R ms_corr = cpw(&cb->blcc, elke);
        chia2 *= etap;
        lambda /= (etap * ((R)1.0 + chia2));
        lambda *= ms_corr;
        R chilog = M<R>::log((R)1.0 + (R)1.0 / chia2);
        R q1 = (R)2.0 * chia2 * (chilog * ((R)1.0 + chia2) - (R)1.0);
        gamma = (R)6.0 * chia2 * ((R)1.0 + chia2) * (chilog * ((R)1.0 + (R)2.0 * chia2) - (R)2.0) / q1 * gamma;
        R xi = q1 * lambda;
        bool find_index = true, spin_index = true;
        R w1 = (R)0, sint1 = (R)0, w2 = (R)0, sint2 = (R)0, cphi1, sphi1, cphi2, sphi2;
        mscat(imed, qel, spin_index, find_index, elke, beta2, xi, lambda, chia2, w1, sint1, ms, sr);
        azimuth(cphi1, sphi1);
        mscat(imed, qel, spin_index, find_index, elke, beta2, xi, lambda, chia2, w2, sint2, ms, sr);
        azimuth(cphi2, sphi2);
        R u2 = sint2 * cphi2, v2 = sint2 * sphi2;
        R u2p = w1 * u2 + sint1 * w2;
        R us = u2p * cphi1 - v2 * sphi1;
        R vs = u2p * sphi1 + v2 * cphi1;
        R ws = w1 * w2 - sint1 * u2;
        xi *= (R)2 * xi_corr;
        R eta = rnd();
        R eta1 = (R)0.5 * ((R)1.0 - eta);
        R delta = (R)0.9082483 - ((R)0.1020621 - (R)0.0263747 * gamma) * xi;
        R temp1 = (R)2.0 + tau;
        R temp = ((R)2.0 + tau * temp1) / ((tau + (R)1.0) * temp1);
        temp -= (tau + (R)1.0) / ((tau + (R)2.0) * (chilog * ((R)1.0 + chia2) - (R)1.0));
        temp *= epsilonp;
        temp1 = (R)1.0 - temp;
        delta += (R)0.40824829 * (epsilon * (tau + (R)1.0) /
                                      ((tau + (R)2.0) * (chilog * ((R)1.0 + chia2) - (R)1.0) * (chilog * ((R)1.0 + (R)2.0 * chia2) - (R)2.0)) -
                                  (R)0.25 * (temp * temp));
        R b = eta * delta, c = eta * ((R)1.0 - delta);
        R w1v2 = w1 * v2;
        R ut = b * sint1 * cphi1 + c * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
        R vt = b * sint1 * sphi1 + c * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
        R wt = eta1 * ((R)1.0 + temp) + b * w1 + c * w2 + eta1 * ws * temp1;
        R ustep = tustep * M<R>::sqrt(ut * ut + vt * vt + wt * wt);
        R u0 = p.u, v0 = p.v, w0 = p.w;
        R sint02 = u0 * u0 + v0 * v0;
        if (sint02 > (R)1.0E-20) {
            R sint0 = M<R>::sqrt(sint02), sint0i = (R)1.0 / sint0;
            R cphi0 = sint0i * u0, sphi0 = sint0i * v0;
            u2p = w0 * us + sint0 * ws;
            ws = w0 * ws - sint0 * us;
            us = u2p * cphi0 - vs * sphi0;
            vs = u2p * sphi0 + vs * cphi0;
            u2p = w0 * ut + sint0 * wt;
            wt = w0 * wt - sint0 * ut;
            ut = u2p * cphi0 - vt * sphi0;
            vt = u2p * sphi0 + vt * cphi0;
        } else {
            wt = w0 * wt;
            ws = w0 * ws;
        }
        out.x = p.x + tustep * ut;
        out.y = p.y + tustep * vt;
        out.z = p.z + tustep * wt;
        out.u = us;
        out.v = vs;
        out.w = ws;
        return ustep;
    }

    // CSDA path length inside one PWLF interval (src/ompmc.c:3997-4032)
    __device__ OMC_INL R drange(int imed, int iq, int lelke, R ekei, R ekef, R elkei, R elkef) {
        return drange(cld(&cbin((1 + iq) / 2, imed, lelke)->dedx), ekei, ekef, elkei, elkef);
    }
    // the same with the interval's stopping-power entry already loaded
    __device__ OMC_INL R drange(const PW<R> &q, R ekei, R ekef, R elkei, R elkef) {
        R fedep = (R)1.0 - ekef / ekei;
        R elktmp = (R)0.5 * (elkei + elkef + (R)0.25 * fedep * fedep * ((R)1.0 + fedep * ((R)1.0 + (R)0.875 * fedep)));
        R dedxmid = (R)1.0 / pwlf(q, elktmp);
        R aux = q.c1 * dedxmid;
        aux = aux * ((R)1.0 + (R)2.0 * aux) * fedep * fedep / ((R)6.0 * ((R)2.0 - fedep) * ((R)2.0 - fedep));
        return fedep * ekei * dedxmid * ((R)1.0 + aux);
    }

    // sub-threshold energy loss over a path tustep (src/ompmc.c:4034-4126)
    __device__ OMC_INL R eloss(int imed, int iq, R rhof, R tustep, R range, R eke, R elke, int lelke) {
        const CBin<R> *cb = cbin((1 + iq) / 2, imed, lelke);
        return eloss(imed, iq, rhof, tustep, range, eke, elke, lelke, cld(&cb->dedx), cld(&cb->range));
    }
    // the same with the stopping-power entry and the range at the lower edge
    // of the energy bin already loaded
    __device__ OMC_INL R eloss(int imed, int iq, R rhof, R tustep, R range, R eke, R elke, int lelke,
                               const PW<R> &q_lo, R rep_lo) {
        const Med<R> &m = t.med[imed];
        int qel = (1 + iq) / 2;
        const R *rep = t.range_ep + ((size_t)qel * t.nmed + imed) * MXEKE;
        const PW<R> *dx = iq < 0 ? t.ededx : t.pdedx;
        R de, fedep;
        R tuss = range - rep_lo / rhof;
        if (tuss >= tustep) {
            const PW<R> q = q_lo;
            R dedxmid = pwlf(q, elke);
            R aux = q.c1 / dedxmid;
            de = dedxmid * tustep * rhof;
            fedep = de / eke;
            de *= ((R)1.0 - (R)0.5 * fedep * aux * ((R)1.0 - (R)0.333333 * fedep * (aux - (R)1.0 - (R)0.25 * fedep * ((R)2.0 - aux * ((R)4.0 - aux)))));
        } else {
            int l = lelke;
            tuss = (range - tustep) * rhof;
            if (tuss <= (R)0) {
                de = eke - m.te * (R)0.99;
            } else {
                while (l > 0 && tuss < gld(rep + (l))) l -= 1;
                R elktmp = ((R)(l + 2) - m.eke0) / m.eke1;
                R eketmp = gld(t.e_array + (imed * MXEKE + l + 1));
                tuss = (gld(rep + (l + 1)) - tuss) / rhof;
                const PW<R> q = gld(dx + (MXEKE * imed + l));
                R dedxmid = pwlf(q, elktmp);
                R aux = q.c1 / dedxmid;
                de = dedxmid * tuss * rhof;
                fedep = de / eketmp;
                de *= ((R)1.0 - (R)0.5 * fedep * aux * ((R)1.0 - (R)0.333333 * fedep * (aux - (R)1.0 - (R)0.25 * fedep * ((R)2.0 - aux * ((R)4.0 - aux)))));
                de += eke - eketmp;
            }
        }
        return de;
    }

    // Russian roulette of a photon produced by an electron interaction when
    // splitting is on (src/ompmc.c:4167-4182, 4356-4371, 4645-4660): killed
    // (weight and energy 0) with probability 1 - 1/nsplit, else its weight
    // times nsplit.  Called in stack order, one draw per photon.
    __device__ __forceinline__ void roulette(Part<R> &x) {
        if (nsplit > 1 && x.iq == 0) {
            R rn = rnd();
            if (rn * (R)nsplit > (R)1.0) {
                x.wt = (R)0.0;
                x.e = (R)0.0;
            } else {
                x.wt *= (R)nsplit;
            }
        }
    }

    // two-photon annihilation at rest (src/ompmc.c:4129-4185)
    __device__ OMC_INL void rannih(Part<R> &a, Part<R> &b) {
        const R RM = (R)0.5109989461;
        R rn = rnd();
        R costhe = (R)2.0 * rn - (R)1;
        R sinthe = M<R>::sqrt(M<R>::fmax((R)0.0, ((R)1.0 - costhe) * ((R)1.0 + costhe)));
        (void)rnd();                    // drawn and unused, as in the reference
        R cphi, sphi;
        azimuth(cphi, sphi);
        a.e = RM; a.iq = 0;
        a.u = sinthe * cphi; a.v = sinthe * sphi; a.w = costhe;
        b.e = RM; b.iq = 0;
        transfer(b, a);
        b.u = (R)-1.0 * a.u; b.v = (R)-1.0 * a.v; b.w = (R)-1.0 * a.w;
        roulette(a);
        roulette(b);
    }

    // bremsstrahlung (src/ompmc.c:4188-4374)
    __device__ OMC_INL void brems(Part<R> &pe, Part<R> &pg) {
        const R RM = (R)0.5109989461;
        int imed = medium(pe);
        const Med<R> &m = t.med[imed];
        R eie = pe.e;
        int l = eie < (R)50.0 ? 1 : 3, l1 = l + 1;
        R ekin = eie - RM, brmin = m.ap / ekin, waux = -M<R>::log(brmin);
        R a = pe.u, b = pe.v, c = pe.w;
        R sinpsi = a * a + b * b, sindel = (R)0, cosdel = (R)0;
        if (sinpsi > (R)1.0E-20) {
            sinpsi = M<R>::sqrt(sinpsi);
            sindel = b / sinpsi;
            cosdel = a / sinpsi;
        }
        R ztarg = m.zbrang, tteie = eie / RM;
        R beta = M<R>::sqrt((tteie - (R)1.0) * (tteie + (R)1.0)) / tteie;
        R y2max = (R)2.0 * beta * ((R)1.0 + beta) * tteie * tteie;
        R y2maxi = (R)1.0 / y2max;
        R z2max = y2max + (R)1.0;
        R z2maxi = M<R>::sqrt(z2max);
        R aux, br, delta, r7, rejf, ese, esg, phi1, phi2;
        int trials = 0;
        do {
            R r6 = rnd();
            r7 = rnd();
            br = brmin * M<R>::exp(r6 * waux);
            esg = ekin * br;
            ese = eie - esg;
            delta = esg / eie / ese * m.delcm;
            aux = ese / eie;
            if (delta < (R)1.0) {
                phi1 = m.dl1[l - 1] + delta * (m.dl2[l - 1] + delta * m.dl3[l - 1]);
                phi2 = m.dl1[l1 - 1] + delta * (m.dl2[l1 - 1] + delta * m.dl3[l1 - 1]);
            } else {
                phi1 = m.dl4[l - 1] + m.dl5[l - 1] * M<R>::log(delta + m.dl6[l - 1]);
                phi2 = phi1;
            }
            rejf = ((R)1.0 + aux * aux) * phi1 - (R)2.0 * aux * phi2 / (R)3.0;
        } while ((r7 >= rejf) && !stuck(trials));
        pg.e = esg;
        pg.iq = 0;
        transfer(pg, pe);
        R y2tst = (R)0, ttese = ese / RM, esedei = ttese / tteie;
        R rjarg1 = (R)1.0 + esedei * esedei, rjarg2 = rjarg1 + (R)2.0 * esedei, rjarg3;
        R rtest = (R)1.0, rejtst = (R)0.0;
        aux = (R)2.0 * ese * tteie / esg;
        aux = aux * aux;
        R aux1 = aux * ztarg;
        if (aux1 > (R)10.0) rjarg3 = -M<R>::log(m.zbrang) + ((R)1.0 - aux1) / (aux1 * aux1);
        else rjarg3 = M<R>::log(aux / ((R)1.0 + aux1));
        R rejmax = rjarg1 * rjarg3 - rjarg2;
        int trials_ang = 0;
        while (rtest >= rejtst && !stuck(trials_ang)) {
            y2tst = rnd();
            rtest = rnd();
            R aux3 = z2maxi / (y2tst + ((R)1.0 - y2tst) * z2maxi);
            rtest = rtest * aux3 * rejmax;
            y2tst = aux3 * aux3 - (R)1.0;
            R a4 = M<R>::pow4(aux3);
            R y2tst1 = esedei * y2tst / a4;
            R aux4 = (R)16.0 * y2tst1 - rjarg2;
            R aux5 = rjarg1 - (R)4.0 * y2tst1;
            if (rtest < aux4 + aux5 * rjarg3) break;
            R aux2 = M<R>::log(aux / ((R)1.0 + aux1 / a4));
            rejtst = aux4 + aux5 * aux2;
        }
        R costhe = (R)1.0 - (R)2.0 * y2tst * y2maxi;
        R sinthe = M<R>::sqrt(M<R>::fmax((R)0.0, ((R)1.0 - costhe * costhe)));
        R cphi, sphi;
        azimuth(cphi, sphi);
        if (sinpsi >= (R)1.0E-10) {
            R us = sinthe * cphi, vs = sinthe * sphi;
            pg.u = c * cosdel * us - sindel * vs + a * costhe;
            pg.v = c * sindel * us + cosdel * vs + b * costhe;
            pg.w = c * costhe - sinpsi * us;
        } else {
            pg.u = sinthe * cphi;
            pg.v = sinthe * sphi;
            pg.w = c * costhe;
        }
        pe.e = ese;
        roulette(pe);
        roulette(pg);
    }

    // Moller (src/ompmc.c:4377-4453); nothing happens at or below 2*TE (quirk Q4)
    __device__ OMC_INL int moller(Part<R> &p0, Part<R> &p1) {
        const R RM = (R)0.5109989461;
        const Med<R> &m = t.med[medium(p0)];
        R eie = p0.e, ekin = eie - RM;
        if (ekin <= (R)2.0 * m.te) return 0;
        R t0 = ekin / RM, e0 = t0 + (R)1.0, extrae = eie - m.thmoll;
        R g2 = (t0 * t0) / (e0 * e0), g3 = ((R)2.0 * t0 + (R)1.0) / (e0 * e0);
        R br, gmax = ((R)1.0 + (R)1.25 * g2), rejf4, r28;
        int trials = 0;
        do {
            R r27 = rnd();
            br = m.te / (ekin - extrae * r27);
            R r = br / ((R)1.0 - br);
            r28 = rnd();
            rejf4 = ((R)1.0 + g2 * (br * br) + r * (r - g3));
            r28 *= gmax;
        } while ((r28 > rejf4) && !stuck(trials));
        R ekse2 = br * ekin, ese1 = eie - ekse2, ese2 = ekse2 + RM;
        p0.e = ese1;
        p1.e = ese2;
        R h1 = (eie + RM) / ekin;
        R costh = h1 * (ese1 - RM) / (ese1 + RM);
        Uphi q;
        uphi21(p0, q, M<R>::sqrt(costh), M<R>::sqrt((R)1.0 - costh));
        p1.iq = -1;
        costh = h1 * (ese2 - RM) / (ese2 + RM);
        uphi32(p1, p0, q, M<R>::sqrt(costh), -M<R>::sqrt((R)1.0 - costh));
        return 1;
    }

    // Bhabha (src/ompmc.c:4456-4543)
    __device__ OMC_INL void bhabha(Part<R> &p0, Part<R> &p1) {
        const R RM = (R)0.5109989461;
        const Med<R> &m = t.med[medium(p0)];
        R eip = p0.e, ekin = eip - RM, t0 = ekin / RM, e0 = t0 + (R)1.0;
        R yy = (R)1.0 / (t0 + (R)2.0);
        R beta2 = (e0 * e0 - (R)1.0) / (e0 * e0);
        R ep0 = m.te / ekin, ep0c = (R)1.0 - ep0, yp = (R)1.0 - (R)2.0 * yy;
        R b4 = M<R>::pow3(yp);
        R b3 = b4 + yp * yp;
        R b2 = yp * ((R)3.0 + yy * yy);
        R b1 = (R)2.0 - yy * yy;
        R br, rejf2, r4;
        int trials = 0;
        do {
            R r3 = rnd();
            br = ep0 / ((R)1.0 - ep0c * r3);
            r4 = rnd();
            rejf2 = ((R)1.0 - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4))));
        } while ((r4 > rejf2) && !stuck(trials));
        p1.iq = 1;
        if (br < (R)0.5) {
            p1.iq = -1;
        } else {
            p0.iq = -1;
            p1.iq = 1;
            br = (R)1.0 - br;
        }
        br = M<R>::fmax(br, (R)0.0);
        R ekse2 = br * ekin, ese1 = eip - ekse2, ese2 = ekse2 + RM;
        p0.e = ese1;
        p1.e = ese2;
        R h1 = (eip + RM) / ekin;
        R costh = M<R>::fmin((R)1.0, h1 * (ese1 - RM) / (ese1 + RM));
        Uphi q;
        uphi21(p0, q, M<R>::sqrt(costh), M<R>::sqrt((R)1.0 - costh));
        costh = h1 * (ese2 - RM) / (ese2 + RM);
        uphi32(p1, p0, q, M<R>::sqrt(costh), -M<R>::sqrt((R)1.0 - costh));
    }

    // two-photon annihilation in flight (src/ompmc.c:4546-4663)
    __device__ OMC_INL void annih(Part<R> &p0, Part<R> &p1) {
        const R RM = (R)0.5109989461;
        R avip = p0.e + RM, a = avip / RM;
        R g = a - (R)1.0, tt = g - (R)1.0, pp = M<R>::sqrt(a * tt);
        R pot = pp / tt, ep0 = (R)1.0 / (a + pp), wsamp = M<R>::log(((R)1.0 - ep0) / ep0);
        R aa = p0.u, bb = p0.v, cc = p0.w;
        R sinpsi = aa * aa + bb * bb, sindel = (R)0, cosdel = (R)0;
        if (sinpsi > (R)1.0E-20) {
            sinpsi = M<R>::sqrt(sinpsi);
            sindel = bb / sinpsi;
            cosdel = aa / sinpsi;
        }
        R ep, rejf, r2;
        int trials = 0;
        do {
            R r1 = rnd();
            ep = ep0 * M<R>::exp(r1 * wsamp);
            r2 = rnd();
            R d = ep * a - (R)1.0;
            rejf = (R)1.0 - d * d / (ep * (a * a - (R)2.0));
        } while ((r2 > rejf) && !stuck(trials));
        R esg1 = avip * ep;
        p0.e = esg1;
        p0.iq = 0;
        R costhe = M<R>::fmin((R)1.0, (esg1 - RM) * pot / esg1);
        R sinthe = M<R>::sqrt((R)1.0 - costhe * costhe);
        R sphi, cphi;
        azimuth(cphi, sphi);
        R us, vs;
        if (sinpsi >= (R)1.0E-10) {
            us = sinthe * cphi;
            vs = sinthe * sphi;
            p0.u = cc * cosdel * us - sindel * vs + aa * costhe;
            p0.v = cc * sindel * us + cosdel * vs + bb * costhe;
            p0.w = cc * costhe - sinpsi * us;
        } else {
            p0.u = sinthe * cphi;
            p0.v = sinthe * sphi;
            p0.w = cc * costhe;
        }
        R esg2 = avip - esg1;
        p1.e = esg2;
        p1.iq = 0;
        transfer(p1, p0);
        costhe = M<R>::fmin((R)1.0, (esg2 - RM) * pot / esg2);
        sinthe = -M<R>::sqrt((R)1.0 - costhe * costhe);
        if (sinpsi >= (R)1.0E-10) {
            us = sinthe * cphi;
            vs = sinthe * sphi;
            p1.u = cc * cosdel * us - sindel * vs + aa * costhe;
            p1.v = cc * sindel * us + cosdel * vs + bb * costhe;
            p1.w = cc * costhe - sinpsi * us;
        } else {
            p1.u = sinthe * cphi;
            p1.v = sinthe * sphi;
            p1.w = cc * costhe;
        }
        roulette(p0);
        roulette(p1);
    }

    // electron below cut-off or discarded: deposit, annihilate a positron
    // (src/ompmc.c:4683-4705, 5079-5106)
    __device__ OMC_INL int electron_stop(Part<R> &p, Part<R> &q, int iq, R eie, bool discard) {
        const R RM = (R)0.5109989461;
        R edep = (discard && iq > 0) ? p.e + RM : p.e - RM;
        ausgab(p, edep);
        if (iq > 0 && edep < eie) {
            rannih(p, q);
            return 1;
        }
        return -1;
    }

    // Resumable state of an electron inside the reference's ustep loop: the
    // remaining number of mean free paths and the cross section of the
    // fictitious method.  phase 0: enter electron() afresh; 1: resume;
    // 2: resume with the distance to the interaction still to be computed
    // (after a step taken with a zero cross section).
    struct EState { R demfp, sig0; int phase; };

    // condensed-history electron step (src/ompmc.c:4666-5451).  The variable
    // scopes of the reference are kept: the accumulated path to the next
    // interaction (total_tstep) restarts at zero in every ustep iteration, so
    // a step that does not reach the interaction site is followed by a
    // zero-length step that resamples the distance.
    __device__ OMC_INL int electron(Part<R> &p, Part<R> &q) {
        EState es = {(R)0, (R)0, 0};
        return electron_run(p, q, es, 0x7fffffff);
    }

    // The same routine, yielding after `budget` ustep iterations: returns 0
    // with es.phase != 0 and the loop state saved; called again with that state
    // it continues exactly where the uninterrupted routine would.  Every other
    // exit leaves es.phase = 0.
    __device__ __forceinline__ int electron_run(Part<R> &p, Part<R> &q, EState &es, int budget) {
        const R RM = (R)0.5109989461;
        if (count) cnt.nelectron++;
        int imed = medium(p);
        R rhof = rhof_of(p);
        R rhof_voxel = rhof;             // density of the current voxel, loaded with its medium at a move
        R eie = p.e;
        int iq = p.iq, qel = (1 + iq) / 2, medold = imed;
        bool resume = es.phase != 0;
        bool resume_compute = es.phase == 2;
        es.phase = 0;
        if (!resume && eie <= ecut_of(imed)) return electron_stop(p, q, iq, eie, false);
        R elke = (R)0;
        int lelke = 0;
        if (resume) {
            elke = M<R>::log(eie - RM);
            lelke = pwlf_i(elke, t.med[imed].eke1, t.med[imed].eke0) - 1;
        }
        R sigratio = (R)0, rfict = (R)0;
        bool first = true;
        int iters = 0;
        do {
            // Every pass of this loop and of the step loop below counts
            // against the budget; the start of a later pass is the state of a
            // fresh call (everything read below is recomputed first), so the
            // routine may yield here with phase 0.  A call that makes no
            // progress for 2^24 passes is abandoned: the particle is dropped
            // and `aborted` reports it.
            if (!first) {
                if (--budget <= 0) {
                    es.phase = 0;
                    return 0;
                }
                if (++iters > (1 << 24) || aborted) {     // or a sampler of the last pass was stuck
                    if (!aborted) {
                        aborted = 2;
                        abort_site = __LINE__;
                    }
                    p.e = eie;
                    lost += charged_energy(p);
                    return -1;
                }
            }
            first = false;
            bool compute_tstep = resume ? resume_compute : true;
            R eke = eie - RM, demfp = (R)0, sig0 = (R)0, ustep = (R)0;
            if (resume) {
                demfp = es.demfp;
                sig0 = es.sig0;
                resume = false;
            } else if (imed != -1) {
                const Med<R> &m = t.med[imed];
                R rn = rnd();
                if (rn == (R)0.0) rn = (R)1.0E-30;
                demfp = M<R>::fmax(-M<R>::log(rn), (R)1.0E-5);
                elke = M<R>::log(eke);
                lelke = pwlf_i(elke, m.eke1, m.eke0) - 1;
                if (m.sig_mono[qel]) {
                    const CBin<R> *cb = cbin(qel, imed, lelke);
                    sig0 = cpw(&cb->sig, elke) / cpw(&cb->dedx, elke);
                } else {
                    sig0 = iq < 0 ? m.esig_e : m.psig_e;
                }
            }
            do {
                bool call_howfar = false, do_single = false, called_msdist = false;
                R tstep, tustep, total_tstep = (R)0, range = (R)0, tvstep = (R)0, de = (R)0, tperp = (R)0;
                PW<R> q_dedx = {(R)0, (R)0};     // stopping power and range at this step's energy bin
                R rep_lo = (R)0;                 // (loaded below for a step in a medium)
                Part<R> fin;
                if (imed == -1) {
                    tstep = (R)10.0E8;
                    ustep = tstep;
                    tustep = ustep;
                    call_howfar = true;
                } else {
                    const Med<R> &m = t.med[imed];
                    const R *rep = t.range_ep + ((size_t)qel * t.nmed + imed) * MXEKE;
                    // this step's table entries at (medium, energy bin), loaded
                    // together ahead of the branches below, whose outcome they
                    // do not depend on (same values as loading them where used)
                    const CBin<R> *cb = cbin(qel, imed, lelke);
                    q_dedx = cld(&cb->dedx);
                    const PW<R> q_tmxs = cld(&cb->tmxs);
                    const PW<R> q_eta = cld(&cb->eta);
                    const PW<R> q_msc = cld(&cb->blcc);
                    const R e_lo = cld(&cb->e_lo);
                    rep_lo = cld(&cb->range);
                    rhof = rhof_voxel;           // the reference reloads the density here, not at the move
                    if (sig0 <= (R)0.0) {
                        tstep = (R)10.0E8;
                        sig0 = (R)1.0E-15;
                    } else {
                        if (compute_tstep) {
                            R total_de = demfp / sig0;
                            R ekef = eke - total_de;
                            if (ekef <= m.e_array0) {
                                tstep = (R)10.0E8;
                            } else {
                                R elkef = M<R>::log(ekef);
                                int lelkef = pwlf_i(elkef, m.eke1, m.eke0) - 1;
                                if (lelkef == lelke) {
                                    tstep = drange(q_dedx, eke, ekef, elke, elkef);
                                } else {
                                    R ekei = e_lo;
                                    R elkei = ((R)(lelke + 1) - m.eke0) / m.eke1;
                                    R tuss = drange(q_dedx, eke, ekei, elke, elkei);
                                    ekei = gld(t.e_array + (imed * MXEKE + lelkef + 1));
                                    elkei = ((R)(lelkef + 2) - m.eke0) / m.eke1;
                                    tstep = drange(imed, iq, lelkef, ekei, ekef, elkei, elkef);
                                    tstep += tuss + rep_lo - gld(rep + (lelkef + 1));
                                }
                            }
                            total_tstep = tstep;
                            compute_tstep = false;
                        }
                        tstep = total_tstep / rhof;
                    }
                    R dedx0 = pwlf(q_dedx, elke);
                    R dedx = rhof * dedx0;
                    R tmxs = pwlf(q_tmxs, elke);
                    tmxs /= rhof;
                    R ekei = e_lo;
                    R elkei = ((R)(lelke + 1) - m.eke0) / m.eke1;
                    range = drange(q_dedx, eke, ekei, elke, elkei);
                    range += rep_lo;
                    range /= rhof;
                    tustep = M<R>::fmin(M<R>::fmin(tstep, tmxs), range);
                    tperp = hownear(p);
                    R blccl = rhof * m.blcc, xccl = rhof * m.xcc;
                    R p2 = eke * (eke + (R)2.0 * RM);
                    R beta2 = p2 / (p2 + RM * RM);
                    R etap = pwlf(q_eta, elke);
                    R ms_corr = pwlf(q_msc, elke);
                    blccl = blccl / etap / ((R)1.0 + (R)0.25 * etap * xccl / blccl / p2) * ms_corr;
                    R ssmfp = beta2 / blccl;
                    R skindepth = (R)3 * ssmfp;
                    tustep = M<R>::fmin(tustep, M<R>::fmax(tperp, skindepth));
                    if ((tustep <= tperp) && (tustep > skindepth)) {
                        called_msdist = true;
                        de = eloss(imed, iq, rhof, tustep, range, eke, elke, lelke, q_dedx, rep_lo);
                        tvstep = tustep;
                        ustep = msdist(p, imed, iq, rhof, de, tustep, eke, fin);
                    } else {
                        R rn = rnd();
                        if (rn < (R)1.0E-30) rn = (R)1.0E-30;
                        R lambda = (R)(-1.0) * M<R>::log((R)1.0 - rn);
                        R lambda_max = (R)0.5 * blccl * RM / dedx;
                        R ek1 = eke / RM + (R)1.0;
                        lambda_max *= ek1 * ek1 * ek1;
                        if (lambda >= (R)0.0 && lambda_max > (R)0.0) {
                            R tuss;
                            if (lambda < lambda_max) tuss = lambda * ssmfp * ((R)1.0 - (R)0.5 * lambda / lambda_max);
                            else tuss = (R)0.5 * lambda * ssmfp;
                            if (tuss < tustep) {
                                tustep = tuss;
                                do_single = true;
                            }
                        } else {
                            p.e = eie;
                            lost += charged_energy(p);
                            return -1;       // reference warns and drops (quirk Q6)
                        }
                        ustep = tustep;
                        call_howfar = !(ustep < tperp);
                    }
                }
                int mv = 0;
                bool leaves = false;
                if (call_howfar) {
                    if (p.ix < 0) {
                        if (count) cnt.nhowfar++;
                        return electron_stop(p, q, iq, eie, true);
                    }
                    mv = howfar(p, ustep, leaves);
                }
                if (ustep < (R)0) ustep = (R)0.0;
                R vstep;
                if (ustep == (R)0.0 || imed == -1) {
                    if (ustep != (R)0.0) {
                        vstep = ustep;
                        p.x += p.u * vstep;
                        p.y += p.v * vstep;
                        p.z += p.w * vstep;
                    }
                    if (mv != 0) {
                        apply_move(p, mv, leaves);
                        imed = medium(p);
                        rhof_voxel = rhof_of(p);    // taken into rhof where the reference reloads it
                    }
                    if (eie <= ecut_of(imed)) return electron_stop(p, q, iq, eie, false);
                    break;
                }
                vstep = ustep;
                if (call_howfar) {
                    tvstep = vstep;
                    if (tvstep != tustep) do_single = false;
                    de = eloss(imed, iq, rhof, tvstep, range, eke, elke, lelke, q_dedx, rep_lo);
                } else {
                    tvstep = tustep;
                    if (!called_msdist) de = eloss(imed, iq, rhof, tvstep, range, eke, elke, lelke, q_dedx, rep_lo);
                }
                R edep = de;
                R ekef = eke - de;
                if (!called_msdist) {
                    R costhe = (R)1.0, sinthe = (R)0.0;
                    if (do_single) {
                        const Med<R> &m = t.med[imed];
                        R ekems = M<R>::fmax(ekef, g.ecut[imed] - RM);
                        R p2 = ekems * (ekems + (R)2.0 * RM);
                        R beta2 = p2 / (p2 + RM * RM);
                        R chia2 = m.xcc / ((R)4.0 * m.blcc * p2);
                        R elkems = M<R>::log(ekems);
                        int lelkems = pwlf_i(elkems, m.eke1, m.eke0) - 1;
                        R etap = iq < 0 ? pwlf(t.etae + (MXEKE * imed + lelkems), elkems)
                                        : pwlf(t.etap + (MXEKE * imed + lelkems), elkems);
                        chia2 *= etap;
                        sscat(imed, qel, chia2, elkems, beta2, costhe, sinthe);
                    }
                    fin.x = p.x + p.u * vstep;
                    fin.y = p.y + p.v * vstep;
                    fin.z = p.z + p.w * vstep;
                    if (do_single) {
                        Uphi qf;
                        uphi21(p, qf, costhe, sinthe);
                    }
                    fin.u = p.u; fin.v = p.v; fin.w = p.w;
                }
                ausgab(p, edep);
                p.x = fin.x; p.y = fin.y; p.z = fin.z;
                p.u = fin.u; p.v = fin.v; p.w = fin.w;
                eie -= edep;
                p.e = eie;
                if (mv == 0 && eie <= ecut_of(imed)) return electron_stop(p, q, iq, eie, false);
                medold = imed;
                if (imed != -1) {
                    const Med<R> &m = t.med[imed];
                    eke = eie - RM;
                    elke = M<R>::log(eke);
                    lelke = pwlf_i(elke, m.eke1, m.eke0) - 1;
                }
                if (mv != 0) {
                    apply_move(p, mv, leaves);
                    imed = medium(p);
                    rhof_voxel = rhof_of(p);    // taken into rhof where the reference reloads it
                }
                if (eie <= ecut_of(imed)) return electron_stop(p, q, iq, eie, false);
                if (imed != medold) break;
                demfp -= de * sig0;
                R tt0 = total_tstep;
                total_tstep -= tvstep * rhof;
                if (total_tstep < M<R>::reach_tol(tt0)) demfp = (R)0.0;
                if (++iters > (1 << 24) || aborted) {     // or a sampler of this step was stuck
                    if (!aborted) {
                        aborted = 3;
                        abort_site = __LINE__;
                    }
                    lost += charged_energy(p);
                    return -1;
                }
                if (demfp >= (R)1.0E-5 && --budget <= 0) {
                    es.demfp = demfp;
                    es.sig0 = sig0;
                    es.phase = compute_tstep ? 2 : 1;
                    return 0;
                }
            } while (demfp >= (R)1.0E-5);
            if ((imed != medold) || (ustep == (R)0.0) || (imed == -1)) continue;
            const CBin<R> *cbf = cbin(qel, imed, lelke);
            R sigf = cpw(&cbf->sig, elke) / cpw(&cbf->dedx, elke);
            sigratio = sigf / sig0;
            rfict = rnd();
        } while (rfict >= sigratio);
        const CBin<R> *cbi = cbin(qel, imed, lelke);
        if (iq < 0) {
            R ebr1 = cpw(&cbi->br1, elke);
            R rn = rnd();
            if (rn <= ebr1) {
                brems(p, q);
                return 1;
            }
            if (p.e <= t.med[imed].thmoll) {
                if (ebr1 <= (R)0) return 0;
                brems(p, q);
                return 1;
            }
            return moller(p, q);
        }
        R pbr1 = cpw(&cbi->br1, elke);
        R rn = rnd();
        if (rn < pbr1) {
            brems(p, q);
            return 1;
        }
        R pbr2 = cpw(&cbi->br2, elke);
        if (rn < pbr2) bhabha(p, q);
        else annih(p, q);
        return 1;
    }

    // one pass of shower() (src/ompmc.c:5454-5465) on the particle on top:
    // returns the change of stack depth (-1, 0, +1); q receives the new top
    __device__ __forceinline__ int step(Part<R> &p, Part<R> &q) {
        return p.iq == 0 ? photon(p, q) : electron(p, q);
    }
};

}  // namespace omc
