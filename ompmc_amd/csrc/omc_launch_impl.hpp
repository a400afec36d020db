// Definitions of the precision-templated launchers; included once per
// precision by omc_k32.hip and omc_k64.hip, which instantiate them.
#pragma once
#include "omc_launch.hpp"

#include <cstring>

namespace omc {

template <class R> hipError_t launch_history(const HistArgs<R> &a, int rng, bool count, int grid, int block,
                                             hipStream_t st) {
    if (a.nsplit > 1 && a.stack) {
        if (rng == RNG_REPLAY) {
            if (count) hipLaunchKernelGGL((k_history_split<R, Replay, true, true>), dim3(grid), dim3(block), 0, st, a);
            else hipLaunchKernelGGL((k_history_split<R, Replay, false, true>), dim3(grid), dim3(block), 0, st, a);
        } else {
            if (count) hipLaunchKernelGGL((k_history_split<R, Philox, true, true>), dim3(grid), dim3(block), 0, st, a);
            else hipLaunchKernelGGL((k_history_split<R, Philox, false, true>), dim3(grid), dim3(block), 0, st, a);
        }
        return hipGetLastError();
    }
    if (a.nsplit > 1) {
        if (rng == RNG_REPLAY) {
            if (count) hipLaunchKernelGGL((k_history_split<R, Replay, true, false>), dim3(grid), dim3(block), 0, st, a);
            else hipLaunchKernelGGL((k_history_split<R, Replay, false, false>), dim3(grid), dim3(block), 0, st, a);
        } else {
            if (count) hipLaunchKernelGGL((k_history_split<R, Philox, true, false>), dim3(grid), dim3(block), 0, st, a);
            else hipLaunchKernelGGL((k_history_split<R, Philox, false, false>), dim3(grid), dim3(block), 0, st, a);
        }
        return hipGetLastError();
    }
    if (rng == RNG_REPLAY) {
        if (count) hipLaunchKernelGGL((k_history<R, Replay, true>), dim3(grid), dim3(block), 0, st, a);
        else hipLaunchKernelGGL((k_history<R, Replay, false>), dim3(grid), dim3(block), 0, st, a);
    } else {
        if (count) hipLaunchKernelGGL((k_history<R, Philox, true>), dim3(grid), dim3(block), 0, st, a);
        else hipLaunchKernelGGL((k_history<R, Philox, false>), dim3(grid), dim3(block), 0, st, a);
    }
    return hipGetLastError();
}

template <class R> hipError_t launch_sampler(const SampArgs<R> &a, hipStream_t st) {
    int block = 64;
    int grid = (a.ncase + block - 1) / block;
    hipLaunchKernelGGL((k_sampler<R>), dim3(grid), dim3(block), 0, st, a);
    return hipGetLastError();
}

// One kernel of round r: its function, shape and by-value arguments.
// Part 0 samples the round's primaries, 1 transports photons, 2 charged
// particles, 3 ends the round.  Arguments depend on r only through its
// parity, which selects the input and output pools.
template <class R> struct RoundCall {
    const void *fn = nullptr;
    dim3 grid, block;
    SrcArgs<R> src;
    EvArgs<R> ev;
    EvCtl *ctl = nullptr;
    int par = 0;
    void *args[2] = {nullptr, nullptr};
    int part = 0;
    unsigned shmem = 0;                  // dynamic LDS bytes per workgroup
    // kernelParams of a graph node: pointers to the argument values above
    void **bind() {
        if (part == 0) args[0] = &src;
        else if (part == 3) { args[0] = &ctl; args[1] = &par; }
        else args[0] = &ev;
        return args;
    }
};

template <class R> void round_call(const EvLaunch<R> &L, long r, int part, bool count, RoundCall<R> &c) {
    const int par = (int)(r & 1), nx = 1 - par;
    c.part = part;
    c.block = dim3(256);
    if (part == 0) {
        SrcArgs<R> &s = c.src;
        s = L.src;
        s.ctl = L.ctl;
        s.ebal = L.ebal;
        s.out = L.pool[par][L.src.dos.charge != 0 && !L.src.use_beamlets ? 1 : 0];
        c.fn = reinterpret_cast<const void *>(&k_source<R>);
        c.grid = dim3(L.grid_source);
    } else if (part == 1 || part == 2) {
        EvArgs<R> &a = c.ev;
        a.t = L.t;
        a.g = L.g;
        a.seed = L.seed;
        a.outp = L.pool[nx][0];
        a.oute = L.pool[nx][1];
        a.n_out = &L.ctl->cnt[nx][0];
        a.cap = L.cap;
        a.endep = L.endep;
        a.slot_stride = L.slot_stride;
        a.touched = L.touched;
        a.touch_stride = L.touch_stride;
        a.ebal = L.ebal;
        a.counters = L.counters ? L.counters + (part == 1 ? 0 : 8) : nullptr;
        a.err = &L.ctl->err;
        a.ctl = L.ctl;
        a.budget = L.budget;
        a.turns = L.turns;
        a.max_age = L.max_age;
        a.delta = L.delta;
        a.nsplit = L.nsplit > 1 ? L.nsplit : 1;
        a.ebin = L.ebin;
        a.eb_words = part == 2 ? L.eb_words : 0;
        a.tlist = part == 2 ? L.tlist : 0;
        c.shmem = (unsigned)a.eb_words * 16u + (unsigned)a.tlist * c.block.x * 16u;
        a.in = L.pool[par][part - 1];
        a.n_in = &L.ctl->cnt[par][part - 1];
        a.claim = &L.ctl->claim[par][part - 1];
        if (part == 1 && a.nsplit > 1)
            c.fn = count ? reinterpret_cast<const void *>(&k_event_split<R, true>)
                         : reinterpret_cast<const void *>(&k_event_split<R, false>);
        else if (part == 1)
            c.fn = count ? reinterpret_cast<const void *>(&k_event<R, true, true>)
                         : reinterpret_cast<const void *>(&k_event<R, true, false>);
        else
            c.fn = count ? reinterpret_cast<const void *>(&k_event<R, false, true>)
                         : reinterpret_cast<const void *>(&k_event<R, false, false>);
        c.grid = dim3(part == 2 ? L.grid_charged : L.grid_event);
    } else {
        c.ctl = L.ctl;
        c.par = par;
        c.fn = reinterpret_cast<const void *>(&k_round_end);
        c.grid = dim3(1);
        c.block = dim3(64);
    }
}

template <class R> hipError_t launch_round_part(const EvLaunch<R> &L, long r, int part, bool count, hipStream_t st) {
    RoundCall<R> c;
    round_call(L, r, part, count, c);
    if (part == 0) {
        hipLaunchKernelGGL((k_source<R>), c.grid, c.block, 0, st, c.src);
    } else if (part == 1 || part == 2) {
        if (c.fn == reinterpret_cast<const void *>(&k_event_split<R, true>))
            hipLaunchKernelGGL((k_event_split<R, true>), c.grid, c.block, 0, st, c.ev);
        else if (c.fn == reinterpret_cast<const void *>(&k_event_split<R, false>))
            hipLaunchKernelGGL((k_event_split<R, false>), c.grid, c.block, 0, st, c.ev);
        else if (c.fn == reinterpret_cast<const void *>(&k_event<R, true, true>))
            hipLaunchKernelGGL((k_event<R, true, true>), c.grid, c.block, 0, st, c.ev);
        else if (c.fn == reinterpret_cast<const void *>(&k_event<R, true, false>))
            hipLaunchKernelGGL((k_event<R, true, false>), c.grid, c.block, 0, st, c.ev);
        else if (c.fn == reinterpret_cast<const void *>(&k_event<R, false, true>))
            hipLaunchKernelGGL((k_event<R, false, true>), c.grid, c.block, c.shmem, st, c.ev);
        else
            hipLaunchKernelGGL((k_event<R, false, false>), c.grid, c.block, c.shmem, st, c.ev);
    } else {
        hipLaunchKernelGGL(k_round_end, c.grid, c.block, 0, st, c.ctl, c.par);
    }
    return hipGetLastError();
}

// A graph of rounds r0 .. r0+nround-1, built node by node (no stream
// capture, so enqueues and allocations of other threads cannot disturb it):
// each kernel depends on the one before, as on a stream.
template <class R> hipError_t build_rounds_graph(const EvLaunch<R> &L, long r0, int nround, bool count,
                                                 hipGraph_t *out) {
    hipGraph_t g = nullptr;
    hipError_t e = hipGraphCreate(&g, 0);
    if (e != hipSuccess) return e;
    hipGraphNode_t prev = nullptr;
    for (long r = r0; r < r0 + nround && e == hipSuccess; r++)
        for (int part = 0; part < 4 && e == hipSuccess; part++) {
            RoundCall<R> c;
            round_call(L, r, part, count, c);
            hipKernelNodeParams p;
            memset(&p, 0, sizeof p);
            p.func = const_cast<void *>(c.fn);
            p.gridDim = c.grid;
            p.blockDim = c.block;
            p.sharedMemBytes = c.shmem;
            p.kernelParams = c.bind();
            p.extra = nullptr;
            hipGraphNode_t n = nullptr;
            e = hipGraphAddKernelNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, &p);
            prev = n;
        }
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return e;
    }
    *out = g;
    return hipSuccess;
}

template <class R> hipError_t launch_rounds(const EvLaunch<R> &L, long r0, int nround, bool count, hipStream_t st) {
    for (long r = r0; r < r0 + nround; r++)
        for (int part = 0; part < 4; part++) {
            hipError_t e = launch_round_part(L, r, part, count, st);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

}  // namespace omc
