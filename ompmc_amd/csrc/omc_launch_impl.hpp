// Definitions of the precision-templated launchers; included once per
// precision by omc_k32.hip and omc_k64.hip, which instantiate them.
#pragma once
#include "omc_launch.hpp"

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

template <class R> hipError_t launch_round_part(const EvLaunch<R> &L, long r, int part, bool count, hipStream_t st) {
    const int par = (int)(r & 1), nx = 1 - par;
    if (part == 0) {
        SrcArgs<R> s = L.src;
        s.ctl = L.ctl;
        s.ebal = L.ebal;
        s.out = L.pool[par][L.src.dos.charge != 0 && !L.src.use_beamlets ? 1 : 0];
        hipLaunchKernelGGL((k_source<R>), dim3(L.grid_source), dim3(256), 0, st, s);
    } else if (part == 1 || part == 2) {
        EvArgs<R> a;
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
        a.budget = L.budget;
        a.turns = L.turns;
        a.max_age = L.max_age;
        a.delta = L.delta;
        a.nsplit = L.nsplit > 1 ? L.nsplit : 1;
        a.in = L.pool[par][part - 1];
        a.n_in = &L.ctl->cnt[par][part - 1];
        a.claim = &L.ctl->claim[par][part - 1];
        if (part == 1 && a.nsplit > 1) {
            if (count) hipLaunchKernelGGL((k_event_split<R, true>), dim3(L.grid_event), dim3(256), 0, st, a);
            else hipLaunchKernelGGL((k_event_split<R, false>), dim3(L.grid_event), dim3(256), 0, st, a);
        } else if (part == 1) {
            if (count) hipLaunchKernelGGL((k_event<R, true, true>), dim3(L.grid_event), dim3(256), 0, st, a);
            else hipLaunchKernelGGL((k_event<R, true, false>), dim3(L.grid_event), dim3(256), 0, st, a);
        } else {
            if (count) hipLaunchKernelGGL((k_event<R, false, true>), dim3(L.grid_event), dim3(256), 0, st, a);
            else hipLaunchKernelGGL((k_event<R, false, false>), dim3(L.grid_event), dim3(256), 0, st, a);
        }
    } else {
        hipLaunchKernelGGL(k_round_end, dim3(1), dim3(64), 0, st, L.ctl, par);
    }
    return hipGetLastError();
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
