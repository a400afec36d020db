"""Autotuner of the event-queue form's launch parameters.

The batched form's queue knobs (``omc_run_opts``) are the particles in
flight per round (``event_target``), the electron steps between lane refills
(``electron_budget``), the refills an electron may hold its lane before it
is requeued (``electron_turns``) and the beamlet groups transported
concurrently (``streams``).  ``tune_dij`` searches them by coordinate
descent, timing settings on a sample of the caller's own beamlets at the
caller's photon splitting factor, and returns the fastest as keyword
arguments for ``Context.run_dij``.  Results are cached per workload shape in
a JSON file, so a later process reuses them.  None of the knobs changes a
result beyond FP64 summation order: every history keeps its random-stream
key.

    from ompmc_amd import tune
    best = tune.tune_dij(ctx, phantom, spectrum, beamlets)
    ctx.run_dij(phantom, spectrum, beamlets, **best)
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import FORM_EVENT, FP32

CACHE = os.environ.get("OMC_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "ompmc_amd_tune.json"))

# Ladders of the knobs.  The search starts at the engine's defaults (marked)
# and moves along each ladder one rung at a time.
TARGETS = (262144, 524288, 1048576, 2097152, 4194304)     # default 4096 per CU = 1048576 on 256 CUs
BUDGETS = (4, 8, 12, 16, 24)                              # default 12
TURNS = (12, 24, 48)                                      # default 24
STREAMS = (1, 2, 3, 4)                                    # default 3 when the run has several groups


def _sample(bix, n):
    from .phantoms import Beamlets
    sel = np.linspace(0, bix.n - 1, min(n, bix.n)).astype(int)
    return Beamlets(bix.ibeam[sel], bix.src, bix.corner[sel], bix.side1[sel], bix.side2[sel],
                    np.asarray(bix.nhist)[sel])


def _key(phantom, bix, precision, sample, nsplit):
    nvox = phantom.nreg - 1
    return (f"dij/{precision}/nvox={nvox}/beamlets={min(sample, bix.n)}/hist={int(np.median(bix.nhist))}"
            + (f"/nsplit={nsplit}" if nsplit > 1 else ""))


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(d, f, indent=1)
    except OSError:
        pass


def _start(ladder, default):
    return ladder.index(default) if default in ladder else len(ladder) // 2


def tune_dij(ctx, phantom, spectrum, bix, nbatch=10, precision=FP32, sample=200, targets=None, budgets=BUDGETS,
             turns=TURNS, streams=STREAMS, use_cache=True, verbose=False, nsplit=1, max_evals=24, margin=0.02):
    """Returns the fastest {event_target, electron_budget, electron_turns,
    streams} found (plus form, precision and nsplit) for ``Context.run_dij``
    on a sample of ``bix``, by coordinate descent: from the defaults, each
    knob in turn tries the next rung up and down its ladder and moves while
    that is faster by more than ``margin`` (measurement noise), until no knob
    moves or ``max_evals`` settings have been timed.  Every setting computes
    the same matrix (the random streams do not depend on these knobs); only
    the time differs.  The default targets are four times larger with
    photon splitting, which carries nsplit times more electrons per
    history."""
    if targets is None:
        targets = tuple(t * (4 if nsplit > 1 else 1) for t in TARGETS)
    key = _key(phantom, bix, precision, sample, nsplit)
    cache = _load() if use_cache else {}
    if key in cache:
        return dict(cache[key]["opts"])
    sub = _sample(bix, sample)
    ctx.run_dij(phantom, spectrum, _sample(bix, 4), nbatch=nbatch, precision=precision, form=FORM_EVENT,
                nsplit=nsplit)
    ladders = [tuple(targets), tuple(budgets), tuple(turns), tuple(streams)]
    names = ("event_target", "electron_budget", "electron_turns", "streams")
    pos = [_start(ladders[0], 1048576 * (4 if nsplit > 1 else 1)), _start(ladders[1], 12), _start(ladders[2], 24),
           _start(ladders[3], 3)]
    timed = {}

    def rate_of(p):
        k = tuple(p)
        if k not in timed:
            kw = {n: ladders[i][p[i]] for i, n in enumerate(names)}
            t0 = time.perf_counter()
            _, _, _, st = ctx.run_dij(phantom, spectrum, sub, nbatch=nbatch, precision=precision, form=FORM_EVENT,
                                      nsplit=nsplit, **kw)
            timed[k] = st.histories / (time.perf_counter() - t0)
            if verbose:
                print(f"tune: {kw}: {timed[k]:.4g} histories/s")
        return timed[k]

    best = rate_of(pos)
    moved = True
    while moved and len(timed) < max_evals:
        moved = False
        for i in range(len(ladders)):
            for step in (1, -1):
                q = list(pos)
                q[i] += step
                if not 0 <= q[i] < len(ladders[i]) or len(timed) >= max_evals:
                    continue
                r = rate_of(q)
                if r > best * (1.0 + margin):
                    pos, best, moved = q, r, True
                    break
    opts = dict(form=FORM_EVENT, precision=precision, nsplit=nsplit,
                **{n: ladders[i][pos[i]] for i, n in enumerate(names)})
    if use_cache:
        cache[key] = {"opts": opts, "histories_per_s": best,
                      "table": [dict(rate=r, **{n: ladders[i][k[i]] for i, n in enumerate(names)})
                                for k, r in timed.items()]}
        _store(cache)
    return opts
