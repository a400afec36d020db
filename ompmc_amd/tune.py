"""Autotuner of the event-queue form's launch parameters.

The batched form's queue knobs (``omc_run_opts``) are the particles in
flight per round (``event_target``), the electron steps between lane refills
(``electron_budget``) and the refills an electron may hold its lane before it
is requeued (``electron_turns``).  ``tune_dij`` times candidate settings on a
sample of the caller's own beamlets, at the caller's photon splitting factor,
and returns the fastest as keyword arguments for ``Context.run_dij``.  Results are cached per workload shape in
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

TARGETS = (524288, 1048576, 2097152)
BUDGETS = (8, 12)
TURNS = (12, 24, 48)


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


def tune_dij(ctx, phantom, spectrum, bix, nbatch=10, precision=FP32, sample=200, targets=None, budgets=BUDGETS,
             turns=TURNS, use_cache=True, verbose=False, nsplit=1):
    """Returns the fastest {event_target, electron_budget, electron_turns}
    (plus form, precision and nsplit) for ``Context.run_dij`` on a sample of
    ``bix``.  Every candidate computes the same matrix (the random streams do
    not depend on these settings); only the time differs.  The default
    targets are four times larger with photon splitting, which carries
    nsplit times more electrons per history."""
    if targets is None:
        targets = tuple(t * (4 if nsplit > 1 else 1) for t in TARGETS)
    key = _key(phantom, bix, precision, sample, nsplit)
    cache = _load() if use_cache else {}
    if key in cache:
        return dict(cache[key]["opts"])
    sub = _sample(bix, sample)
    ctx.run_dij(phantom, spectrum, _sample(bix, 4), nbatch=nbatch, precision=precision, form=FORM_EVENT,
                nsplit=nsplit)
    results = []
    for tgt in targets:
        for bud in budgets:
            for tu in turns:
                t0 = time.perf_counter()
                _, _, _, st = ctx.run_dij(phantom, spectrum, sub, nbatch=nbatch, precision=precision,
                                          form=FORM_EVENT, event_target=tgt, electron_budget=bud, electron_turns=tu,
                                          nsplit=nsplit)
                rate = st.histories / (time.perf_counter() - t0)
                results.append((rate, tgt, bud, tu))
                if verbose:
                    print(f"tune: target {tgt} budget {bud} turns {tu}: {rate:.4g} histories/s")
    rate, tgt, bud, tu = max(results)
    opts = dict(form=FORM_EVENT, precision=precision, event_target=tgt, electron_budget=bud, electron_turns=tu,
                nsplit=nsplit)
    if use_cache:
        cache[key] = {"opts": opts, "histories_per_s": rate,
                      "table": [dict(rate=r, event_target=t, electron_budget=b, electron_turns=u)
                                for r, t, b, u in results]}
        _store(cache)
    return opts
