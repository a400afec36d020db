#!/usr/bin/env python3
"""Throughput benchmark of the GPU transport path.

Default workload (c3): the dose-influence matrix of a ~2000-beamlet plan on a
160x160x100 synthetic CT (five beams, 0.5 cm bixels, lognormal histories per
beamlet in 1e3..5e4, synthetic 6 MV spectrum, 10 batches, threshold 1e-4 of
each beamlet's maximum), computed by the batched form (beamlet groups, event
queues, FP32 transport with FP64 tallies).  A "step" is the whole matrix,
phantom and tables resident in HBM.  Beside it the line reports:

* the plain GPU history-by-history form running the reference's loop order
  (one beamlet at a time) on every 20th beamlet, as histories/s;
* the plan-dose statistical uncertainty in voxels above 50% of the maximum,
  and both forms again at the metric's operating point: every beamlet's
  histories scaled so that this uncertainty reaches 1% (beamlet subsets);
* the FP32/FP64 plan-dose difference on every 10th beamlet (three seeds
  per precision, region fixed by an independent FP64 pilot run);
* p50/p99 latency and throughput of a short served request mix;
* the roofline of the dominant transport kernel (algorithmic bytes from
  on-device event counters over its duration, timed with HIP events on its
  stream inside the timed steps) and its measured HBM traffic per launch
  (from the committed rocprofv3 PMC summary under profiles/);
* a CPU baseline: the oracle restatement (OpenMP, the reference's loop) on a
  sample of the same beamlets.

    python bench.py [--steps K] [--warmup W] [--config c3|c2|c1|c5] [--form event|history]
                    [--precision fp32|fp64] [--no-cpu] [--no-serve] [--no-plain]
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {0: "k_history", 1: "k_source", 2: "k_event (photon)", 3: "k_event (charged)", 4: "post-processing"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_cores():
    """Host cores the CPU baseline uses: those this process may run on, at
    most 16 -- the host share that goes with one GPU on the measurement
    hosts, whatever nproc reports there (bench.py reports both)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def kernel_bytes(st, k):
    """Algorithmic bytes of event kernel class k (0 photon, 1 charged):
    72 per transport event (howfar/hownear call: read + write of the 36-byte
    particle state), 5 per voxel entered (density + medium), 16 per FP64
    deposit (read-modify-write)."""
    return 72.0 * (st.ev_howfar[k] + st.ev_hownear[k]) + 5.0 * st.ev_cross[k] + 16.0 * st.ev_dep[k]


def total_bytes(st):
    return 72.0 * (st.n_howfar + st.n_hownear) + 5.0 * st.n_cross + 16.0 * st.n_dep


def subset(bix, every, offset=0):
    from ompmc_amd import phantoms as P
    sel = np.arange(offset, bix.n, every)
    return P.Beamlets(bix.ibeam[sel], bix.src, bix.corner[sel], bix.side1[sel], bix.side2[sel], bix.nhist[sel])


def scaled(bix, k):
    from ompmc_amd import phantoms as P
    return P.Beamlets(bix.ibeam, bix.src, bix.corner, bix.side1, bix.side2,
                      (np.asarray(bix.nhist, np.int64) * k).astype(np.int64))


def c3_config():
    from ompmc_amd import phantoms as P
    import ompmc_amd as omc
    ph = P.synthetic_ct()
    bix = P.beamlets()
    e, w = P.synthetic_6mv()
    return ph, bix, omc.Spectrum.from_histogram(e, w), (e, w)


# ------------------------------------------------------------------ serving mix
_PHANTOM_POOL = {}


def serve_mix(ctx, rate_hist_per_s, nreq=12, load=0.8, seed=20261206, precision=32, max_hist=1e8, max_inflight=4):
    """A Poisson-arrival request mix with the ranges of the serving workload:
    phantoms 64^3 .. 256x256x160 at 0.2-0.5 cm (a pool of at most 8, reused),
    media water/lung/bone/air, 1e5..max_hist histories per request
    (log-uniform), half full-dose requests (10 batches) and half Dij requests
    over 1-9 beams.  The offered load is `load` times the measured batched
    throughput.  200 requests per load reproduce the full workload; the
    default takes 12."""
    import ompmc_amd as omc
    from ompmc_amd import phantoms as P
    rng = np.random.default_rng(seed)
    e, w = P.synthetic_6mv()
    spec = omc.Spectrum.from_histogram(e, w)
    key = (seed, min(nreq, 8))
    if key not in _PHANTOM_POOL:          # the same phantoms at every offered load
        pool = []
        for _ in range(key[1]):
            nx = int(rng.integers(64, 257)); nz = int(rng.integers(64, 161))
            vox = float(rng.uniform(0.2, 0.5))
            pool.append((P.synthetic_ct(nx=nx, ny=nx, nz=nz, vx=vox, vy=vox, vz=vox,
                                        seed=int(rng.integers(1 << 30)), target_radius=min(2.0, 0.2 * nx * vox)),
                         nx, vox))
        _PHANTOM_POOL[key] = pool
    pool = _PHANTOM_POOL[key]
    rng = np.random.default_rng(seed + 1)
    reqs = []
    for i in range(nreq):
        ph, nx, vox = pool[i % len(pool)]
        nh = int(10 ** rng.uniform(5, math.log10(max_hist)))
        if i % 2 == 0:
            half = min(5.0, 0.3 * nx * vox)
            reqs.append(("dose", ph, nh // 10 * 10, dict(collimator=(-half, half, -half, half))))
        else:
            nb = int(rng.integers(1, 10))
            bix = P.beamlets(nbeams=nb, radius=min(2.0, 0.25 * nx * vox), hist_median=3000, hist_sigma_log10=0.3,
                             hist_range=(1000, 20000), seed=int(rng.integers(1 << 30)))
            bix.nhist = np.full(bix.n, max(10, nh // bix.n), np.int64)
            reqs.append(("dij", ph, int(bix.nhist.sum()), dict(bix=bix)))
    mean_h = np.mean([r[2] for r in reqs])
    rate = load * rate_hist_per_s / mean_h               # requests per second
    gaps = rng.exponential(1.0 / rate, nreq)
    srv = omc.Server(ctx, max_inflight=max_inflight)
    ids = []
    t0 = time.perf_counter()
    t_next = t0
    for k, ((kind, ph, nh, kw), gap) in enumerate(zip(reqs, gaps)):
        t_next += gap
        dt = t_next - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        if kind == "dose":
            ids.append(srv.submit_dose(ph, spec, max(1, nh // 10), 10, precision=precision, **kw))
        else:
            ids.append(srv.submit_dij(ph, spec, kw["bix"], nbatch=10, precision=precision))
        if (k + 1) % 20 == 0:
            log(f"served mix at load {load}: {k + 1} of {nreq} submitted")
    res = []
    for k, i in enumerate(ids):
        res.append(srv.wait(i))
        if (k + 1) % 20 == 0:
            log(f"served mix at load {load}: {k + 1} of {nreq} done")
    wall = time.perf_counter() - t0
    srv.close()
    bad = [r for r in res if r["error"]]
    lat = np.array([r["end_t"] - r["submit_t"] for r in res])
    hist = sum(r["histories"] for r in res)
    kinds = np.array([r[0] for r in reqs])
    per_kind = {k: {"p50_s": float(np.percentile(lat[kinds == k], 50)), "p99_s": float(np.percentile(lat[kinds == k], 99))}
                for k in ("dose", "dij")}
    return dict(requests=nreq, failed=len(bad), offered_load=load, p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), histories_per_s=hist / wall, wall_s=wall, by_kind=per_kind,
                histories=float(hist),
                mix=f"Poisson arrivals; phantoms 64^3..256x256x160 at 0.2-0.5 cm; 1e5..{max_hist:.0e} histories "
                    "per request (log-uniform); half dose, half Dij over 1-9 beams")


# ------------------------------------------------------------------ CPU baseline
def cpu_baseline_c3(ph, bix, spec, nbatch, seconds_target=12.0):
    """The oracle restatement of the reference's matRad loop (OpenMP over the
    histories of each batch, as omc_matrad.c:1578-1585) on the first beamlets
    of the plan, on the host cores."""
    from oracle import bindings as OB
    from ompmc_amd import phantoms as P
    cores = cpu_cores()
    with tempfile.TemporaryDirectory() as tmp:
        phf = os.path.join(tmp, "ct.egsphant")
        P.write_egsphant(phf, ph)
        d = P.default_inp(phf, tmp + "/", ncase=1000, nbatch=nbatch)
        s = OB.Session(P.write_inp(os.path.join(tmp, "ct"), d)[:-4])
        bd = dict(ibeam=bix.ibeam, src=bix.src.reshape(-1), corner=bix.corner.reshape(-1),
                  side1=bix.side1.reshape(-1), side2=bix.side2.reshape(-1), gaussian=0, sigma=0.0)
        nper = np.maximum(np.asarray(bix.nhist) // nbatch, 1).astype(np.int32)

        def run(nb, threads, div=1):
            sub = {k: (v[:nb] if k == "ibeam" else v[:3 * nb] if k in ("corner", "side1", "side2") else v)
                   for k, v in bd.items()}
            np_ = np.maximum(nper[:nb] // div, 1).astype(np.int32)
            _, _, secs = s.matrad_omp(spec, sub, np_, nbatch, threads)
            return secs, int(np_.sum()) * nbatch
        secs, nh = run(2, cores)
        nb = int(min(bix.n, max(2, 2 * seconds_target / 3 / max(secs, 1e-3))))
        res = [run(nb, cores) for _ in range(3)]                 # median of three runs
        nh = res[0][1]
        runs = sorted(r[0] for r in res)
        secs = runs[1]
        # one thread on the same beamlets with 1/cores of their histories
        secs1, nh1 = run(nb, 1, cores)
        try:
            xc = cpu_cross_check(cores)
        except Exception as e:          # reported, not required
            xc = {"error": str(e)}
        return dict(value=nh / secs, unit="histories/s", cores=cores, kind="port", cross_check=xc,
                    sample=f"first {nb} beamlets of the plan ({nh} histories, {nbatch} batches), oracle "
                           f"restatement of the reference's matRad loop, OpenMP {cores} threads, median of 3 runs "
                           f"({', '.join(f'{x:.1f}' for x in runs)} s)",
                    one_core={"value": nh1 / secs1, "sample": f"the same {nb} beamlets with 1/{cores} of their "
                                                               f"histories ({nh1}), 1 thread, {secs1:.1f} s"},
                    host={"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads_allowed": cores})


def cpu_cross_check(cores, nhist=1000000):
    """The reference's own omc_dosxyz, compiled from its sources into
    oracle/_ref, against the oracle restatement on the same input: the 60^3
    water cube with 2 MeV photons in a 10x10 cm field, nhist histories in 10
    batches, both with OpenMP on the same `cores` host cores (the reference
    pinned with taskset, since it takes omp_get_num_procs() threads).
    Transport time only: the reference's two "Execution time up to this
    point" lines around its batch loop (omc_dosxyz.c:1234,1267), the
    restatement's own timer around the same loop.  None where the reference
    binary was not built."""
    import re
    import subprocess
    from oracle import bindings as OB
    from ompmc_amd import phantoms as P
    ref = os.path.join(REPO, "oracle", "_ref", "omc_dosxyz")
    if not os.path.exists(ref):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        phf = os.path.join(tmp, "w60.egsphant")
        P.write_egsphant(phf, P.water_cube(60, 0.5))
        d = P.default_inp(phf, tmp + "/", energy=2.0, ncase=nhist, nbatch=10, collimator=(-5, 5, -5, 5))
        stem = P.write_inp(os.path.join(tmp, "w60"), d)[:-4]
        cpus = ",".join(str(c) for c in sorted(os.sched_getaffinity(0))[:cores])
        t_ref = []
        for _ in range(3):
            r = subprocess.run(["taskset", "-c", cpus, ref, "-i", stem, "-o", "ref"], stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=600, cwd=tmp)
            t = [float(x) for x in re.findall(r"Execution time up to this point :\s*([0-9.]+)", r.stdout)]
            if r.returncode != 0 or len(t) < 2:
                return {"error": f"reference run failed ({r.returncode})"}
            t_ref.append(t[1] - t[0])
        s = OB.Session(stem)
        t_orc = sorted(s.dose_omp(nhist // 10, 10, cores)[2] for _ in range(3))
        s.close()
    t_ref.sort()
    return {"reference_seconds": t_ref[1], "oracle_seconds": t_orc[1], "ratio_oracle_over_reference": t_orc[1] / t_ref[1],
            "cores": cores, "sample": f"60^3 water, 2 MeV, 10x10 cm, {nhist} histories in 10 batches, median of 3",
            "reference_histories_per_s": nhist / t_ref[1], "oracle_histories_per_s": nhist / t_orc[1]}


# ------------------------------------------------------------------ c3
def bench_c3(a):
    import ompmc_amd as omc
    prec = omc.FP32 if a.precision == "fp32" else omc.FP64
    form = omc.FORM_EVENT if a.form == "event" else omc.FORM_HISTORY
    nbatch, rel = 10, 1e-4
    t_setup = time.perf_counter()
    ph, bix, spec, _ = c3_config()
    if a.subset > 1:
        bix = subset(bix, a.subset)
    ctx = omc.Context(ph.media)
    gph = ctx.phantom(ph, ecut=0.7, pcut=0.01)
    log(f"c3: {bix.n} beamlets, {int(np.sum(np.maximum(bix.nhist // nbatch, 1)) * nbatch)} histories per step, "
        f"setup {time.perf_counter() - t_setup:.1f} s")
    # the matrices stay in device memory (inputs and outputs resident); the
    # host-output rate, with the device-to-host copy, is reported beside it
    opts = dict(precision=prec, form=form, group=a.group, event_target=a.target, electron_budget=a.budget,
                electron_turns=a.turns, batches_per_job=a.batches_per_job, csc_on_device=True, download=False)
    # event counts of the transport kernels (algorithmic bytes per history)
    sub20 = subset(bix, 20)
    _, _, _, stc = ctx.run_dij(gph, spec, sub20, nbatch=nbatch, rel_threshold=rel, count_events=True, **opts)
    log(f"counts: {stc.histories:.0f} histories, {total_bytes(stc) / stc.histories:.0f} B/history")
    for _ in range(a.warmup):
        ctx.run_dij(gph, spec, bix, nbatch=nbatch, rel_threshold=rel, **opts)
    log("warmup done")
    walls, hist, plan, st_last = [], 0.0, None, None
    ksec = np.zeros(5)
    kbusy = np.zeros(5)
    klaunch = np.zeros(5)
    for k in range(a.steps):
        t0 = time.perf_counter()
        _, _, pl, st = ctx.run_dij(gph, spec, bix, nbatch=nbatch, rel_threshold=rel, want_plan=(k == a.steps - 1),
                                   count_events="time", **opts)
        walls.append(time.perf_counter() - t0)
        hist += st.histories
        ksec += np.array(st.kernel_seconds[:])
        kbusy += np.array(st.kernel_busy_seconds[:])
        klaunch += np.array(st.kernel_launches[:])
        if pl is not None:
            plan = pl
        st_last = st
        log(f"step {k}: {st.histories / walls[-1]:.4g} histories/s, {st.groups} groups of {st.group_size}, "
            f"{st.batches_per_job} batches per job, nnz {st.nnz}")
    wall = sum(walls)
    value = hist / wall
    pd, pv = plan
    sel = pd > 0.5 * pd.max()
    unc50 = float(np.mean(np.sqrt(pv[sel]) / pd[sel]))
    # roofline of the dominant transport kernel
    if form == omc.FORM_EVENT:
        dom = 3 if ksec[3] >= ksec[2] else 2
        bph = kernel_bytes(stc, dom - 2) / stc.histories
    else:
        dom = 0
        bph = total_bytes(stc) / stc.histories
    # launches of the three group streams overlap: the kernel's time on the
    # device is the union of its launches (busy), not their sum; the history
    # form runs one launch at a time
    busy = kbusy[dom] if kbusy[dom] > 0 else ksec[dom]
    achieved = bph * hist / busy / 1e9
    achieved_per_launch = bph * hist / ksec[dom] / 1e9
    traffic = read_traffic(dom, hist, klaunch[dom]) if (a.config == "c3" and form == omc.FORM_EVENT
                                                        and prec == omc.FP32) else None
    out = {
        "metric": "histories/s + p50/p99 request latency vs plain GPU form at <=1% dose unc.",
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == omc.FP32 else "f64",
        "tally_dtype": "f64",
        "data": "synthetic CT, beamlets and 6 MV spectrum (seeded)",
        "config": {"workload": "c3: Dij of a 5-beam plan, 160x160x100 synthetic CT at 0.25x0.25x0.3 cm, "
                               f"{bix.n} beamlets of 0.5 cm, lognormal 1e3..5e4 histories each, 10 batches, "
                               "threshold 1e-4", "beamlets": int(bix.n), "histories_per_step": hist / a.steps,
                   "form": a.form, "precision": a.precision, "group": int(st_last.group_size),
                   "batches_per_job": int(st_last.batches_per_job)},
        "plan_dose_rel_unc_gt50": unc50,
        "nnz": int(st_last.nnz),
        "kernels": {KERNEL_NAMES[k]: {"seconds": float(ksec[k]), "launches": int(klaunch[k])}
                    for k in range(5) if klaunch[k] > 0},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_history": bph,
                     "busy_seconds": float(busy), "summed_launch_seconds": float(ksec[dom]),
                     "per_launch_bytes": bph * hist / max(klaunch[dom], 1),
                     "per_launch_seconds": ksec[dom] / max(klaunch[dom], 1),
                     "per_launch_frac": achieved_per_launch / HBM_PEAK_GBS,
                     "what": "algorithmic bytes of the kernel's histories over the time the device ran at least "
                             "one of its launches (union over the concurrent streams); per_launch_frac divides by "
                             "the summed launch durations instead"},
    }
    valu = read_valu(dom, value)
    if valu:
        out["valu_roofline"] = valu
    # secondary roofline: FP64 tally atomics.  Every deposit call counts as one
    # atomic (an upper bound: a lane folds repeated deposits to one region into
    # one add), against the device's measured rate of scattered FP64 atomic adds
    ndep = stc.n_dep / stc.histories
    peak_small = omc.atomic_probe(1 << 18, 1 << 26)
    peak_large = omc.atomic_probe(1 << 27, 1 << 27)
    out["atomic_roofline"] = {"bound": "fp64 atomics", "deposits_per_history": ndep, "achieved": ndep * value,
                              "peak": peak_large, "unit": "atomics/s", "frac": ndep * value / peak_large,
                              "probe": {"table_2MiB": peak_small, "table_1GiB": peak_large}}
    log(f"atomics: {ndep:.1f} deposits/history -> {ndep * value:.3g}/s against {peak_large:.3g}/s measured")
    # the same step with the matrix returned in host memory, as the mex
    # interface receives it (device-to-host copy and column stitching)
    ho = dict(opts, csc_on_device=False, download=True)
    rates = []
    for _ in range(2):          # the first call also pins the pooled host blocks it returns
        t0 = time.perf_counter()
        _, _, _, sth = ctx.run_dij(gph, spec, bix, nbatch=nbatch, rel_threshold=rel, **ho)
        rates.append(sth.histories / (time.perf_counter() - t0))
    out["host_output"] = {"histories_per_s": rates[1], "first_call_histories_per_s": rates[0], "nnz": int(sth.nnz),
                          "what": "steps returning the CSC in host memory (device-to-host copy and the copy into "
                                  "numpy arrays included); the first call also pins the pooled host blocks"}
    log(f"host output: {rates[1]:.4g} histories/s (first call {rates[0]:.4g})")
    if not a.no_plain:
        sub = subset(bix, 20)
        _, _, _, sp = ctx.run_dij(gph, spec, sub, nbatch=nbatch, rel_threshold=rel, precision=prec,
                                  form=omc.FORM_HISTORY, group=1, csc_on_device=True, download=False)
        out["plain_form"] = {"histories_per_s": sp.histories / sp.wall_seconds, "beamlets": int(sub.n),
                             "what": "history kernel, one beamlet at a time (reference loop order), every 20th beamlet"}
        out["speedup_vs_plain"] = value / out["plain_form"]["histories_per_s"]
        log(f"plain: {out['plain_form']['histories_per_s']:.4g} histories/s")
        # the metric's operating point: every beamlet's histories scaled by k so
        # that the plan dose reaches <= 1% statistical uncertainty in voxels above
        # 50% of its maximum (uncertainty ~ N^-1/2); both forms at that scale,
        # the batched one on every 20th beamlet, the plain one on every 400th
        k = int(math.ceil(max(1.0, (unc50 / 0.01) ** 2)))
        sb = scaled(subset(bix, 20), k)
        sp_ = scaled(subset(bix, 400, offset=7), k)
        _, _, _, stb = ctx.run_dij(gph, spec, sb, nbatch=nbatch, rel_threshold=rel, **opts)
        _, _, _, stp = ctx.run_dij(gph, spec, sp_, nbatch=nbatch, rel_threshold=rel, precision=prec,
                                   form=omc.FORM_HISTORY, group=1, csc_on_device=True, download=False)
        rb, rp = stb.histories / stb.wall_seconds, stp.histories / stp.wall_seconds
        out["at_1pct_unc"] = {"history_scale": k, "histories_for_plan": float(hist / a.steps * k),
                              "batched_histories_per_s": rb, "batched_beamlets": int(sb.n),
                              "plain_histories_per_s": rp, "plain_beamlets": int(sp_.n),
                              "speedup_vs_plain": rb / rp,
                              "plan_seconds_batched": hist / a.steps * k / rb}
        log(f"at 1% unc (x{k} histories): batched {rb:.4g}, plain {rp:.4g}, ratio {rb / rp:.1f}")
    # energy balance of the last timed step (FP64 tallies): primaries' energy
    # against deposits in every region, photons that left and dropped energy
    out["energy_balance"] = {"source_MeV": st_last.e_source, "deposit_MeV": st_last.e_deposit,
                             "escape_MeV": st_last.e_escape, "lost_MeV": st_last.e_lost,
                             "residual": st_last.energy_residual()}
    # the same step without kernel timing, its rounds replayed from HIP graphs
    # against eager launches (the timed steps above launch eagerly: the HIP
    # events around every kernel cannot sit in a graph)
    modes = {}
    for g in (True, False, True, False):
        t0 = time.perf_counter()
        _, _, _, stg = ctx.run_dij(gph, spec, bix, nbatch=nbatch, rel_threshold=rel, graph=g, **opts)
        modes.setdefault("graph" if g else "eager", []).append(stg.histories / (time.perf_counter() - t0))
        if g:
            modes.setdefault("graph_replays", []).append(int(stg.graph_replays))
    out["launch_modes"] = {"graph_histories_per_s": max(modes["graph"]), "eager_histories_per_s": max(modes["eager"]),
                           "graph_replays_per_step": modes["graph_replays"][-1],
                           "what": "untimed steps, best of two each; graph: each job's chunks of four rounds "
                                   "(16 kernels) replay one HIP graph"}
    log(f"launch modes: {out['launch_modes']}")
    if a.tune and form == omc.FORM_EVENT:
        # the autotuner's choice on a 200-beamlet sample of this plan, timed on
        # the whole plan against the engine's defaults (untimed kernels)
        from ompmc_amd import tune
        t0 = time.perf_counter()
        best = tune.tune_dij(ctx, gph, spec, bix, nbatch=nbatch, precision=prec, use_cache=False)
        t_tune = time.perf_counter() - t0
        kn = {k: best[k] for k in ("event_target", "electron_budget", "electron_turns", "streams")}
        rates = {}
        for name, kw in (("tuned", kn), ("default", {}), ("tuned", kn), ("default", {})):
            o2 = dict(opts)
            o2.update(kw)
            t0 = time.perf_counter()
            _, _, _, stt = ctx.run_dij(gph, spec, bix, nbatch=nbatch, rel_threshold=rel, **o2)
            rates.setdefault(name, []).append(stt.histories / (time.perf_counter() - t0))
        out["tuner"] = {"choice": kn, "tuning_seconds": t_tune, "tuned_histories_per_s": max(rates["tuned"]),
                        "default_histories_per_s": max(rates["default"]),
                        "what": "coordinate descent on a 200-beamlet sample; both settings on the whole plan, "
                                "best of two untimed steps each"}
        log(f"tuner: {out['tuner']}")
    if not a.no_fp64 and prec == omc.FP32:
        # FP32 against FP64 on the whole plan: three independent seeds per
        # precision; region (voxels above 20% of the plan maximum) fixed by an
        # FP64 pilot run with its own seed.  The FP64 runs also give the FP64
        # throughput of the step.
        kw = dict(nbatch=nbatch, rel_threshold=rel, want_plan=True, form=form, group=a.group, csc_on_device=True,
                  download=False)
        pilot = ctx.run_dij(gph, spec, bix, precision=omc.FP64, seed=70, **kw)[2][0]
        reg = pilot > 0.2 * pilot.max()
        sums, rates = {}, {}
        for p in (omc.FP32, omc.FP64):
            v = []
            for i in range(3):
                t0 = time.perf_counter()
                _, _, pl_, st_ = ctx.run_dij(gph, spec, bix, precision=p, seed=71 + 3 * i + (p == omc.FP64), **kw)
                rates.setdefault(p, []).append(st_.histories / (time.perf_counter() - t0))
                v.append(pl_[0][reg].sum())
            sums[p] = np.array(v)
        d = sums[omc.FP32].mean() / sums[omc.FP64].mean() - 1.0
        sig = math.hypot(sums[omc.FP32].std(ddof=1), sums[omc.FP64].std(ddof=1)) / math.sqrt(3) / sums[omc.FP64].mean()
        out["fp64_plan_dose_rel_diff"] = {"value": d, "sigma": sig, "beamlets": int(bix.n), "region": "> 20% of max",
                                          "seeds": 3, "voxels": int(reg.sum())}
        out["fp64"] = {"histories_per_s": float(np.median(rates[omc.FP64])),
                       "what": "the whole plan in FP64 transport, matrices in HBM, no kernel timing, median of 3"}
        log(f"fp32/fp64: {d:+.3e} +- {sig:.1e}; fp64 {out['fp64']['histories_per_s']:.4g} histories/s")
    if not a.no_split and form == omc.FORM_EVENT:
        # photon splitting (the reference's nsplit): the same beamlets and
        # histories at several splitting factors; efficiency = 1 / (plan dose
        # uncertainty^2 x seconds), the figure splitting exists to raise.
        # Each factor runs once untimed first, like the main step's warm-up
        # (split runs keep four times the particles in flight, and the first
        # one allocates those pools)
        sub = subset(bix, 20)
        kw = dict(nbatch=nbatch, rel_threshold=rel, want_plan=True, precision=prec, group=a.group,
                  csc_on_device=True, download=False)
        res = {}
        for ns in (1, 4, 20):
            ctx.run_dij(gph, spec, sub, nsplit=ns, **kw)
            t0 = time.perf_counter()
            pl = ctx.run_dij(gph, spec, sub, nsplit=ns, **kw)[2]
            secs = time.perf_counter() - t0
            spd, spv = pl
            sel = spd > 0.5 * spd.max()
            u = float(np.mean(np.sqrt(spv[sel]) / spd[sel]))
            res[str(ns)] = {"seconds": secs, "plan_dose_rel_unc_gt50": u, "efficiency": 1.0 / (u * u * secs)}
        for v in res.values():
            v["efficiency_vs_nsplit1"] = v["efficiency"] / res["1"]["efficiency"]
        out["photon_splitting"] = {"beamlets": int(sub.n), "histories": int(np.sum(np.maximum(sub.nhist // nbatch, 1))
                                                                                * nbatch),
                                   "nsplit": res, "what": "event form, every 20th beamlet, the same histories at "
                                                          "each splitting factor, after one untimed run each; "
                                                          "efficiency = 1/(unc^2 s)"}
        log("splitting: " + ", ".join(f"nsplit {k} {v['seconds']:.2f} s unc {v['plan_dose_rel_unc_gt50']:.4f} "
                                      f"eff x{v['efficiency_vs_nsplit1']:.2f}" for k, v in res.items()))
    if not a.no_serve:
        loads = [float(x) for x in a.serve_loads.split(",")]
        runs = [serve_mix(ctx, value, precision=prec, load=ld, nreq=a.serve_requests, max_hist=a.serve_max_hist)
                for ld in loads]
        out["serving"] = runs[0] if len(runs) == 1 else {"loads": runs}
        log(f"serving: {out['serving']}")
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline_c3(ph, bix, spec, nbatch)
        except Exception as e:          # the baseline is reported, not required
            log(f"cpu baseline failed: {e}")
    out["cpu_baseline"] = cpu
    ctx.close()
    return out


VALU_LANES_PER_S = 256 * 4 * 16 * 2.4e9   # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (non-packed FP32/FP64)


def read_valu(kernel_class, histories_per_s):
    """VALU roofline of a kernel from the committed issue-counter summary:
    lane-instructions per history (SQ_THREAD_CYCLES_VALU: one thread-cycle
    per active lane of a VALU instruction) times the measured histories/s,
    against the device's lane rate; and the issue rate of wave instructions
    (SQ_INSTS_VALU) against its wave-instruction rate (a wave64 instruction
    occupies a 16-lane SIMD for 4 cycles).  None without a summary."""
    path = os.path.join(REPO, "profiles", "c3_charged_issue_counters.md")
    try:
        text = open(path).read()
        t = json.loads(text.split("```json", 1)[1].split("```", 1)[0]).get(KERNEL_NAMES[kernel_class])
    except (OSError, IndexError, ValueError):
        return None
    if not t:
        return None
    lanes = t["valu_thread_cycles_per_history"] * histories_per_s
    waves = t["valu_insts_per_history"] * histories_per_s
    return {"bound": "valu", "kernel": KERNEL_NAMES[kernel_class], "achieved": lanes, "peak": VALU_LANES_PER_S,
            "unit": "lane-instructions/s", "frac": lanes / VALU_LANES_PER_S,
            "issue_frac": waves / (VALU_LANES_PER_S / 64), "lane_utilisation": t["lane_utilisation"],
            "per_history": {"valu_insts": t["valu_insts_per_history"],
                            "valu_thread_cycles": t["valu_thread_cycles_per_history"]},
            "source": "profiles/c3_charged_issue_counters.md (counters per history) x this run's histories/s"}


def read_traffic(kernel_class, histories, launches):
    """HBM bytes per launch of a kernel from the committed PMC summary:
    (FETCH_SIZE + WRITE_SIZE) per history, each scaled by what the counter
    reads for a copy of known size, times the histories of the timed steps
    over their launches.  None when no summary covers the kernel."""
    path = os.path.join(REPO, "profiles", "c3_event_fp32_pmc_traffic.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        t = json.loads(block).get(KERNEL_NAMES[kernel_class])
    except (OSError, IndexError, ValueError):
        return None
    if not t or launches <= 0:
        return None
    return (t["read_bytes_per_history"] + t["write_bytes_per_history"]) * histories / launches


# ------------------------------------------------------------------ c5 (large plan)
def bench_c5(a):
    """The large plan: Dij of 10044 beamlets over 9 beams on a 256x256x200 CT
    at 0.2 cm, 1e9 histories (lognormal per beamlet), 10 batches, threshold
    1e-4.  One step is the whole plan; the step runs in FP32 transport (the
    value) and once in FP64, and the plan doses of the two are compared in
    the voxels above 20% of the FP64 maximum."""
    import ompmc_amd as omc
    from ompmc_amd import phantoms as P
    nbatch, rel = 10, 1e-4
    t_setup = time.perf_counter()
    ph, bix = P.large_plan(a.c5_histories)
    spec = omc.Spectrum.from_histogram(*P.synthetic_6mv())
    ctx = omc.Context(ph.media)
    gph = ctx.phantom(ph, ecut=0.7, pcut=0.01)
    log(f"c5: {bix.n} beamlets, {int(bix.nhist.sum())} histories, setup {time.perf_counter() - t_setup:.1f} s")
    opts = dict(form=omc.FORM_EVENT, csc_on_device=True, download=False)
    sub = subset(bix, 100)
    _, _, _, stc = ctx.run_dij(gph, spec, sub, nbatch=nbatch, rel_threshold=rel, count_events=True, precision=omc.FP32,
                               **opts)
    ctx.run_dij(gph, spec, sub, nbatch=nbatch, rel_threshold=rel, precision=omc.FP64, **opts)   # warm-up
    res = {}
    for prec in (omc.FP32, omc.FP64):
        t0 = time.perf_counter()
        _, _, pl, st = ctx.run_dij(gph, spec, bix, nbatch=nbatch, rel_threshold=rel, want_plan=True,
                                   count_events="time", precision=prec, seed=1802 + prec, **opts)
        wall = time.perf_counter() - t0
        res[prec] = (pl, st, wall)
        log(f"c5 fp{prec}: {st.histories / wall:.4g} histories/s, {wall:.1f} s, {st.groups} groups of "
            f"{st.group_size}, nnz {st.nnz}, energy residual {st.energy_residual():.2e}")
    (p32, v32), st32, w32 = res[omc.FP32]
    (p64, v64), st64, w64 = res[omc.FP64]
    reg = p64 > 0.2 * p64.max()
    d = p32[reg].sum() / p64[reg].sum() - 1.0
    sig = math.sqrt(v32[reg].sum() + v64[reg].sum()) / p64[reg].sum()
    z = np.abs(p32[reg] - p64[reg]) / np.maximum(np.sqrt(v32[reg] + v64[reg]), 1e-300)
    sel50 = p64 > 0.5 * p64.max()
    unc50 = float(np.mean(np.sqrt(v32[sel50]) / p32[sel50]))
    ksec = np.array(st32.kernel_seconds[:])
    kbusy = np.array(st32.kernel_busy_seconds[:])
    dom = 3 if ksec[3] >= ksec[2] else 2
    bph = kernel_bytes(stc, dom - 2) / stc.histories
    busy = kbusy[dom] if kbusy[dom] > 0 else ksec[dom]
    achieved = bph * st32.histories / busy / 1e9
    value = st32.histories / w32
    out = {
        "metric": "histories/s + p50/p99 request latency vs plain GPU form at <=1% dose unc.",
        "value": value, "unit": "histories/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * w32,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "tally_dtype": "f64",
        "data": "synthetic CT, beamlets and 6 MV spectrum (seeded)",
        "config": {"workload": "c5: Dij of a 9-beam plan, 256x256x200 synthetic CT at 0.2 cm, "
                               f"{bix.n} beamlets of 0.5 cm, {bix.nhist.sum():.3g} histories, 10 batches, "
                               "threshold 1e-4", "beamlets": int(bix.n), "histories_per_step": st32.histories,
                   "group": int(st32.group_size), "groups": int(st32.groups)},
        "fp64": {"histories_per_s": st64.histories / w64, "seconds": w64, "nnz": int(st64.nnz)},
        "fp64_plan_dose_rel_diff": {"value": d, "sigma_independent_voxels": sig,
                                    "region": "> 20% of the FP64 plan maximum",
                                    "voxels": int(reg.sum()), "frac_within_2sigma": float(np.mean(z <= 2)),
                                    "what": "one FP32 and one FP64 run of the whole plan, different seeds; "
                                            "sigma_independent_voxels sums the per-voxel batch variances over "
                                            "the region, which understates the sum's uncertainty (voxels share "
                                            "histories); frac_within_2sigma compares voxel by voxel"},
        "plan_dose_rel_unc_gt50": unc50,
        "nnz": int(st32.nnz),
        "energy_residual": {"fp32": st32.energy_residual(), "fp64": st64.energy_residual()},
        "kernels": {KERNEL_NAMES[k]: {"seconds": float(ksec[k]), "busy_seconds": float(kbusy[k]),
                                      "launches": int(st32.kernel_launches[k])}
                    for k in range(5) if st32.kernel_launches[k] > 0},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_history": bph,
                     "busy_seconds": float(busy)},
        "cpu_baseline": None,
    }
    log(f"c5: fp32 {value:.4g}, fp64 {out['fp64']['histories_per_s']:.4g} histories/s; fp32/fp64 {d:+.2e} +- "
        f"{sig:.1e}")
    ctx.close()
    return out


# ------------------------------------------------------------------ c1 / c2 (single dose runs)
def bench_dose(a):
    import ompmc_amd as omc
    from ompmc_amd import phantoms as P
    if a.config == "c1":
        ph, spec, nper = P.water_cube(60, 0.5), omc.Spectrum.mono(2.0), 100000
        wl = "c1: water 60^3 x 0.5 cm, 2 MeV photons, 10x10 cm parallel field"
    else:
        ph = P.slab_phantom(128, 0.3)
        spec, nper = omc.Spectrum.from_histogram(*P.synthetic_6mv()), 1000000
        wl = "c2: slab 128^3 x 0.3 cm (water/lung/bone/water), synthetic 6 MV, 10x10 cm parallel field"
    prec = omc.FP32 if a.precision == "fp32" else omc.FP64
    form = omc.FORM_EVENT if a.form == "event" else omc.FORM_HISTORY
    ctx = omc.Context(ph.media)
    gph = ctx.phantom(ph, ecut=0.7, pcut=0.01)
    kw = dict(collimator=(-5, 5, -5, 5), precision=prec, form=form, event_target=a.target, electron_budget=a.budget)
    _, _, stc = ctx.run_dosxyz(gph, spec, max(10000, nper // 10), 2, seed=7, count_events=True, **kw)
    for _ in range(a.warmup):
        ctx.run_dosxyz(gph, spec, nper, 1, seed=5, **kw)
    t0 = time.perf_counter()
    s1, s2, st = ctx.run_dosxyz(gph, spec, nper, max(2, a.steps), seed=1802, count_events="time", **kw)
    wall = time.perf_counter() - t0
    ksec = np.array(st.kernel_seconds[:])
    dom = 0 if form == omc.FORM_HISTORY else (3 if ksec[3] >= ksec[2] else 2)
    bph = (total_bytes(stc) if dom == 0 else kernel_bytes(stc, dom - 2)) / stc.histories
    achieved = bph * st.histories / ksec[dom] / 1e9
    dose, unc = ctx.dose(gph, s1, s2, nper, max(2, a.steps))
    m = dose > 0.5 * dose.max()
    out = {"metric": "histories/s (dose run)", "value": st.histories / wall, "unit": "histories/s", "n_gpus": 1,
           "steps": max(2, a.steps), "warmup": a.warmup, "ms_per_step": 1e3 * wall / max(2, a.steps),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32" if prec == omc.FP32 else "f64",
        "tally_dtype": "f64", "data": "synthetic (seeded)",
           "config": {"workload": wl, "histories_per_step": nper, "form": a.form, "precision": a.precision},
           "mean_rel_unc_gt50": float(np.mean(unc[m])),
           "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "bytes_per_history": bph},
           "cpu_baseline": None}
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=("c1", "c2", "c3", "c5"))
    ap.add_argument("--c5-histories", type=float, default=1e9, help="c5: histories of the whole plan")
    ap.add_argument("--form", default="event", choices=("event", "history"))
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--group", type=int, default=0, help="beamlets per group (0: by HBM budget)")
    ap.add_argument("--turns", type=int, default=0, help="electron refills per lane before requeue (0 auto, -1 never)")
    ap.add_argument("--batches-per-job", type=int, default=0, help="batches per transport job (0: by HBM budget)")
    ap.add_argument("--target", type=int, default=0, help="event form: particles in flight (0: automatic)")
    ap.add_argument("--budget", type=int, default=0, help="event form: electron steps per event (0: automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="c3: skip the photon-splitting measurement")
    ap.add_argument("--tune", action="store_true", help="c3: also run the autotuner and time its choice")
    ap.add_argument("--subset", type=int, default=1, help="c3: every n-th beamlet only (profiling runs)")
    ap.add_argument("--serve-loads", default="0.5,0.8,1.0",
                    help="offered loads of the served mix, as fractions of the batched throughput (e.g. 0.5,0.8,1.0)")
    ap.add_argument("--serve-requests", type=int, default=12, help="requests per served mix (200: the full workload)")
    ap.add_argument("--serve-max-hist", type=float, default=1e8, help="largest request, in histories")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device once (PMC scale of FETCH_SIZE / WRITE_SIZE) and exit")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        x = torch.ones(a.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.calibrate_copy << 20, "copy_bytes_written": a.calibrate_copy << 20}))
        return
    if a.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {a.gpus} is not supported "
              "(beamlets and batches would shard across GPUs with no reduction; see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    out = bench_c3(a) if a.config == "c3" else bench_c5(a) if a.config == "c5" else bench_dose(a)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
