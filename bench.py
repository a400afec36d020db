#!/usr/bin/env python3
"""Throughput benchmark of the GPU transport path.

A "step" is one statistical batch of histories of the configured workload,
with phantom and tables already resident in HBM.  Prints one JSON line:
histories/s of the timed steps, the roofline of the transport kernel
(algorithmic bytes per history from on-device event counters, over the
kernel's time measured with HIP events on its stream), and a CPU baseline
(the compiled reference, or the oracle restatement, on the host cores).

    python bench.py [--steps K] [--warmup W] [--config c2] [--form event|history]
                    [--precision fp32|fp64] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def build_config(name):
    from ompmc_amd import phantoms as P
    import ompmc_amd as omc
    if name == "c1":
        ph = P.water_cube(60, 0.5)
        spec = omc.Spectrum.mono(2.0)
        return dict(workload="water 60^3 x 0.5 cm, 2 MeV photons, 10x10 cm parallel field",
                    ph=ph, spec=spec, spec_file=None, energy=2.0, collimator=(-5, 5, -5, 5), nper=200000)
    if name == "c2":
        ph = P.slab_phantom(128, 0.3)
        e, w = P.synthetic_6mv()
        spec = omc.Spectrum.from_histogram(e, w)
        return dict(workload="slab 128^3 x 0.3 cm (water/lung/bone/water), synthetic 6 MV, 10x10 cm parallel field",
                    ph=ph, spec=spec, spec_hist=(e, w), energy=None, collimator=(-5, 5, -5, 5), nper=1000000)
    raise SystemExit(f"unknown config {name}")


def algorithmic_bytes(st):
    """B = 72 N_ev + 5 N_cross + 16 N_dep per history: a read and a write of
    the 36-byte particle state per transport event (howfar/hownear call), one
    density word and medium byte per voxel entered, one FP64
    read-modify-write per energy deposit."""
    h = st.histories
    n_ev = (st.n_howfar + st.n_hownear) / h
    return 72.0 * n_ev + 5.0 * st.n_cross / h + 16.0 * st.n_dep / h, n_ev


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_reference(cfg, seconds_target=12.0):
    """The compiled reference (oracle/_ref/omc_dosxyz, OpenMP) on a bounded
    sample of the same workload, restricted with taskset to the host share."""
    exe = os.path.join(REPO, "oracle", "_ref", "omc_dosxyz")
    if not os.path.exists(exe):
        return None
    from ompmc_amd import phantoms as P
    cores = cpu_cores()
    with tempfile.TemporaryDirectory() as tmp:
        phf = os.path.join(tmp, "ph.egsphant")
        P.write_egsphant(phf, cfg["ph"])
        kw = {}
        if cfg.get("spec_hist") is not None:
            sf = os.path.join(tmp, "s.spectrum")
            P.write_spectrum(sf, "synthetic", *cfg["spec_hist"])
            kw["spectrum"] = sf
        else:
            kw["energy"] = cfg["energy"]

        def run(n):
            d = P.default_inp(phf, tmp + "/", ncase=n, nbatch=10, collimator=cfg["collimator"], **kw)
            stem = P.write_inp(os.path.join(tmp, "in"), d)[:-4]
            cmd = ["taskset", "-c", f"0-{cores - 1}", exe, "-i", stem, "-o", "out"]
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
            if r.returncode != 0:
                return None
            # transport phase: the two "Execution time up to this point" prints
            ts = [float(l.split(":")[1].split()[0]) for l in r.stdout.splitlines()
                  if "Execution time up to this point" in l]
            return ts[-1] - ts[0] if len(ts) >= 2 else None
        t = run(20000)
        if not t or t <= 0:
            return None
        n = int(min(5e7, max(20000, 20000 * seconds_target / t)))
        n -= n % 10
        t = run(n)
        if not t:
            return None
        return dict(value=n / t, unit="histories/s", cores=cores, kind="reference",
                    sample=f"{n} histories of the same workload, reference omc_dosxyz (OpenMP, {cores} threads), "
                           f"transport phase {t:.2f} s")


def cpu_baseline_port(cfg, seconds_target=12.0):
    """The oracle restatement (FP64, OpenMP) when the reference binary is absent."""
    from oracle import bindings as OB
    from ompmc_amd import phantoms as P
    cores = cpu_cores()
    with tempfile.TemporaryDirectory() as tmp:
        phf = os.path.join(tmp, "ph.egsphant")
        P.write_egsphant(phf, cfg["ph"])
        kw = {}
        if cfg.get("spec_hist") is not None:
            sf = os.path.join(tmp, "s.spectrum")
            P.write_spectrum(sf, "synthetic", *cfg["spec_hist"])
            kw["spectrum"] = sf
        else:
            kw["energy"] = cfg["energy"]
        d = P.default_inp(phf, tmp + "/", ncase=1000, nbatch=10, collimator=cfg["collimator"], **kw)
        s = OB.Session(P.write_inp(os.path.join(tmp, "in"), d)[:-4])
        _, _, t = s.dose_omp(2000, 1, cores)
        n = int(max(2000, 2000 * seconds_target / max(t, 1e-3)))
        _, _, t = s.dose_omp(n, 1, cores)
        return dict(value=n / t, unit="histories/s", cores=cores, kind="port",
                    sample=f"{n} histories of the same workload, oracle restatement (OpenMP, {cores} threads)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--form", default="history", choices=("event", "history"))
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--nper", type=int, default=0, help="histories per step (default: config's)")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(histories shard across GPUs by batch, see the design notes)", file=sys.stderr)
        sys.exit(2)
    import ompmc_amd as omc
    cfg = build_config(a.config)
    nper = a.nper or cfg["nper"]
    prec = omc.FP32 if a.precision == "fp32" else omc.FP64
    form = omc.FORM_EVENT if a.form == "event" else omc.FORM_HISTORY
    ph = cfg["ph"]
    ctx = omc.Context(ph.media)
    gph = ctx.phantom(ph, ecut=0.7, pcut=0.01)
    kw = dict(collimator=cfg["collimator"], precision=prec, form=form)
    # warmup (module load, first-touch) and event counts for the roofline
    _, _, stc = ctx.run_dosxyz(gph, cfg["spec"], max(10000, nper // 10), max(2, a.warmup), seed=7, count_events=True, **kw)
    bytes_per_hist, n_ev = algorithmic_bytes(stc)
    t0 = time.perf_counter()
    s1, s2, st = ctx.run_dosxyz(gph, cfg["spec"], nper, max(2, a.steps), seed=1802, **kw)
    wall = time.perf_counter() - t0
    steps = max(2, a.steps)
    hist = st.histories
    value = hist / wall
    kernel_rate = hist / st.gpu_seconds
    achieved = bytes_per_hist * kernel_rate / 1e9
    # statistical quality of the timed run (voxels above 50% of the maximum)
    dose, unc = ctx.dose(gph, s1, s2, nper, steps)
    m = dose > 0.5 * dose.max()
    unc50 = float(np.mean(unc[m])) if m.any() else float("nan")
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline_reference(cfg) or cpu_baseline_port(cfg)
    out = {
        "metric": "histories/s (transport, one GPU)",
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transport, f64 tallies" if prec == omc.FP32 else "f64",
        "data": "synthetic phantom and spectrum (seeded)",
        "config": {"workload": cfg["workload"], "histories_per_step": nper, "form": a.form,
                   "precision": a.precision},
        "kernel_histories_per_s": kernel_rate,
        "mean_rel_unc_gt50": unc50,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_history": bytes_per_hist, "events_per_history": n_ev},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
