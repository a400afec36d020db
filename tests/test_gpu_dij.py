"""Dose-influence (Dij) runs against the oracle's restatement of omc_matrad.

Replay: the history form in FP64, each history fed its own slice of the
oracle's RANMAR stream (the reference's serial loop order beamlet, batch,
history), reproduces the oracle's sparse matrix -- its transport
(oracle/omc_oracle.c) and its scoring and assembly (oracle/orc_matrad.c) --:
the same columns and rows, values within 1e-9, including the variance output
with the reference's carry-over from one beamlet to the next.  The batched forms (beamlet groups,
event queues) are checked against the plain per-beamlet form."""
import os

import numpy as np
import pytest

from tests.conftest import assert_agreement, fmt_agreement, z_agreement

pytestmark = pytest.mark.gpu

omc = pytest.importorskip("ompmc_amd")
B = pytest.importorskip("oracle.bindings")
P = omc.phantoms


def _case(tmp_path, nbeams=2, radius=1.0, nhist=(150, 400), seed=3):
    ph = P.synthetic_ct(nx=20, ny=20, nz=12, vx=0.5, vy=0.5, vz=0.5, seed=seed, target_radius=1.0)
    bix = P.beamlets(nbeams=nbeams, sad=100.0, bixel=0.5, radius=radius, seed=seed,
                     hist_median=np.sqrt(nhist[0] * nhist[1]), hist_sigma_log10=0.3, hist_range=nhist)
    phf = os.path.join(str(tmp_path), "ct.egsphant")
    P.write_egsphant(phf, ph)
    d = P.default_inp(phf, str(tmp_path) + "/", ncase=1000, nbatch=10)
    stem = P.write_inp(os.path.join(str(tmp_path), "ct"), d)[:-4]
    return P.read_egsphant(phf), bix, stem


def _bix_dict(bix, gaussian=0, sigma=0.0):
    return dict(ibeam=bix.ibeam, src=bix.src.reshape(-1), corner=bix.corner.reshape(-1),
                side1=bix.side1.reshape(-1), side2=bix.side2.reshape(-1), gaussian=gaussian, sigma=sigma)


def _nper(bix, nbatch):
    nh = np.asarray(bix.nhist, np.int64)
    nh = np.where(nh // nbatch == 0, nbatch, nh)
    return (nh // nbatch).astype(np.int32)


def _compare_csc(g, o, what):
    jc_g, ir_g, v_g = g
    jc_o, ir_o, v_o = o
    assert len(jc_g) == len(jc_o)
    same = np.array_equal(jc_g, jc_o) and np.array_equal(ir_g, ir_o)
    if not same:
        # a threshold decision may flip on a last-bit difference: compare as dense
        n = max(ir_g.max(initial=0), ir_o.max(initial=0)) + 1
        ncol = len(jc_g) - 1
        dg = np.zeros((ncol, n)); do = np.zeros((ncol, n))
        for j in range(ncol):
            dg[j, ir_g[jc_g[j]:jc_g[j + 1]]] = v_g[jc_g[j]:jc_g[j + 1]]
            do[j, ir_o[jc_o[j]:jc_o[j + 1]]] = v_o[jc_o[j]:jc_o[j + 1]]
        diff = np.abs(dg - do) > 1e-9 * np.maximum(np.abs(do), 1e-300)
        assert diff.mean() < 1e-3, f"{what}: {diff.sum()} entries differ"
        return
    rel = np.abs(v_g - v_o) / np.maximum(np.abs(v_o), 1e-300)
    frac = float(np.mean(rel < 1e-9)) if len(rel) else 1.0
    assert frac >= 0.99, f"{what}: {frac:.4f} of entries within 1e-9, worst {rel.max():.3g}"


@pytest.mark.parametrize("group,gaussian", [(1, 0), (3, 0), (3, 1)])
def test_dij_replay_fp64_matches_oracle(tmp_path, group, gaussian):
    nbatch, rel, sigma = 4, 1e-4, 0.3
    ph_h, bix, stem = _case(tmp_path)
    sess = B.Session(stem)
    spec = omc.Spectrum.read(os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    nper = _nper(bix, nbatch)
    s, s2, cnt, bat = sess.matrad(spec, _bix_dict(bix, gaussian, sigma), nper, nbatch, want_counts=True,
                                  want_batches=True)
    jc, ir, val, var = B.matrad_dij(ph_h, bat, nper, nbatch, rel)
    off = np.zeros(len(cnt) + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    stream = sess.ranmar(int(off[-1])).astype(np.uint32)
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    (g, gv, _, st) = c.run_dij(ph, spec, bix, nbatch=nbatch, rel_threshold=rel, want_var=True,
                               precision=omc.FP64, form=omc.FORM_HISTORY, group=group, rng=omc.RNG_REPLAY,
                               replay=stream, replay_offsets=off, gaussian=gaussian, sigma=sigma)
    c.close()
    assert st.errors == 0
    assert st.group_size == min(group, bix.n)
    assert len(val) > 0
    _compare_csc(g, (jc, ir, val), "dose")
    _compare_csc(gv, (jc, ir, var), "variance")


def test_dij_layout_invariants(tmp_path):
    ph_h, bix, _ = _case(tmp_path)
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    (jc, ir, val), _, (pd, pv), st = c.run_dij(ph, spec, bix, nbatch=4, rel_threshold=1e-3, want_plan=True,
                                                form=omc.FORM_EVENT, precision=omc.FP32)
    c.close()
    nvox = ph_h.med.size
    assert jc[0] == 0 and np.all(np.diff(jc) >= 0) and jc[-1] == len(ir) == st.nnz
    for j in range(len(jc) - 1):
        col = ir[jc[j]:jc[j + 1]]
        assert np.all(np.diff(col) > 0)                    # strictly increasing rows
        v = val[jc[j]:jc[j + 1]]
        if len(v):
            assert np.all(v > 1e-3 * v.max())              # strict threshold of the column max
    assert ir.min() >= 0 and ir.max() < nvox
    # the beamlet-summed dose equals the column sum of the unthresholded matrix
    # to within the thresholded remainder
    dense = np.bincount(ir, weights=val, minlength=nvox)
    assert dense.sum() <= pd.sum() * (1 + 1e-9)
    assert dense.sum() >= 0.99 * pd.sum()


def test_dij_grouping_and_forms_agree(tmp_path):
    """Plain per-beamlet history form (the reference's loop), history form over
    beamlet groups and the event form over beamlet groups: the two history
    variants see identical random streams (same matrix to FP64 summation
    order); the event form agrees statistically through the summed dose."""
    ph_h, bix, _ = _case(tmp_path, nhist=(2000, 4000))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    kw = dict(nbatch=5, rel_threshold=1e-4, want_plan=True, precision=omc.FP64)
    plain = c.run_dij(ph, spec, bix, form=omc.FORM_HISTORY, group=1, **kw)
    grouped = c.run_dij(ph, spec, bix, form=omc.FORM_HISTORY, group=5, **kw)
    event = c.run_dij(ph, spec, bix, form=omc.FORM_EVENT, group=4, **kw)
    c.close()
    _compare_csc(grouped[0], plain[0], "grouped vs plain")
    pd_h, pv_h = plain[2]
    pd_e, pv_e = event[2]
    sel = pd_h > 0.5 * pd_h.max()
    ag = z_agreement(pd_h[sel], np.sqrt(pv_h[sel]), pd_e[sel], np.sqrt(pv_e[sel]), 5)
    print(f"event vs history plan dose: {fmt_agreement(ag)}")
    assert_agreement(ag)


def test_autotuner_picks_a_measured_setting(tmp_path, monkeypatch):
    """The tuner times its candidates on a sample of the caller's beamlets,
    returns the fastest and caches it; the matrix does not depend on the
    setting (random streams are keyed by beamlet, batch and history)."""
    from ompmc_amd import tune
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "tune.json"))
    ph_h, bix, _ = _case(tmp_path, nhist=(500, 1000))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    best = tune.tune_dij(c, ph, spec, bix, nbatch=4, sample=8, targets=(4096, 65536), budgets=(2, 8), turns=(3, -1),
                         streams=(1, 2), max_evals=6)
    assert best["event_target"] in (4096, 65536) and best["electron_budget"] in (2, 8)
    assert best["electron_turns"] in (3, -1) and best["streams"] in (1, 2)
    assert (tmp_path / "tune.json").exists()
    import json
    table = json.load(open(tmp_path / "tune.json")).popitem()[1]["table"]
    assert 2 <= len(table) <= 6                              # the descent stops within its budget
    assert tune.tune_dij(c, ph, spec, bix, nbatch=4, sample=8) == best          # from the cache
    # a split plan is tuned with splitting, under its own cache entry
    bs = tune.tune_dij(c, ph, spec, bix, nbatch=4, sample=8, targets=(4096, 65536), budgets=(8,), turns=(24,),
                       streams=(1,), nsplit=4)
    assert bs["nsplit"] == 4 and best["nsplit"] == 1
    assert tune.tune_dij(c, ph, spec, bix, nbatch=4, sample=8, nsplit=4) == bs
    a = c.run_dij(ph, spec, bix, nbatch=4, seed=5, **best)[0]
    b = c.run_dij(ph, spec, bix, nbatch=4, seed=5, event_target=999, electron_budget=3, electron_turns=2)[0]
    c.close()
    _compare_csc(a, b, "tuned vs other setting")


def test_dij_batches_per_job_same_matrix(tmp_path):
    """Event-form Dij with every batch of a group in one job (a tally slot per
    beamlet and batch), two batches per job, and one batch per job: the same
    histories with the same stream keys, so the matrix and its variance agree
    to FP64 summation order."""
    ph_h, bix, _ = _case(tmp_path, nhist=(400, 800))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    kw = dict(nbatch=5, rel_threshold=1e-4, want_var=True, precision=omc.FP64, form=omc.FORM_EVENT, seed=11)
    one = c.run_dij(ph, spec, bix, batches_per_job=1, group=3, **kw)
    two = c.run_dij(ph, spec, bix, batches_per_job=2, group=4, **kw)
    all_ = c.run_dij(ph, spec, bix, **kw)
    lanes = c.run_dij(ph, spec, bix, group=2, streams=2, want_plan=True, **kw)
    lane1 = c.run_dij(ph, spec, bix, group=2, streams=1, want_plan=True, **kw)
    c.close()
    assert one[3].batches_per_job == 1 and two[3].batches_per_job == 2 and all_[3].batches_per_job == 5
    assert lanes[3].streams == 2 and lanes[3].groups == (bix.n + 1) // 2 and lane1[3].streams == 1
    # concurrent groups: the finish still runs in group order (variance carry, plan totals)
    assert np.allclose(lanes[2][0], lane1[2][0], rtol=1e-12, atol=0)
    assert np.allclose(lanes[2][1], lane1[2][1], rtol=1e-12, atol=0)
    for other, what in ((two, "2 batches per job"), (all_, "all batches per job"), (lanes, "two streams")):
        _compare_csc(other[0], one[0], what)
        _compare_csc(other[1], one[1], what + " (variance)")


def test_repeated_runs_reuse_cached_memory(tmp_path):
    """Run-scoped device blocks go back to a per-device cache and are handed
    out again with whatever the last run left in them: a Dij run, a dose run,
    and both again must repeat their results (to FP64 summation order), so
    every buffer that has to start at zero is cleared by its run."""
    ph_h, bix, _ = _case(tmp_path, nhist=(400, 800))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    kw = dict(nbatch=5, rel_threshold=1e-4, want_var=True, want_plan=True, precision=omc.FP64,
              form=omc.FORM_EVENT, seed=3, group=3)
    dkw = dict(collimator=(-1, 1, -1, 1), precision=omc.FP64, form=omc.FORM_EVENT, seed=9)
    a = c.run_dij(ph, spec, bix, **kw)
    s1, q1, _ = c.run_dosxyz(ph, spec, 1000, 5, **dkw)
    b = c.run_dij(ph, spec, bix, **kw)
    s2, q2, _ = c.run_dosxyz(ph, spec, 1000, 5, **dkw)
    c.close()
    _compare_csc(b[0], a[0], "repeated Dij")
    _compare_csc(b[1], a[1], "repeated Dij (variance)")
    assert np.allclose(b[2][0], a[2][0], rtol=1e-10, atol=0)
    assert np.allclose(s2, s1, rtol=1e-10, atol=1e-12 * s1.max())
    assert np.allclose(q2, q1, rtol=1e-10, atol=1e-12 * q1.max())


def test_dij_output_kept_on_device_matches_host_output(tmp_path):
    """csc_on_device leaves the matrices in device memory (omc_csc.device = 1);
    downloaded, they equal the host-output run's to FP64 summation order."""
    ph_h, bix, _ = _case(tmp_path, nhist=(400, 800))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    kw = dict(nbatch=5, rel_threshold=1e-4, want_var=True, precision=omc.FP64, form=omc.FORM_EVENT, seed=13, group=3)
    host = c.run_dij(ph, spec, bix, **kw)
    dev = c.run_dij(ph, spec, bix, csc_on_device=True, **kw)
    jc_only = c.run_dij(ph, spec, bix, csc_on_device=True, download=False, **kw)
    c.close()
    _compare_csc(dev[0], host[0], "device output")
    _compare_csc(dev[1], host[1], "device output (variance)")
    assert np.array_equal(jc_only[0][0], dev[0][0]) and jc_only[0][1] is None
    assert dev[3].nnz == len(dev[0][1]) and abs(dev[3].nnz - host[3].nnz) <= max(1, host[3].nnz // 1000)
