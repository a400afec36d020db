"""Photon splitting in the event-queue form.

The event form splits each photon it takes from the queue (nsplit sub-photons
at stratified mean free paths along its ray, one survivor among the scattered
photons) and rouletting the photons that charged particles produce, like the
reference's photon() and brems/annih/rannih with nsplit > 1
(src/ompmc.c:1916-2136, 4167-4182, 4356-4371, 4645-4660).  The particles are
the reference's, in another order and with other random streams, so the
checks are statistical:

* against the oracle (RANMAR, the reference's algorithm) on the two golden
  splitting cases, per voxel above 10% of the maximum and in the total energy;
* against unsplit transport: splitting changes the variance, not the mean;
* the schedule (queue target, grid, electron step budget) changes nothing
  beyond FP64 summation order, since every particle's stream is keyed by its
  history and its parent's draw position;
* Dij columns with splitting agree with the history form's.
"""
import numpy as np
import pytest

from tests.conftest import DOSE_CASES, assert_agreement, fmt_agreement, golden, write_case_inp, z_agreement

pytestmark = pytest.mark.gpu

omc = pytest.importorskip("ompmc_amd")
B = pytest.importorskip("oracle.bindings")
P = omc.phantoms


def _spectrum(kw):
    return omc.Spectrum.read(omc.os.path.join(omc.DATA, "spectra", "mohan6.spectrum")) \
        if kw.get("spectrum") else omc.Spectrum.mono(kw["energy"])


def _run(ctx, ph, name, nh, nb, **opts):
    kw = DOSE_CASES[name][0]
    opts.setdefault("nsplit", kw.get("nsplit", 1))
    return ctx.run_dosxyz(ph, _spectrum(kw), nh // nb, nb, collimator=kw.get("collimator", (-5, 5, -5, 5)),
                          charge=kw.get("charge", 0), **opts)


def _batch_stats(s1, s2, nb):
    m = s1 / nb
    return m, np.sqrt(np.maximum(s2 / nb - m * m, 0.0) / (nb - 1))


def _agreement(a1, a2, b1, b2, nb):
    """Per-voxel agreement above 10% of the maximum, at the levels the
    t distribution of the batch estimates gives (tests/conftest.py)."""
    ma, ua = _batch_stats(a1[1:], a2[1:], nb)
    mb, ub = _batch_stats(b1[1:], b2[1:], nb)
    sel = np.maximum(ma, mb) > 0.1 * max(ma.max(), mb.max())
    return z_agreement(ma[sel], ua[sel], mb[sel], ub[sel], nb)


def _ctx(name):
    ph_h = P.read_egsphant(golden(f"{name}.egsphant"))
    c = omc.Context(ph_h.media)
    return c, c.phantom(ph_h, ecut=0.7, pcut=0.01)


@pytest.mark.parametrize("name", ["split", "split_slab", "electron_split", "positron_split"])
def test_event_split_fp64_agrees_with_oracle(tmp_path, name):
    """Event form with splitting (FP64, keyed streams) against the oracle with the
    same nsplit (FP64, RANMAR): per voxel, and the total energy against the
    spread of the GPU total over independent seeds."""
    nh, nb = 200000, 20
    stem = write_case_inp(tmp_path, name)
    s1, s2, _ = B.Session(stem).dose_omp(nh // nb, nb)
    c, ph = _ctx(name)
    g1, g2, st = _run(c, ph, name, nh, nb, precision=omc.FP64, form=omc.FORM_EVENT, seed=21)
    assert st.errors == 0
    ag = _agreement(s1, s2, g1, g2, nb)
    print(f"{name}: {fmt_agreement(ag)}")
    assert ag["n"] > 20
    assert_agreement(ag, name)
    tots = [g1[1:].sum()]
    for k in range(7):
        tots.append(_run(c, ph, name, nh, nb, precision=omc.FP64, form=omc.FORM_EVENT, seed=200 + k)[0][1:].sum())
    c.close()
    sd = float(np.std(tots, ddof=1))
    ta = s1[1:].sum()
    print(f"{name}: oracle total {ta:.1f}, GPU {np.mean(tots):.1f} +- {sd:.1f} per run")
    assert abs(ta - np.mean(tots)) <= 3.0 * sd * np.sqrt(1 + 1 / len(tots))


def test_event_split_matches_unsplit_and_lowers_variance():
    """nsplit 8 against nsplit 1 in the event form (FP32): the same dose per
    voxel within the batch uncertainties, and a smaller uncertainty per
    history where the photons interact (the point of splitting)."""
    name = "split_slab"
    c, ph = _ctx(name)
    nh, nb = 400000, 20
    a1, a2, sa = _run(c, ph, name, nh, nb, precision=omc.FP32, form=omc.FORM_EVENT, seed=31, nsplit=1)
    b1, b2, sb = _run(c, ph, name, nh, nb, precision=omc.FP32, form=omc.FORM_EVENT, seed=32, nsplit=8)
    c.close()
    assert sa.errors == 0 and sb.errors == 0
    ag = _agreement(a1, a2, b1, b2, nb)
    ma, ua = _batch_stats(a1[1:], a2[1:], nb)
    mb, ub = _batch_stats(b1[1:], b2[1:], nb)
    sel = np.maximum(ma, mb) > 0.5 * max(ma.max(), mb.max())
    ratio = float(np.median(ub[sel] / np.maximum(ua[sel], 1e-300)))
    print(f"nsplit 8 vs 1: {fmt_agreement(ag)}; median uncertainty ratio {ratio:.3f} "
          f"(charged events {sb.events_charged / max(sa.events_charged, 1):.2f}x)")
    assert_agreement(ag)
    assert ratio < 0.85
    assert abs(mb[sel].sum() / ma[sel].sum() - 1.0) < 0.01


def test_event_split_result_independent_of_schedule():
    """Queue target, grid and electron step budget change the schedule of a
    split run, not its particles."""
    name = "split"
    c, ph = _ctx(name)
    ref = _run(c, ph, name, 4000, 4, precision=omc.FP64, form=omc.FORM_EVENT, seed=3)[0]
    for kw in (dict(event_target=100, electron_budget=1), dict(event_target=777, electron_budget=3, grid=7),
               dict(event_target=1 << 16, electron_budget=1000), dict(electron_turns=1, electron_budget=2),
               dict(batches_per_job=1), dict(count_events=True)):
        g, _, st = _run(c, ph, name, 4000, 4, precision=omc.FP64, form=omc.FORM_EVENT, seed=3, **kw)
        assert np.allclose(g, ref, rtol=1e-10, atol=1e-12 * ref.max()), kw
        if kw.get("count_events"):
            # every split counts once as a photon step; its sub-photons cross voxels
            assert st.n_photon >= st.events_photon > 0 and st.ev_cross[0] > 0 and st.n_electron > 0
    c.close()


def test_event_split_largest_factor_fits_the_pools():
    """nsplit 64 (the largest accepted) with the default queue target: the
    injection is scaled down so that the charged products of one photon
    launch fit the pools without a rerun.  Two runs at nsplit 64 against four
    unsplit ones, independent seeds, each side summed over its runs.  (One
    unsplit run of this size carries a fluctuation common to the voxels,
    which share histories; against a single such run the mean z of an
    unbiased comparison came out at +0.16 to +0.25.)"""
    name = "split"
    c, ph = _ctx(name)

    def runs(ns, seeds):
        s1 = s2 = 0.0
        for sd in seeds:
            a1, a2, st = _run(c, ph, name, 20000, 10, precision=omc.FP32, form=omc.FORM_EVENT, seed=sd, nsplit=ns)
            assert st.errors == 0, (ns, sd)
            s1, s2 = s1 + a1, s2 + a2
        return s1, s2, 10 * len(seeds), st
    g1, g2, na, st = runs(64, (41, 43))
    h1, h2, nb, _ = runs(1, (42, 44, 46, 48))
    c.close()
    ma, ua = _batch_stats(g1[1:], g2[1:], na)
    mb, ub = _batch_stats(h1[1:], h2[1:], nb)
    sel = np.maximum(ma, mb) > 0.1 * max(ma.max(), mb.max())
    ag = z_agreement(ma[sel], ua[sel], mb[sel], ub[sel], na, nb)
    print(f"nsplit 64 vs 1: {fmt_agreement(ag)}; peak in flight {st.peak_inflight}")
    assert_agreement(ag)


@pytest.mark.parametrize("ns", [4, 12])
def test_dij_split_event_matches_history_form(tmp_path, ns):
    """Dij with photon splitting: the event form's beamlet-summed dose agrees
    with the plain per-beamlet history form's within their uncertainties (at
    nsplit 12 the history form keeps its stacks in device memory)."""
    ph_h = P.synthetic_ct(nx=20, ny=20, nz=12, vx=0.5, vy=0.5, vz=0.5, seed=3, target_radius=1.0)
    bix = P.beamlets(nbeams=2, sad=100.0, bixel=0.5, radius=1.0, seed=3, hist_median=3000.0,
                     hist_sigma_log10=0.2, hist_range=(2000, 4000))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    kw = dict(nbatch=5, rel_threshold=1e-4, want_plan=True, precision=omc.FP64, nsplit=ns)
    plain = c.run_dij(ph, spec, bix, form=omc.FORM_HISTORY, group=1, seed=7, **kw)
    event = c.run_dij(ph, spec, bix, form=omc.FORM_EVENT, group=4, seed=8, **kw)
    c.close()
    assert plain[3].errors == 0 and event[3].errors == 0
    pd_h, pv_h = plain[2]
    pd_e, pv_e = event[2]
    sel = pd_h > 0.5 * pd_h.max()
    ag = z_agreement(pd_h[sel], np.sqrt(pv_h[sel]), pd_e[sel], np.sqrt(pv_e[sel]), 5)
    print(f"Dij nsplit {ns}, history vs event plan dose: {fmt_agreement(ag)}")
    assert ag["n"] > 20
    assert_agreement(ag)


def test_history_form_deep_splitting_matches_event_form():
    """Beyond nsplit 8 the history form keeps each history's particles in a
    stack in device memory (the scratch stack holds 96).  At nsplit 20, the
    reference's sample-input value, its dose agrees with the event form's;
    factors above 64 are refused in both forms."""
    name = "split_slab"
    c, ph = _ctx(name)
    nh, nb = 100000, 10
    h1, h2, sh = _run(c, ph, name, nh, nb, precision=omc.FP32, form=omc.FORM_HISTORY, seed=61, nsplit=20)
    e1, e2, se = _run(c, ph, name, nh, nb, precision=omc.FP32, form=omc.FORM_EVENT, seed=62, nsplit=20)
    for form in (omc.FORM_HISTORY, omc.FORM_EVENT):
        with pytest.raises(omc.EngineError, match="nsplit"):
            _run(c, ph, name, 100, 2, precision=omc.FP32, form=form, nsplit=65)
    c.close()
    assert sh.errors == 0 and se.errors == 0
    ag = _agreement(h1, h2, e1, e2, nb)
    print(f"nsplit 20, history vs event: {fmt_agreement(ag)}")
    assert_agreement(ag)
