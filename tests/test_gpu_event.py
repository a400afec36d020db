"""Event-queue form against the oracle and against the plain history form.

The event form keys every particle's random stream by its history and the
draw position of its parent, so (a) its dose is independent of the queue
target, the grid and the electron step budget up to the order of FP64 atomic
additions, and (b) it agrees with the oracle (RANMAR, the reference's CPU
algorithm) statistically: per voxel above 10% of the maximum, the difference
of the two batch means against their combined batch uncertainty."""
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
    return ctx.run_dosxyz(ph, _spectrum(kw), nh // nb, nb, collimator=kw.get("collimator", (-5, 5, -5, 5)),
                          charge=kw.get("charge", 0), **opts)


def _batch_stats(s1, s2, nb):
    """Mean per batch and the uncertainty of that mean (batch method)."""
    m = s1 / nb
    var = np.maximum(s2 / nb - m * m, 0.0) / (nb - 1)
    return m, np.sqrt(var)


def _agreement(a1, a2, b1, b2, nb):
    """Per-voxel agreement above 10% of the maximum, at the levels the
    t distribution of the batch estimates gives (tests/conftest.py)."""
    ma, ua = _batch_stats(a1[1:], a2[1:], nb)
    mb, ub = _batch_stats(b1[1:], b2[1:], nb)
    sel = np.maximum(ma, mb) > 0.1 * max(ma.max(), mb.max())
    return z_agreement(ma[sel], ua[sel], mb[sel], ub[sel], nb)


@pytest.fixture(scope="module")
def slab(tmp_path_factory):
    ph_h = P.read_egsphant(golden("slab.egsphant"))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    yield c, ph
    c.close()


@pytest.mark.parametrize("name", ["slab", "electron", "positron", "vacuum"])
def test_event_fp64_agrees_with_oracle(tmp_path, name):
    """Event form (FP64, keyed streams) against the oracle (FP64, RANMAR) on the same
    phantom, source and history count."""
    nh, nb = 20000, 20
    stem = write_case_inp(tmp_path, name)
    s1, s2 = B.Session(stem).dose(nh // nb, nb)
    ph_h = P.read_egsphant(golden(f"{name}.egsphant"))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    g1, g2, st = _run(c, ph, name, nh, nb, precision=omc.FP64, form=omc.FORM_EVENT, seed=11)
    assert st.errors == 0
    ag = _agreement(s1, s2, g1, g2, nb)
    print(f"{name}: {fmt_agreement(ag)}")
    assert ag["n"] > 20
    assert_agreement(ag, name)
    # total energy deposited in the phantom: the spread of the GPU total over
    # independent seeds gives its standard deviation for nh histories
    tots = [g1[1:].sum()]
    for k in range(7):
        tots.append(_run(c, ph, name, nh, nb, precision=omc.FP64, form=omc.FORM_EVENT, seed=100 + k)[0][1:].sum())
    c.close()
    sd = float(np.std(tots, ddof=1))
    ta = s1[1:].sum()
    print(f"{name}: oracle total {ta:.1f}, GPU {np.mean(tots):.1f} +- {sd:.1f} per run")
    assert abs(ta - np.mean(tots)) <= 4.0 * sd * np.sqrt(1 + 1 / len(tots))


def test_event_result_independent_of_schedule(slab):
    """Queue target, grid and electron step budget change the schedule, not the
    particles: the dose agrees to FP64 summation order."""
    c, ph = slab
    ref = _run(c, ph, "slab", 4000, 4, precision=omc.FP64, form=omc.FORM_EVENT, seed=3)[0]
    for kw in (dict(event_target=100, electron_budget=1), dict(event_target=777, electron_budget=3, grid=7),
               dict(event_target=1 << 16, electron_budget=1000), dict(electron_turns=1, electron_budget=2),
               dict(electron_turns=-1), dict(batches_per_job=1), dict(batches_per_job=3, electron_turns=2)):
        g = _run(c, ph, "slab", 4000, 4, precision=omc.FP64, form=omc.FORM_EVENT, seed=3, **kw)[0]
        assert np.allclose(g, ref, rtol=1e-10, atol=1e-12 * ref.max()), kw


def test_event_batches_per_job_keep_batch_statistics(slab):
    """All batches in one job (a tally copy per batch) and one job per batch
    give the same per-region sums and sums of squares: the batch statistics
    are taken batch by batch in both."""
    c, ph = slab
    a, a2, sa = _run(c, ph, "slab", 4000, 5, precision=omc.FP64, form=omc.FORM_EVENT, seed=4, batches_per_job=1)
    b, b2, sb = _run(c, ph, "slab", 4000, 5, precision=omc.FP64, form=omc.FORM_EVENT, seed=4)
    assert sa.batches_per_job == 1 and sb.batches_per_job == 5
    assert np.allclose(a, b, rtol=1e-10, atol=1e-12 * b.max())
    assert np.allclose(a2, b2, rtol=1e-10, atol=1e-12 * b2.max())


def test_event_pool_growth_reruns_job():
    """6 MeV electrons with a tiny queue target overflow the pools (each lane
    runs its electron to the end and emits every bremsstrahlung photon); the
    job is rerun with larger pools from zeroed tallies and gives the dose of
    the default run to FP64 summation order."""
    ph_h = P.read_egsphant(golden("electron.egsphant"))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    ref, _, st0 = _run(c, ph, "electron", 4000, 4, precision=omc.FP64, form=omc.FORM_EVENT, seed=8)
    g, _, st1 = _run(c, ph, "electron", 4000, 4, precision=omc.FP64, form=omc.FORM_EVENT, seed=8, event_target=64)
    c.close()
    assert st0.errors == 0 and st1.errors == 0
    assert np.allclose(g, ref, rtol=1e-10, atol=1e-12 * ref.max())


def test_event_matches_history_form_fp32(slab):
    """Both GPU forms sample the same physics; FP32 transport."""
    c, ph = slab
    nh, nb = 200000, 20
    h1, h2, sh = _run(c, ph, "slab", nh, nb, precision=omc.FP32, form=omc.FORM_HISTORY, seed=5)
    e1, e2, se = _run(c, ph, "slab", nh, nb, precision=omc.FP32, form=omc.FORM_EVENT, seed=6)
    assert sh.errors == 0 and se.errors == 0
    ag = _agreement(h1, h2, e1, e2, nb)
    print(f"history vs event: {fmt_agreement(ag)}")
    assert_agreement(ag)


@pytest.fixture(scope="module")
def ct():
    """Synthetic CT: every voxel's density differs from its medium's nominal
    one, so FP32 rounding of density-scaled path lengths is exercised."""
    ph_h = P.synthetic_ct(nx=64, ny=64, nz=48, vx=0.3, vy=0.3, vz=0.3, seed=9, target_radius=2.0)
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    yield c, ph
    c.close()


def test_fp32_dose_within_half_percent_of_fp64(ct):
    """FP32 transport against the FP64 GPU run on the synthetic CT: the dose
    summed over voxels above 20% of the maximum (region fixed by an
    independent pilot run) differs by less than 0.5%; eight independent seeds
    of 500k histories per precision put the statistical uncertainty of the
    difference near 0.15%."""
    c, ph = ct
    spec = omc.Spectrum.read(omc.os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))

    def run(prec, seed, n=500000):
        return c.run_dosxyz(ph, spec, n // 4, 4, collimator=(-3, 3, -3, 3), precision=prec, form=omc.FORM_EVENT,
                            seed=seed)[0][1:]
    pilot = run(omc.FP64, 999)
    sel = pilot > 0.2 * pilot.max()
    reg = {prec: np.array([run(prec, 200 + k)[sel].sum() for k in range(8)]) for prec in (omc.FP32, omc.FP64)}
    a, b = reg[omc.FP32], reg[omc.FP64]
    rel = a.mean() / b.mean() - 1.0
    sem = np.hypot(a.std(ddof=1), b.std(ddof=1)) / np.sqrt(8) / b.mean()
    print(f"FP32 vs FP64 dose (voxels > 20% max): {rel:+.3e} +- {sem:.1e}")
    assert sem < 2.5e-3
    assert abs(rel) < 5e-3


def test_delta_tracking_matches_voxel_walk(ct):
    """Photon delta tracking (the event form's default) against the
    reference's voxel-by-voxel flight, on a CT with air, lung, bone and
    per-voxel densities: per-voxel agreement of the batch means, and the dose
    summed over voxels above 20% of the maximum (region fixed by a pilot run)
    compared through the spread of independent seeds."""
    c, ph = ct
    spec = omc.Spectrum.read(omc.os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    kw = dict(collimator=(-3, 3, -3, 3), precision=omc.FP64, form=omc.FORM_EVENT)
    a1, a2, _ = c.run_dosxyz(ph, spec, 20000, 20, seed=41, **kw)
    b1, b2, _ = c.run_dosxyz(ph, spec, 20000, 20, seed=42, voxel_walk=True, **kw)
    ag = _agreement(a1, a2, b1, b2, 20)
    print(f"delta vs voxel walk: {fmt_agreement(ag)}")
    assert ag["n"] > 100
    assert_agreement(ag)
    pilot = c.run_dosxyz(ph, spec, 50000, 4, seed=40, **kw)[0][1:]
    sel = pilot > 0.2 * pilot.max()
    sums = {vw: np.array([c.run_dosxyz(ph, spec, 50000, 4, seed=50 + 2 * k + vw, voxel_walk=vw, **kw)[0][1:][sel].sum()
                          for k in range(6)]) for vw in (False, True)}
    a, b = sums[False], sums[True]
    d = a.mean() / b.mean() - 1.0
    sem = np.hypot(a.std(ddof=1), b.std(ddof=1)) / np.sqrt(6) / b.mean()
    print(f"delta vs voxel walk, dose > 20% of max: {d:+.2e} +- {sem:.1e}")
    assert abs(d) < 4.0 * sem + 1e-3


def test_event_counts_and_kernel_times(slab):
    c, ph = slab
    _, _, st = _run(c, ph, "slab", 20000, 2, precision=omc.FP32, form=omc.FORM_EVENT, count_events=True)
    assert st.n_howfar > 0 and st.n_hownear > 0 and st.n_dep > 0
    assert st.ev_howfar[0] + st.ev_howfar[1] == st.n_howfar
    assert st.events_photon > 0 and st.events_charged > 0 and st.rounds > 0
    assert st.kernel_seconds[2] > 0 and st.kernel_seconds[3] > 0
    assert st.peak_inflight >= 1


def test_atomic_probe_measures_a_rate():
    """The FP64 atomic probe (secondary roofline of the tallies) runs and gives
    a plausible rate: scattered adds over a 512 MiB table are no faster than
    over a 512 KiB one, and both are between 1e8 and 1e13 per second."""
    small = omc.atomic_probe(1 << 16, 1 << 24)
    large = omc.atomic_probe(1 << 26, 1 << 26)
    print(f"FP64 atomics: {small:.3g}/s over 512 KiB, {large:.3g}/s over 512 MiB")
    assert 1e8 < large < 1e13 and 1e8 < small < 1e13
    assert large <= 1.2 * small


def test_stuck_particle_guard_reports_limit():
    """An electron still going after max_particle_steps condensed-history
    steps is dropped and the run fails with OMC_ERR_LIMIT (4), the kernels
    exiting normally; forced here from the host with a bound of 16 steps,
    which 6 MeV electrons exceed, and checked not to trip at the default."""
    ph_h = P.read_egsphant(golden("electron.egsphant"))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    with pytest.raises(omc.EngineError) as ei:
        _run(c, ph, "electron", 400, 2, precision=omc.FP32, form=omc.FORM_EVENT, seed=4, electron_turns=1,
             electron_budget=2, max_particle_steps=16)
    assert ei.value.code == 4 and "progress" in str(ei.value)
    _, _, st = _run(c, ph, "electron", 400, 2, precision=omc.FP32, form=omc.FORM_EVENT, seed=4, electron_turns=1,
                    electron_budget=2)
    c.close()
    assert st.errors == 0


def test_history_step_guard_reports_limit():
    """The history form's step guard: a history that takes more than
    max_particle_steps transport calls is cut and the run fails with
    OMC_ERR_LIMIT (4) whether or not statistics are asked for; 6 MeV electron
    histories take more than two such calls.  At the default the same run is
    clean."""
    ph_h = P.read_egsphant(golden("electron.egsphant"))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    for prec in (omc.FP32, omc.FP64):
        with pytest.raises(omc.EngineError) as ei:
            _run(c, ph, "electron", 400, 2, precision=prec, form=omc.FORM_HISTORY, seed=4, max_particle_steps=2)
        assert ei.value.code == 4 and "guard" in str(ei.value)
    _, _, st = _run(c, ph, "electron", 400, 2, precision=omc.FP32, form=omc.FORM_HISTORY, seed=4)
    c.close()
    assert st.errors == 0


def test_unequal_voxel_spacing_agrees_with_oracle(tmp_path):
    """Voxel boundaries at unequal spacing, which the egsphant format allows:
    the event form then finds the voxel of a delta-tracking site by bisection
    instead of by multiplication.  FP64 event form against the oracle on the
    same phantom file, and FP32 event form against the history form's
    voxel-by-voxel walk, per voxel within the batch uncertainties."""
    rng = np.random.default_rng(17)

    def bounds(n, lo, hi):
        w = rng.uniform(0.5, 1.5, n)
        b = lo + np.concatenate([[0.0], np.cumsum(w)]) * (hi - lo) / w.sum()
        b[-1] = hi
        return b
    n = 16
    xb, yb, zb = bounds(n, -4.0, 4.0), bounds(n, -4.0, 4.0), bounds(n, 0.0, 10.0)
    assert np.ptp(np.diff(xb)) > 0.2 and np.ptp(np.diff(zb)) > 0.2
    zc = 0.5 * (zb[1:] + zb[:-1])
    lay = np.where(zc < 3.0, 1, np.where(zc < 6.0, 2, 3)).astype(np.int32)
    med = np.broadcast_to(lay[:, None, None], (n, n, n)).copy()
    rho = np.where(med == 1, 1.0, np.where(med == 2, 0.26, 1.85))
    phf = str(tmp_path / "uneven.egsphant")
    P.write_egsphant(phf, P.Phantom([P.WATER, P.LUNG, P.BONE], xb, yb, zb, med, rho, "uneven"))
    ph_h = P.read_egsphant(phf)
    nh, nb = 50000, 10
    d = P.default_inp(phf, str(tmp_path) + "/", energy=2.0, ncase=nh, nbatch=nb, collimator=(-2, 2, -2, 2))
    stem = P.write_inp(str(tmp_path / "uneven"), d)[:-4]
    s1, s2, _ = B.Session(stem).dose_omp(nh // nb, nb)
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.mono(2.0)
    kw = dict(collimator=(-2, 2, -2, 2))
    g1, g2, st = c.run_dosxyz(ph, spec, nh // nb, nb, precision=omc.FP64, form=omc.FORM_EVENT, seed=13, **kw)
    e1, e2, se = c.run_dosxyz(ph, spec, nh // nb, nb, precision=omc.FP32, form=omc.FORM_EVENT, seed=14, **kw)
    h1, h2, sh = c.run_dosxyz(ph, spec, nh // nb, nb, precision=omc.FP32, form=omc.FORM_HISTORY, seed=15, **kw)
    c.close()
    assert st.errors == 0 and se.errors == 0 and sh.errors == 0
    for what, (a1, a2, b1, b2) in (("event FP64 vs oracle", (s1, s2, g1, g2)),
                                   ("event FP32 vs history FP32", (h1, h2, e1, e2))):
        ag = _agreement(a1, a2, b1, b2, nb)
        print(f"unequal spacing, {what}: {fmt_agreement(ag)}")
        assert ag["n"] > 100
        assert_agreement(ag, what)


def test_graph_replay_matches_eager_launches(ct):
    """Without kernel timing the event form enqueues each chunk of rounds by
    replaying a HIP graph built from the job's first chunk.  Several dose
    and Dij runs in a row, graph and eager alternating (eager work and
    synchronisations between the replays), each graph run against the eager
    result to FP64 summation order; the Dij runs use three concurrent lanes,
    each building and replaying its graphs on its own thread."""
    c, ph = ct
    spec = omc.Spectrum.read(omc.os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    kw = dict(collimator=(-3, 3, -3, 3), precision=omc.FP32, form=omc.FORM_EVENT, seed=77, batches_per_job=2)
    e1, e2, _ = c.run_dosxyz(ph, spec, 20000, 6, graph=False, **kw)
    for rep in range(3):
        g1, g2, st = c.run_dosxyz(ph, spec, 20000, 6, graph=True, **kw)
        assert st.errors == 0 and st.rounds > 4
        assert st.graph_replays > 0 and st.graph_fallbacks == 0
        assert np.allclose(g1, e1, rtol=1e-9, atol=1e-12 * e1.max()), rep
        assert np.allclose(g2, e2, rtol=1e-9, atol=1e-12 * e2.max()), rep
        c.run_dosxyz(ph, spec, 5000, 2, graph=False, **kw)          # eager work between replays
    bix = P.beamlets(nbeams=3, radius=1.5, hist_median=800, hist_range=(400, 1600), seed=9)
    dk = dict(nbatch=5, precision=omc.FP32, form=omc.FORM_EVENT, group=4, streams=3, seed=78)
    (jc, ir, val), _, _, _ = c.run_dij(ph, spec, bix, graph=False, **dk)
    for rep in range(2):
        (jc2, ir2, val2), _, _, st = c.run_dij(ph, spec, bix, graph=True, **dk)
        assert st.errors == 0 and st.streams == 3
        assert st.graph_replays > 0 and st.graph_fallbacks == 0
        assert np.array_equal(jc, jc2) and np.array_equal(ir, ir2), rep
        assert np.allclose(val, val2, rtol=1e-9), rep


def test_lds_staged_bins_match_device_memory_reads(ct):
    """The FP32 charged-particle kernel reads the electron coefficients of a
    particle's energy bin from a copy in LDS (lds_tables on, the default) or
    from device memory (off): the same values, so the same seed gives the
    same dose up to FP64 summation order, for dose and Dij runs."""
    c, ph = ct
    spec = omc.Spectrum.read(omc.os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    kw = dict(collimator=(-3, 3, -3, 3), precision=omc.FP32, form=omc.FORM_EVENT, seed=91)
    a1, a2, sa = c.run_dosxyz(ph, spec, 20000, 4, lds_tables=True, **kw)
    b1, b2, sb = c.run_dosxyz(ph, spec, 20000, 4, lds_tables=False, **kw)
    assert sa.errors == 0 and sb.errors == 0
    assert sa.lds_table_bytes > 0 and sb.lds_table_bytes == 0
    assert sa.lds_table_bytes <= 64 * 1024 and sa.lds_table_bytes % 80 == 0
    assert np.allclose(a1, b1, rtol=1e-9, atol=1e-12 * b1.max())
    assert np.allclose(a2, b2, rtol=1e-9, atol=1e-12 * b2.max())
    # FP64 transport reads device memory whatever the option says
    _, _, s64 = c.run_dosxyz(ph, spec, 2000, 2, lds_tables=True, **dict(kw, precision=omc.FP64))
    assert s64.lds_table_bytes == 0
    bix = P.beamlets(nbeams=2, radius=1.5, hist_median=800, hist_range=(400, 1600), seed=19)
    dk = dict(nbatch=4, precision=omc.FP32, form=omc.FORM_EVENT, streams=2, seed=92)
    (jc, ir, val), _, _, st1 = c.run_dij(ph, spec, bix, lds_tables=True, **dk)
    (jc2, ir2, val2), _, _, st2 = c.run_dij(ph, spec, bix, lds_tables=False, **dk)
    assert st1.lds_table_bytes > 0 and st2.lds_table_bytes == 0
    assert np.array_equal(jc, jc2) and np.array_equal(ir, ir2)
    assert np.allclose(val, val2, rtol=1e-9)


def test_nine_media_read_bins_from_device_memory():
    """Nine media, the most a context takes: their electron bins (9 x 151
    records of 80 bytes) exceed the 64 KiB the charged kernel may stage in
    LDS, so it reads them in place.  Nine slabs along the beam, 6 MV: FP32
    transport against FP64 per voxel above 10% of the maximum and in the
    deposited energy (1e6 histories each, independent seeds)."""
    media = [P.WATER, P.LUNG, P.BONE, P.AIR, "AL521ICRU", "PMMA521ICRU", "POLYSTY521ICRU", "CU521ICRU",
             "ICRUTISSUE521ICRU"]
    dens = [1.0, 0.26, 1.85, 0.0012, 2.7, 1.19, 1.06, 8.96, 1.0]
    n, nz, vox = 24, 36, 0.5
    zc = (np.arange(nz) + 0.5) * vox
    lay = np.minimum((zc // 2.0).astype(np.int32), 8)         # 2 cm slabs, the last one to the end
    med = np.broadcast_to((lay + 1)[:, None, None], (nz, n, n)).astype(np.int32).copy()
    rho = np.asarray(dens)[med - 1]
    half = n * vox / 2.0
    edges = -half + vox * np.arange(n + 1, dtype=np.float64)
    ph_h = P.Phantom(media, edges, edges.copy(), vox * np.arange(nz + 1, dtype=np.float64), med, rho, "nine")
    c = omc.Context(media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    spec = omc.Spectrum.read(omc.os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    nb = 10
    kw = dict(collimator=(-4, 4, -4, 4), form=omc.FORM_EVENT)
    a1, a2, sa = c.run_dosxyz(ph, spec, 100000, nb, precision=omc.FP32, seed=301, **kw)
    b1, b2, sb = c.run_dosxyz(ph, spec, 100000, nb, precision=omc.FP64, seed=302, **kw)
    c.close()
    assert sa.errors == 0 and sb.errors == 0
    assert sa.lds_table_bytes == 0
    assert abs(sa.energy_residual()) < 1e-6 and abs(sb.energy_residual()) < 1e-9
    ag = _agreement(a1, a2, b1, b2, nb)
    ea, eb = a1[1:].sum(), b1[1:].sum()
    rel = ea / eb - 1.0
    print(f"nine media, FP32 vs FP64: {fmt_agreement(ag)}; deposit {rel:+.2e}")
    assert ag["n"] > 1000
    assert_agreement(ag)
    assert abs(rel) < 5e-3
