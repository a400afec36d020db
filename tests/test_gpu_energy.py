"""FP64 energy balance of the transport: the primaries' energy (photons E,
electrons E - m, positrons E + m) equals the energy deposited in every region
(region 0 included) plus the photons that left the phantom -- which the
reference discards without a tally (src/ompmc.c:2008-2017) -- plus energy
dropped without a deposit (quirk Q6), up to rounding, when photons are not
split.

* History form, FP64, replaying the reference's RANMAR stream on golden
  cases: each history runs the oracle's particles, so the escaped and
  deposited energy agree with the oracle's tallies (oracle/omc_oracle.c) as
  the deposits do (a libm last-bit difference may flip a rejection in a few
  histories), and the GPU's own balance closes to 1e-10.
* Event form (persistent queues, delta tracking), FP32 and FP64 transport on
  a CT phantom, dose and Dij: the balance over the run closes to the rounding
  of the transport arithmetic (FP32 1e-5, FP64 1e-10).
* With splitting the balance holds in expectation only (the roulette and
  the split walk are unbiased, not conserving)."""
import os

import numpy as np
import pytest

from tests.conftest import DOSE_CASES, golden, write_case_inp

pytestmark = pytest.mark.gpu

omc = pytest.importorskip("ompmc_amd")
B = pytest.importorskip("oracle.bindings")
P = omc.phantoms


def _spec(kw):
    if kw.get("spectrum"):
        return omc.Spectrum.read(os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    return omc.Spectrum.mono(kw["energy"])


@pytest.mark.parametrize("name", ["w20", "slab", "electron", "positron", "ct", "vacuum"])
def test_history_replay_energy_matches_oracle(tmp_path, name):
    kw, nh, nb = DOSE_CASES[name]
    stem = write_case_inp(tmp_path, name)
    sess = B.Session(stem)
    s1, _, cnt, eb = sess.dose(nh // nb, nb, want_counts=True, want_energy=True)
    off = np.zeros(len(cnt) + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    stream = sess.ranmar(int(off[-1])).astype(np.uint32)
    sess.close()
    ph_h = P.read_egsphant(golden(f"{name}.egsphant"))
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    g1, _, st = c.run_dosxyz(ph, _spec(kw), nh // nb, nb, collimator=kw.get("collimator", (-5, 5, -5, 5)),
                             charge=kw.get("charge", 0), precision=omc.FP64, form=omc.FORM_HISTORY,
                             rng=omc.RNG_REPLAY, replay=stream, replay_offsets=off, keep_diverged_replay=True)
    c.close()
    src, esc, lost, dep = eb
    print(f"{name}: source {st.e_source:.9g} ({src:.9g}), escape {st.e_escape:.9g} ({esc:.9g}), "
          f"lost {st.e_lost:.3g} ({lost:.3g}), deposit {st.e_deposit:.9g} ({dep:.9g}) MeV")
    assert st.e_source == pytest.approx(src, rel=1e-12)          # the same primaries
    assert st.e_deposit == pytest.approx(g1.sum(), rel=1e-12)
    assert abs(st.energy_residual()) < 1e-10
    assert st.e_escape == pytest.approx(esc, rel=1e-6, abs=1e-9 * src)
    assert st.e_deposit == pytest.approx(dep, rel=1e-6)


@pytest.fixture(scope="module")
def ct():
    ph_h = P.synthetic_ct(nx=24, ny=24, nz=16, vx=0.5, vy=0.5, vz=0.5, seed=4, target_radius=1.5)
    c = omc.Context(ph_h.media)
    ph = c.phantom(ph_h, ecut=0.7, pcut=0.01)
    yield c, ph, ph_h
    c.close()


@pytest.mark.parametrize("precision,tol", [(omc.FP32, 1e-5), (omc.FP64, 1e-10)])
def test_event_form_dose_energy_balance(ct, precision, tol):
    c, ph, _ = ct
    spec = omc.Spectrum.read(os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    s, _, st = c.run_dosxyz(ph, spec, 20000, 5, collimator=(-3, 3, -3, 3), precision=precision,
                            form=omc.FORM_EVENT, seed=7)
    assert st.errors == 0
    assert st.e_deposit == pytest.approx(s.sum(), rel=1e-12)
    assert st.e_escape > 0.05 * st.e_source
    r = st.energy_residual()
    print(f"event dose fp{precision}: residual {r:.3g}")
    assert abs(r) < tol


@pytest.mark.parametrize("form", [omc.FORM_EVENT, omc.FORM_HISTORY])
def test_dij_energy_balance(ct, form):
    c, ph, _ = ct
    bix = P.beamlets(nbeams=3, radius=1.0, hist_median=600, hist_range=(300, 1200), iso=(0.0, 0.0, 0.0))
    (jc, ir, val), _, _, st = c.run_dij(ph, omc.Spectrum.mono(2.0), bix, nbatch=5, form=form,
                                        precision=omc.FP32, group=4)
    assert len(ir) > 0
    r = st.energy_residual()
    print(f"Dij form {form}: source {st.e_source:.6g} MeV, residual {r:.3g}")
    assert abs(r) < 1e-5


@pytest.mark.parametrize("form", [omc.FORM_EVENT, omc.FORM_HISTORY])
def test_split_energy_balance_in_expectation(ct, form):
    """A sub-photon that leaves the phantom ends its split; the sub-photons
    after it, which would follow it along the same ray, leave with it, so the
    escape counts their weight too.  The roulette of scattered photons and of
    the photons charged particles produce conserves energy in expectation:
    the mean residual over seeds is zero within its standard error."""
    c, ph, _ = ct
    spec = omc.Spectrum.read(os.path.join(omc.DATA, "spectra", "mohan6.spectrum"))
    res = []
    for seed in range(6):
        _, _, st = c.run_dosxyz(ph, spec, 20000, 5, collimator=(-3, 3, -3, 3), precision=omc.FP32,
                                form=form, seed=100 + seed, nsplit=5)
        assert st.errors == 0
        res.append(st.energy_residual())
    m, sd = np.mean(res), np.std(res, ddof=1)
    print(f"split residuals form {form}: {res}")
    assert abs(m) < 4 * sd / np.sqrt(len(res)) + 1e-4
    assert sd < 0.01
