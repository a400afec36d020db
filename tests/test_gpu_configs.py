"""Dose runs at the full sizes of the two single-dose workloads.

* 60^3 water at 0.5 cm, 2 MeV photons, 10x10 cm parallel field, 1e6
  histories in 10 batches, FP64: the event form on the GPU against the
  oracle (the reference's algorithm with RANMAR, OpenMP over the histories of
  each batch) -- per voxel above 10% of the maximum, the fractions of
  voxels whose batch means differ by at most 2 (3) combined standard errors
  reach the levels the t distribution of the batch estimates gives, with no
  common bias (tests/conftest.py), and the deposited energies agree within 1% (about ten
  standard errors of the total at 1e6 histories).
* 128^3 layered slab at 0.3 cm (water, lung 0.26, bone 1.85, water),
  synthetic 6 MV spectrum, 10x10 cm field, 1e7 histories: FP32 transport
  against FP64 -- the dose summed over voxels above 20% of the maximum
  (region fixed by an FP64 pilot run) within 0.5%, four seeds per precision,
  and per-voxel agreement of the two runs within their batch uncertainties.
"""
import os

import numpy as np
import pytest

from tests.conftest import assert_agreement, fmt_agreement, z_agreement

pytestmark = pytest.mark.gpu

omc = pytest.importorskip("ompmc_amd")
B = pytest.importorskip("oracle.bindings")
P = omc.phantoms


def _batch_stats(s1, s2, nb):
    m = s1 / nb
    return m, np.sqrt(np.maximum(s2 / nb - m * m, 0.0) / (nb - 1))


def test_water_1e6_fp64_agrees_with_oracle(tmp_path):
    nper, nb = 100000, 10
    ph = P.water_cube(60, 0.5)
    phf = os.path.join(str(tmp_path), "water60.egsphant")
    P.write_egsphant(phf, ph)
    d = P.default_inp(phf, str(tmp_path) + "/", energy=2.0, ncase=nper * nb, nbatch=nb, collimator=(-5, 5, -5, 5))
    stem = P.write_inp(os.path.join(str(tmp_path), "water60"), d)[:-4]
    o1, o2, secs = B.Session(stem).dose_omp(nper, nb)
    c = omc.Context(ph.media)
    gp = c.phantom(ph, ecut=0.7, pcut=0.01)
    g1, g2, st = c.run_dosxyz(gp, omc.Spectrum.mono(2.0), nper, nb, collimator=(-5, 5, -5, 5), precision=omc.FP64,
                              form=omc.FORM_EVENT, seed=11)
    c.close()
    assert st.errors == 0
    mo, uo = _batch_stats(o1[1:], o2[1:], nb)
    mg, ug = _batch_stats(g1[1:], g2[1:], nb)
    sel = np.maximum(mo, mg) > 0.1 * max(mo.max(), mg.max())
    ag = z_agreement(mo[sel], uo[sel], mg[sel], ug[sel], nb)
    eo, eg = mo.sum(), mg.sum()
    print(f"water 1e6 FP64: {fmt_agreement(ag)}; energy per batch {eg:.6g} vs oracle {eo:.6g} ({(eg / eo - 1):+.2e}); "
          f"oracle {secs:.1f} s")
    assert ag["n"] > 1000
    assert_agreement(ag)
    assert abs(eg / eo - 1.0) < 0.01


def test_slab_1e7_fp32_within_half_percent_of_fp64():
    nper, nb = 1000000, 10
    ph = P.slab_phantom(128, 0.3)
    spec = omc.Spectrum.from_histogram(*P.synthetic_6mv())
    c = omc.Context(ph.media)
    gp = c.phantom(ph, ecut=0.7, pcut=0.01)

    def run(prec, seed):
        s1, s2, st = c.run_dosxyz(gp, spec, nper, nb, collimator=(-5, 5, -5, 5), precision=prec,
                                  form=omc.FORM_EVENT, seed=seed)
        assert st.errors == 0
        return s1[1:], s2[1:]
    pilot = run(omc.FP64, 500)[0]
    reg = pilot > 0.2 * pilot.max()
    runs = {p: [run(p, 501 + 2 * k + (p == omc.FP64)) for k in range(4)] for p in (omc.FP32, omc.FP64)}
    c.close()
    tot = {p: np.array([r[0][reg].sum() for r in runs[p]]) for p in runs}
    a, b = tot[omc.FP32], tot[omc.FP64]
    rel = a.mean() / b.mean() - 1.0
    sem = np.hypot(a.std(ddof=1), b.std(ddof=1)) / 2.0 / b.mean()
    # per voxel: the first FP32 run against the first FP64 run
    m32, u32 = _batch_stats(*runs[omc.FP32][0], nb)
    m64, u64 = _batch_stats(*runs[omc.FP64][0], nb)
    sel = np.maximum(m32, m64) > 0.1 * max(m32.max(), m64.max())
    ag = z_agreement(m32[sel], u32[sel], m64[sel], u64[sel], nb)
    print(f"slab 1e7 FP32 vs FP64: region {reg.sum()} voxels, {rel:+.3e} +- {sem:.1e}; per voxel {fmt_agreement(ag)}")
    assert sem < 2.5e-3
    assert abs(rel) < 5e-3
    assert_agreement(ag)
