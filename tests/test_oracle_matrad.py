"""The oracle's restatement of omc_matrad's scoring and Dij assembly
(oracle/orc_matrad.c, after ucodes/omc_matrad/omc_matrad.c:856-1012 and
:1583-1705), pinned by what the arithmetic fixes rather than by retyping it:

* constant per-batch energies x give dose = x/n * 1.602e-10/(rho V) and the
  Q11 variance floor dose^2 / (nbatch (nbatch - 1));
* a voxel without energy gives 0 and 0; a voxel with rho < 0.044 gives 0
  and 0 whatever it received;
* the variance of beamlet b-1 enters beamlet b's sum of squares (scaled once
  more by the Gy factor), hand-computed; a beamlet without energy in a voxel
  clears what is carried on;
* an entry exactly equal to rel * max is excluded (strict '>'), the next
  representable value above it is kept;
* an empty beamlet repeats jc; rows are 0-based voxel indices in increasing
  order; region 0 (outside) never appears.

The energy balance of the oracle's transport (deposits in every region plus
the photons that left equal the primaries' energy, no splitting) is pinned
here too, on golden whole-run inputs."""
import numpy as np
import pytest

from tests.conftest import write_case_inp

B = pytest.importorskip("oracle.bindings")


class _Ph:
    def __init__(self, nx, ny, nz, dx=0.5, dy=0.25, dz=0.4, rho=None):
        self.xb = np.arange(nx + 1) * dx
        self.yb = np.arange(ny + 1) * dy
        self.zb = np.arange(nz + 1) * dz
        self.rho = np.ones(nx * ny * nz) if rho is None else np.asarray(rho, float)
        self.vol = dx * dy * dz


FAC = 1.602E-10


def test_constant_batches_give_dose_and_the_variance_floor():
    ph = _Ph(4, 3, 2, rho=np.linspace(0.5, 1.9, 24))
    nb, n = 10, 1000
    x = np.zeros(25)
    x[1:] = np.linspace(1.0, 24.0, 24)                   # MeV per batch, same every batch
    e = np.tile(x, (1, nb, 1))
    jc, ir, val, var = B.matrad_dij(ph, e, [n], nb, 0.0)
    dose = x[1:] / n * FAC / (ph.rho * ph.vol)
    assert jc.tolist() == [0, 24] and ir.tolist() == list(range(24))
    assert np.allclose(val, dose, rtol=1e-13, atol=0)
    assert np.allclose(var, dose ** 2 / (nb * (nb - 1)), rtol=1e-9, atol=0)


def test_zero_and_low_density_voxels():
    rho = np.ones(8)
    rho[3] = 0.0439                                       # below the reference's 0.044 cut
    rho[5] = 0.044                                        # at it: kept
    ph = _Ph(2, 2, 2, rho=rho)
    nb = 4
    e = np.zeros((1, nb, 9))
    e[0, :, 1 + 3] = 2.0
    e[0, :, 1 + 5] = 2.0
    e[0, 1, 1 + 6] = 3.0                                  # one batch only
    sc = B.MatradScore(ph, 1)
    for k in range(nb):
        sc.set_endep(e[0, k])
        sc.accum_endep(10)
    sc.accumulate(nb)
    d, u = sc.accum()
    assert d[1 + 0] == 0.0 and u[1 + 0] == 0.0            # no energy
    assert d[1 + 3] == 0.0 and u[1 + 3] == 0.0            # rho < 0.044
    assert d[1 + 5] > 0.0
    # one batch of four with energy: mean a/4, variance (a^2/3 - a^2/16)/4
    a = 3.0 / 10 * FAC / (ph.vol * 1.0)
    assert d[1 + 6] == pytest.approx(a / 4, rel=1e-13)
    assert u[1 + 6] == pytest.approx((a * a / 3 - a * a / 16) / 4, rel=1e-12)


def test_variance_carries_into_the_next_beamlet():
    # The carried array holds Gy^2 and the sums it joins MeV^2: with voxels of
    # 1.602e-10 g the Gy factor is ~1 and the carry-over is of the same order
    # as the sum of squares (for real voxels it is ~1e-20 of it).
    ph = _Ph(3, 1, 1, dx=1.0, dy=1.0, dz=1.602E-10)
    nb, n = 5, 100
    e = np.zeros((3, nb, 4))
    e[0, :, 1] = 4.0                                      # beamlet 0: constant in voxel 0
    e[1, :, 1] = 2.0                                      # beamlet 1: constant in voxel 0 again
    e[1, :, 2] = 1.0
    e[2, :, 2] = 1.0                                      # beamlet 2: voxel 0 empty -> carry cleared
    e[2, :, 1] = 0.0
    e[2, 0, 3] = 5.0
    jc, ir, val, var = B.matrad_dij(ph, e, [n, n, n], nb, 0.0)
    f = FAC / ph.vol
    d0 = 4.0 / n * f
    v0 = d0 * d0 / (nb * (nb - 1))
    d1 = 2.0 / n * f
    # beamlet 1, voxel 0: accum_endep2 = v0 + nb (2/n)^2, then scaled by f^2
    v1 = ((v0 + nb * (2.0 / n) ** 2) * f * f / (nb - 1) - d1 * d1) / nb
    assert v1 > 1.02 * d1 * d1 / (nb * (nb - 1))          # the carry raises it above the floor
    assert jc.tolist()[:3] == [0, 1, 3]
    assert ir[0] == 0 and val[0] == pytest.approx(d0, rel=1e-13) and var[0] == pytest.approx(v0, rel=1e-9)
    assert ir[1] == 0 and val[1] == pytest.approx(d1, rel=1e-13) and var[1] == pytest.approx(v1, rel=1e-9)
    # beamlet 2: voxel 1 carries beamlet 1's floor; voxel 0 (no energy now) is absent
    d2 = 1.0 / n * f
    v_b1_vox1 = d2 * d2 / (nb * (nb - 1))
    v2 = ((v_b1_vox1 + nb * (1.0 / n) ** 2) * f * f / (nb - 1) - d2 * d2) / nb
    col2 = slice(jc[2], jc[3])
    assert ir[col2].tolist() == [1, 2]
    assert var[col2][0] == pytest.approx(v2, rel=1e-9)
    # a fourth beamlet with energy in voxel 0 only: nothing carried there
    # (beamlet 2 had none), so it is back at the floor
    e4 = np.concatenate([e, np.zeros((1, nb, 4))])
    e4[3, :, 1] = 4.0
    _, ir4, val4, var4 = B.matrad_dij(ph, e4, [n] * 4, nb, 0.0)
    assert var4[-1] == pytest.approx(v0, rel=1e-9)


def test_threshold_is_strict_and_empty_beamlets_repeat_jc():
    ph = _Ph(4, 1, 1)
    nb, n = 2, 8
    rel = 0.5
    e = np.zeros((3, nb, 5))
    e[0, :, 1] = 8.0                                      # the maximum
    e[0, :, 2] = 4.0                                      # exactly rel * max (powers of two: exact)
    e[0, :, 3] = np.nextafter(4.0, 5.0)                   # just above
    e[0, :, 0] = 100.0                                    # outside: never in the matrix
    # beamlet 1 empty; beamlet 2 one voxel
    e[2, :, 4] = 1.0
    jc, ir, val, var = B.matrad_dij(ph, e, [n] * 3, nb, rel)
    assert jc.tolist() == [0, 2, 2, 3]
    assert ir.tolist() == [0, 2, 3]
    d_eq = 4.0 / n * FAC / ph.vol
    assert all(v != d_eq for v in val)


def test_step_by_step_matches_the_whole_loop():
    rng = np.random.default_rng(5)
    ph = _Ph(5, 4, 3, rho=rng.uniform(0.01, 2.0, 60))
    nb, nbix = 6, 7
    e = rng.exponential(1.0, (nbix, nb, 61)) * (rng.random((nbix, nb, 61)) < 0.3)
    nper = rng.integers(1, 50, nbix)
    whole = B.matrad_dij(ph, e, nper, nb, 1e-2)
    sc = B.MatradScore(ph, nbix)
    for b in range(nbix):
        for k in range(nb):
            sc.set_endep(e[b, k])
            sc.accum_endep(nper[b])
        sc.accumulate(nb)
        sc.beamlet_end(b, 1e-2)
    step = sc.csc()
    for a, c in zip(whole, step):
        assert np.array_equal(a, c)
    jc, ir, val, _ = whole
    for j in range(nbix):
        col = ir[jc[j]:jc[j + 1]]
        assert np.all(np.diff(col) > 0) and (len(col) == 0 or (col.min() >= 0 and col.max() < 60))


@pytest.mark.parametrize("name", ["w20", "slab", "electron", "positron", "ct", "vacuum"])
def test_oracle_energy_balance(tmp_path, name):
    """Primaries' energy = deposits (all regions, region 0 included) +
    photons that left + energy dropped (quirk Q6), to FP64 rounding."""
    stem = write_case_inp(tmp_path, name, ncase=400, nbatch=4)
    s = B.Session(stem)
    sm, _, eb = s.dose(100, 4, want_energy=True)
    s.close()
    src, esc, lost, dep = eb
    assert src > 0 and dep > 0
    assert abs(dep - sm.sum()) <= 1e-12 * src
    assert abs(dep + esc + lost - src) <= 1e-11 * src, (src, esc, lost, dep)
    if name in ("w20", "slab", "ct"):
        assert esc > 0.05 * src                           # photon beams: a good part leaves
