"""The Dij scoring and assembly kernels (k_dij_fold_finish, or k_accum_slots
with k_dij_finish when a group's batches span several jobs; k_dij_thresh,
k_dij_count, k_dij_compact) on exact inputs through omc_dij_assemble,
against the oracle's restatement of omc_matrad.c:897-1012 and :1603-1705
(oracle/orc_matrad.c, itself pinned by closed forms in
tests/test_oracle_matrad.py).

With every batch of a group in one job the kernels take the reference's
order of operations, so the matrices are equal bit for bit -- values,
variances, rows and columns, ties at the threshold included.  Across
several jobs the carried variance joins the finished sum of squares, one
rounding away from the reference's order: the dose and the pattern stay
exact, the variance agrees to 1e-12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

omc = pytest.importorskip("ompmc_amd")
B = pytest.importorskip("oracle.bindings")
P = omc.phantoms

FAC = 1.602E-10


def _phantom(nx, ny, nz, rho, dx=0.5, dy=0.25, dz=0.4):
    media = [P.WATER]
    return P.Phantom(media, np.arange(nx + 1) * dx, np.arange(ny + 1) * dy, np.arange(nz + 1) * dz,
                     np.ones(nx * ny * nz, np.int32), np.asarray(rho, float))


def _tallies(rng, nbix, nbatch, nreg, fill=0.2, tile_gaps=True):
    e = rng.exponential(1.0, (nbix, nbatch, nreg)) * (rng.random((nbix, nbatch, nreg)) < fill)
    if tile_gaps:                                      # whole 64-region tiles without energy
        for b in range(nbix):
            for k in range(nbatch):
                for t in range(0, nreg, 64):
                    if rng.random() < 0.5:
                        e[b, k, t:t + 64] = 0.0
    return e


def _run(ctx, ph, e, nper, rel, **opts):
    gp = ctx.phantom(ph)
    (jc, ir, val), (vjc, vir, var), st = ctx.dij_assemble(gp, e, nper, rel_threshold=rel, **opts)
    gp.release()
    assert np.array_equal(jc, vjc) and np.array_equal(ir, vir)
    return (jc, ir, val, var), st


@pytest.fixture(scope="module")
def ctx():
    c = omc.Context([P.WATER])
    yield c
    c.close()


@pytest.mark.parametrize("opts", [dict(group=1), dict(group=4, streams=1), dict(group=3, streams=3),
                                  dict(group=0)], ids=["plain", "group4", "3streams", "auto"])
def test_assemble_bitwise_against_oracle(ctx, opts):
    rng = np.random.default_rng(11)
    nx, ny, nz = 13, 11, 7                             # 1001 voxels: tiles and chunks end ragged
    rho = rng.uniform(0.01, 2.0, nx * ny * nz)         # some below the 0.044 cut
    ph = _phantom(nx, ny, nz, rho)
    nbix, nbatch = 10, 6
    e = _tallies(rng, nbix, nbatch, nx * ny * nz + 1)
    e[4] = 0.0                                         # an empty beamlet
    nper = rng.integers(1, 300, nbix)
    rel = 0.03
    o = B.matrad_dij(ph, e, nper, nbatch, rel)
    g, st = _run(ctx, ph, e, nper, rel, **opts)
    assert st.batches_per_job == nbatch                # the fused path
    for a, b, what in zip(g, o, ("jc", "ir", "val", "var")):
        assert a.dtype == b.dtype and np.array_equal(a, b), what
    assert g[0][5] == g[0][4]                          # the empty beamlet repeats jc
    dep = e.sum()
    assert st.e_deposit == pytest.approx(dep, rel=1e-12)


def test_assemble_threshold_ties(ctx):
    """Entries exactly at rel * max are excluded, the next double above kept
    (powers of two keep the scaled values exact)."""
    nx = 70                                            # two tiles
    ph = _phantom(nx, 1, 1, np.ones(nx))
    nbatch, rel = 2, 0.5
    e = np.zeros((2, nbatch, nx + 1))
    e[0, :, 1] = 8.0
    e[0, :, 2] = 4.0
    e[0, :, 66] = np.nextafter(4.0, 5.0)
    e[0, :, 0] = 100.0                                 # outside the phantom
    e[1, :, 65] = 1.0
    o = B.matrad_dij(ph, e, [8, 8], nbatch, rel)
    g, _ = _run(ctx, ph, e, [8, 8], rel, group=2)
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
    assert g[0].tolist() == [0, 2, 3] and g[1].tolist() == [0, 65, 64]


def test_assemble_closed_forms_on_gpu(ctx):
    """Constant batches give dose x/n * 1.602e-10/(rho V) and the variance
    floor dose^2/(nb (nb-1)); the previous beamlet's variance is carried into
    the next one's sum of squares (voxels of 1.602e-10 g make it count)."""
    ph = _phantom(3, 1, 1, np.ones(3), dx=1.0, dy=1.0, dz=1.602E-10)
    nb, n = 5, 100
    e = np.zeros((2, nb, 4))
    e[0, :, 1] = 4.0
    e[1, :, 1] = 2.0
    (jc, ir, val, var), _ = _run(ctx, ph, e, [n, n], 0.0, group=2)
    f = FAC / (1.0 * 1.0 * 1.602E-10)
    d0, d1 = 4.0 / n * f, 2.0 / n * f
    v0 = d0 * d0 / (nb * (nb - 1))
    v1 = ((v0 + nb * (2.0 / n) ** 2) * f * f / (nb - 1) - d1 * d1) / nb
    assert jc.tolist() == [0, 1, 2] and ir.tolist() == [0, 0]
    assert val[0] == pytest.approx(d0, rel=1e-13) and val[1] == pytest.approx(d1, rel=1e-13)
    assert var[0] == pytest.approx(v0, rel=1e-9) and var[1] == pytest.approx(v1, rel=1e-9)
    assert var[1] > 1.02 * d1 * d1 / (nb * (nb - 1))


def test_assemble_across_jobs(ctx):
    """Batches of a group over several jobs (k_accum_slots + k_dij_finish)."""
    rng = np.random.default_rng(12)
    nx, ny, nz = 9, 8, 5
    ph = _phantom(nx, ny, nz, rng.uniform(0.01, 2.0, nx * ny * nz))
    nbix, nbatch = 7, 6
    e = _tallies(rng, nbix, nbatch, nx * ny * nz + 1, tile_gaps=False)
    nper = rng.integers(1, 300, nbix)
    o = B.matrad_dij(ph, e, nper, nbatch, 0.02)
    g, st = _run(ctx, ph, e, nper, 0.02, group=3, batches_per_job=4)
    assert st.batches_per_job == 4
    for k in range(3):
        assert np.array_equal(g[k], o[k])
    assert np.allclose(g[3], o[3], rtol=1e-12, atol=0)


def test_assemble_groups_and_repeats_agree(ctx):
    """A second call with fewer beamlets on the same context starts clean
    (the slots, tile marks and carried variance of the first are gone)."""
    rng = np.random.default_rng(13)
    nx, ny, nz = 16, 16, 4
    ph = _phantom(nx, ny, nz, rng.uniform(0.5, 1.5, nx * ny * nz))
    e = _tallies(rng, 9, 4, nx * ny * nz + 1)
    nper = rng.integers(1, 100, 9)
    first, _ = _run(ctx, ph, e, nper, 1e-3, group=3, streams=2)
    second, _ = _run(ctx, ph, e[:4], nper[:4], 1e-3, group=2)
    o = B.matrad_dij(ph, e[:4], nper[:4], 4, 1e-3)
    for a, b in zip(second, o):
        assert np.array_equal(a, b)
    assert first[0][4] == second[0][4]
