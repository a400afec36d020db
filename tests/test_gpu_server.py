"""Request server: requests submitted while others run are admitted and run
concurrently, and each result equals the same request run alone (to FP64
summation order, since every history's random stream is keyed by the
request's seed and the history)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

omc = pytest.importorskip("ompmc_amd")
P = omc.phantoms


def test_server_results_same_alone_and_in_the_mix():
    media = [P.WATER, P.LUNG, P.BONE, P.AIR]
    ctx = omc.Context(media)
    phs = [P.synthetic_ct(nx=24, ny=24, nz=16, vx=0.5, vy=0.5, vz=0.5, seed=s) for s in (1, 2)]
    spec = omc.Spectrum.mono(2.0)
    bix = P.beamlets(nbeams=2, radius=1.0, hist_median=600, hist_range=(300, 1200), seed=4)
    srv = omc.Server(ctx, max_inflight=3)
    ids = []
    ids.append(("dose", 0, srv.submit_dose(phs[0], spec, 20000, 5, collimator=(-2, 2, -2, 2), seed=31)))
    ids.append(("dij", 1, srv.submit_dij(phs[1], spec, bix, nbatch=4, seed=32)))
    ids.append(("dose", 1, srv.submit_dose(phs[1], spec, 30000, 5, collimator=(-1, 1, -1, 1), seed=33)))
    ids.append(("dij", 0, srv.submit_dij(phs[0], spec, bix, nbatch=4, seed=34)))
    ids.append(("dij", 1, srv.submit_dij(phs[1], spec, bix, nbatch=4, seed=35, nsplit=4)))
    res = [(k, i, srv.wait(r)) for k, i, r in ids]
    srv.close()
    for k, i, r in res:
        assert r["error"] == 0, r
        assert r["end_t"] >= r["start_t"] >= r["submit_t"] >= 0
    # each request alone, through the plain API, with the same seed
    ph0, ph1 = ctx.phantom(phs[0]), ctx.phantom(phs[1])
    gph = [ph0, ph1]
    s1, _, _ = ctx.run_dosxyz(gph[0], spec, 20000, 5, collimator=(-2, 2, -2, 2), seed=31)
    assert np.allclose(res[0][2]["sum"], s1, rtol=1e-9, atol=1e-12 * s1.max())
    s1, _, _ = ctx.run_dosxyz(gph[1], spec, 30000, 5, collimator=(-1, 1, -1, 1), seed=33)
    assert np.allclose(res[2][2]["sum"], s1, rtol=1e-9, atol=1e-12 * s1.max())
    for j, (ph, seed, ns) in ((1, (gph[1], 32, 1)), (3, (gph[0], 34, 1)), (4, (gph[1], 35, 4))):
        (jc, ir, val), _, _, _ = ctx.run_dij(ph, spec, bix, nbatch=4, seed=seed, nsplit=ns)
        jc2, ir2, val2 = res[j][2]["dij"]
        assert np.array_equal(jc, jc2) and np.array_equal(ir, ir2)
        assert np.allclose(val, val2, rtol=1e-9)
    ctx.close()


def test_server_rejects_request_over_budget():
    ctx = omc.Context([P.WATER])
    srv = omc.Server(ctx, hbm_budget=1 << 20, max_inflight=1)
    with pytest.raises(omc.EngineError) as e:
        srv.submit_dose(P.water_cube(20, 0.5), omc.Spectrum.mono(2.0), 1000, 2)
    assert e.value.code == 8
    srv.close()
    ctx.close()
