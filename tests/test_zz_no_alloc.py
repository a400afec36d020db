"""omc_run_dosxyz_noalloc, the dose run that allocates no device memory,
against omc_run_dosxyz in the history form (same kernels, same random
streams): the batch sums agree to FP64 summation order (the tallies are
atomic adds in a schedule-dependent order), in FP32 and FP64 transport, on
two calls in a row, the second with a smaller phantom and fewer histories
(nothing of the first call's scratch contents may leak into it).  Sizes
beyond the scratch array, and the forms and options it does not serve, are
refused with OMC_ERR_ARG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

omc = pytest.importorskip("ompmc_amd")
torch = pytest.importorskip("torch")
P = omc.phantoms


def _noalloc(ctx, ph, nreg, spec, nper, nb, **kw):
    d1 = torch.full((nreg,), 7.0, dtype=torch.float64, device="cuda")     # garbage the call must clear
    d2 = torch.full((nreg,), 7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = ctx.run_dosxyz_noalloc(ph, spec, nper, nb, d1.data_ptr(), d2.data_ptr(), **kw)
    return d1.cpu().numpy(), d2.cpu().numpy(), st


def _same(a, b):
    return np.allclose(a, b, rtol=1e-9, atol=1e-12 * max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("prec", [omc.FP32, omc.FP64])
def test_noalloc_matches_history_form_on_two_calls(prec):
    big = P.synthetic_ct(nx=48, ny=48, nz=32, vx=0.4, vy=0.4, vz=0.4, seed=5, target_radius=2.0)
    small = P.synthetic_ct(nx=20, ny=16, nz=12, vx=0.5, vy=0.5, vz=0.5, seed=6, target_radius=1.0)
    ctx = omc.Context(big.media)
    spec = omc.Spectrum.from_histogram(*P.synthetic_6mv())
    for ph_h, nper, nb, seed in ((big, 20000, 4, 71), (small, 3000, 3, 72)):
        ph = ctx.phantom(ph_h, ecut=0.7, pcut=0.01)
        kw = dict(collimator=(-3, 3, -3, 3), precision=prec, seed=seed)
        n1, n2, st = _noalloc(ctx, ph, ph.nreg, spec, nper, nb, **kw)
        r1, r2, rs = ctx.run_dosxyz(ph, spec, nper, nb, form=omc.FORM_HISTORY, **kw)
        assert st.errors == 0 and st.histories == nper * nb
        assert n1.sum() > 0
        assert _same(n1, r1) and _same(n2, r2)
        assert st.e_source == pytest.approx(rs.e_source, rel=1e-9)
        assert st.e_deposit == pytest.approx(rs.e_deposit, rel=1e-9)
    ctx.close()


def test_noalloc_refuses_what_the_scratch_array_cannot_hold():
    ctx = omc.Context([P.WATER])
    spec = omc.Spectrum.mono(2.0)
    ok = ctx.phantom(P.water_cube(10, 0.5))
    d = torch.zeros(ok.nreg, dtype=torch.float64, device="cuda")
    for kw in (dict(form=omc.FORM_EVENT), dict(nsplit=20)):
        with pytest.raises(omc.EngineError) as e:
            ctx.run_dosxyz_noalloc(ok, spec, 100, 2, d.data_ptr(), d.data_ptr(), **kw)
        assert e.value.code == 5
    # 256 x 256 x 64 voxels + the outside region: one region over the limit
    ph = P.Phantom([P.WATER], np.arange(257) * 0.1, np.arange(257) * 0.1, np.arange(65) * 0.1,
                   np.ones((64, 256, 256), np.int32), np.ones((64, 256, 256)), "over")
    big = ctx.phantom(ph)
    assert big.nreg == omc.NOALLOC_MAX_REGIONS + 1
    dd = torch.zeros(big.nreg, dtype=torch.float64, device="cuda")
    with pytest.raises(omc.EngineError) as e:
        ctx.run_dosxyz_noalloc(big, spec, 100, 2, dd.data_ptr(), dd.data_ptr())
    assert e.value.code == 5 and "scratch" in str(e.value)
    ctx.close()
