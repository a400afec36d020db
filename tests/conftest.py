"""Shared fixtures.  Tests marked `gpu` need an MI355X; the rest run on CPU."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X); run with -m gpu")


def golden(name):
    return os.path.join(GOLDEN, name)


# inputs of the golden whole-run cases (must match tests/golden/make_golden.py)
DOSE_CASES = {
    "w20": (dict(energy=2.0, collimator=(-2, 2, -2, 2)), 2000, 10),
    "small": (dict(energy=2.0), 3000, 10),
    "spec": (dict(spectrum="mohan6"), 2000, 10),
    "electron": (dict(energy=6.0, charge=-1, collimator=(-1, 1, -1, 1)), 1000, 10),
    "positron": (dict(energy=3.0, charge=1, collimator=(-1, 1, -1, 1)), 1000, 10),
    "slab": (dict(spectrum="mohan6", collimator=(-3, 3, -3, 3)), 2000, 10),
    "split": (dict(energy=2.0, collimator=(-2, 2, -2, 2), nsplit=5), 2000, 10),
    "split_slab": (dict(spectrum="mohan6", collimator=(-3, 3, -3, 3), nsplit=3), 2000, 10),
    "split20": (dict(spectrum="mohan6", collimator=(-3, 3, -3, 3), nsplit=20), 500, 10),
    "electron_split": (dict(energy=6.0, charge=-1, collimator=(-1, 1, -1, 1), nsplit=10), 500, 10),
    "positron_split": (dict(energy=3.0, charge=1, collimator=(-1, 1, -1, 1), nsplit=4), 500, 10),
    "ct": (dict(spectrum="mohan6", collimator=(-2, 2, -2, 2)), 2000, 10),
    "vacuum": (dict(energy=6.0, charge=-1, collimator=(-1, 1, -1, 1)), 1000, 10),
}


def write_case_inp(tmp_path, name, data_folder=None, **over):
    """Writes the .inp of a golden case (paths into this repository) and
    returns its stem."""
    from ompmc_amd import phantoms as P
    kw, nh, nb = DOSE_CASES[name]
    kw = dict(kw)
    nh, nb = over.pop("ncase", nh), over.pop("nbatch", nb)
    kw.update(over)
    spec = kw.pop("spectrum", None)
    if spec == "mohan6":
        kw["spectrum"] = os.path.join(REPO, "data", "spectra", "mohan6.spectrum")
    d = P.default_inp(golden(f"{name}.egsphant"), str(tmp_path) + "/", ncase=nh, nbatch=nb,
                      data_folder=data_folder, **kw)
    return P.write_inp(os.path.join(str(tmp_path), name), d)[:-4]


@pytest.fixture(scope="session")
def textured_data(tmp_path_factory):
    """Data folder with the textured spin file (rejection loops exercised)."""
    import subprocess
    d = tmp_path_factory.mktemp("data_textured")
    for f in os.listdir(os.path.join(REPO, "data")):
        if f.endswith(".data") and f != "spinms.data":
            os.symlink(os.path.join(REPO, "data", f), os.path.join(d, f))
    subprocess.check_call([sys.executable, os.path.join(REPO, "data", "gen_spinms.py"), "--variant", "textured",
                           str(d / "spinms.data")])
    return str(d) + "/"


# ------------------------------------------------------------------ statistical agreement
# Two independent batch estimates of a voxel's dose, means ma, mb with
# standard errors ua, ub from na, nb batches (the batch method: ua^2 is the
# sample variance of the batch values over na), give z = (ma - mb) /
# sqrt(ua^2 + ub^2).  With normal batch values and equal means, z follows
# Student's t with the Welch-Satterthwaite degrees of freedom
#     nu = (ua^2 + ub^2)^2 / (ua^4 / (na - 1) + ub^4 / (nb - 1)),
# between min(na, nb) - 1 and na + nb - 2.  The expected fraction of voxels
# with |z| <= k is therefore the mean over voxels of P(|t_nu| <= k): 0.939
# (k = 2) and 0.992 (k = 3) at 10 batches each, against 0.954 and 0.997 for
# exact standard errors.  North of 95% within 2 sigma is out of reach of any
# unbiased estimator at 10-20 batches, and a demand of 99% would fail on
# correct code.  The observed fraction scatters around its expectation with
# variance f(1 - f) / n for n independent voxels; neighbouring voxels share
# particle tracks, so the variance is taken three times larger, and the
# acceptance is the expectation less four of those standard errors (at least
# 0.02).  A bias common to all voxels moves the mean of the signed z away
# from zero long before it moves the fractions: |mean z| must stay within
# four standard errors of that mean (again with the factor three).
def z_agreement(ma, ua, mb, ub, na, nb=None):
    """Per-voxel agreement of two batch estimates (arrays over the selected
    voxels).  Returns a dict with the observed and accepted fractions within
    2 and 3 sigma, the mean signed z and its bound, and the voxel count."""
    from scipy import stats
    import numpy as np
    nb = nb or na
    va, vb = np.asarray(ua, float) ** 2, np.asarray(ub, float) ** 2
    den = np.maximum(va + vb, 1e-300)
    z = (np.asarray(ma, float) - np.asarray(mb, float)) / np.sqrt(den)
    nu = den ** 2 / np.maximum(va ** 2 / (na - 1) + vb ** 2 / (nb - 1), 1e-300)
    nu = np.clip(nu, min(na, nb) - 1, na + nb - 2)
    n = len(z)
    out = {"n": n, "zbar": float(np.mean(z)) if n else 0.0}
    for k in (2, 3):
        e = float(np.mean(2.0 * stats.t.cdf(k, nu) - 1.0)) if n else 1.0
        f = float(np.mean(np.abs(z) <= k)) if n else 1.0
        out[f"f{k}"], out[f"e{k}"] = f, e
        out[f"lim{k}"] = e - max(0.02, 4.0 * np.sqrt(3.0 * e * (1.0 - e) / max(n, 1)))
    sdz = float(np.std(z)) if n > 1 else 1.0
    out["zbar_lim"] = 4.0 * sdz * np.sqrt(3.0 / max(n, 1))
    return out


def assert_agreement(ag, what=""):
    assert ag["f2"] >= ag["lim2"] and ag["f3"] >= ag["lim3"], (what, ag)
    assert abs(ag["zbar"]) <= ag["zbar_lim"], (what, ag)


def fmt_agreement(ag):
    return (f"{ag['n']} voxels, {ag['f2']:.3f} within 2 sigma (expected {ag['e2']:.3f}, accept >= {ag['lim2']:.3f}), "
            f"{ag['f3']:.3f} within 3 (expected {ag['e3']:.3f}, >= {ag['lim3']:.3f}), mean z {ag['zbar']:+.3f} "
            f"(|.| <= {ag['zbar_lim']:.3f})")
