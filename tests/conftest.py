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
