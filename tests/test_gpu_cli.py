"""The drop-in omc_dosxyz CLI (ompmc_amd/omc_dosxyz) on a reference input file:
same arguments, writes <output folder><name>.3ddose in the reference's
layout; its dose agrees with the reference binary's file (oracle restatement
when the reference tree is absent) within the combined statistical
uncertainty that both files carry."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import DOSE_CASES, REPO, assert_agreement, fmt_agreement, write_case_inp, z_agreement

pytestmark = pytest.mark.gpu

B = pytest.importorskip("oracle.bindings")
CLI = os.path.join(REPO, "ompmc_amd", "omc_dosxyz")


def read_3ddose(path):
    """Block layout of the reference writer (omc_dosxyz.c:801-886): counts,
    x/y/z boundaries, dose, relative uncertainty."""
    tok = open(path).read().split()
    nx, ny, nz = (int(t) for t in tok[:3])
    v = np.array(tok[3:], float)
    o = nx + ny + nz + 3
    n = nx * ny * nz
    return (nx, ny, nz), v[:o], v[o:o + n], v[o + n:o + 2 * n]


@pytest.mark.parametrize("name,flags", [("w20", []), ("w20", ["--fp64", "--history"]), ("split", [])])
def test_cli_3ddose_matches_reference_statistically(tmp_path, name, flags):
    """The last case has `nsplit 5` in its input, like the reference's sample
    input, and runs in the default event form."""
    stem = write_case_inp(tmp_path, name, ncase=200000)
    r = subprocess.run([CLI, "-i", stem, "-o", "gpu"] + flags, cwd=str(tmp_path), stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    if DOSE_CASES[name][0].get("nsplit", 1) > 1:
        assert "Photon splitting enabled, nsplit = 5" in r.stdout
    dims, bounds, dg, ug = read_3ddose(str(tmp_path / "gpu.3ddose"))
    # reference file: the same input run by the oracle (bitwise equal to the
    # reference's serial binary) and written by the host layer
    s = B.Session(stem)
    nb = DOSE_CASES[name][2]
    s1, s2 = s.dose(200000 // nb, nb)
    d, u = s.finish(s1, s2, 200000 // nb, nb)
    s.write_3ddose(str(tmp_path / "ref.3ddose"), d, u)
    dims_r, bounds_r, dr, ur = read_3ddose(str(tmp_path / "ref.3ddose"))
    assert dims == dims_r and np.allclose(bounds, bounds_r)
    sel = dr > 0.1 * dr.max()
    ag = z_agreement(dg[sel], ug[sel] * dg[sel], dr[sel], ur[sel] * dr[sel], nb)
    print(f"{flags}: {fmt_agreement(ag)}")
    assert_agreement(ag, flags)
    assert abs(dg[sel].sum() / dr[sel].sum() - 1) < 0.02


def test_cli_reports_missing_input(tmp_path):
    r = subprocess.run([CLI, "-i", str(tmp_path / "absent"), "-o", "x"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode != 0
