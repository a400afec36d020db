"""The C ABI library loads on a CPU-only host, exports every entry point that
include/omc_gpu.h declares, fails cleanly without a device, and its host layer
(input formats) behaves like the reference's.  CPU only."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tests.conftest import REPO, golden

omc = pytest.importorskip("ompmc_amd")


def declared_functions():
    src = open(os.path.join(REPO, "include", "omc_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(omc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = omc.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.omc_api_version() == 1


def test_context_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(omc.EngineError) as e:
        omc.Context(["H2O521ICRU"])
    assert e.value.code == 6


def _host():
    L = omc.lib()
    L.omc_inp_parse.argtypes = [C.c_void_p, C.c_char_p]
    L.omc_inp_get.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]
    return L


def _inp(tmp_path, text):
    (tmp_path / "t.inp").write_text(text)
    buf = C.create_string_buffer(4 + 2 * 80 * 256 + 64)
    assert _host().omc_inp_parse(buf, str(tmp_path / "t").encode()) == 0
    def get(key):
        out = C.create_string_buffer(256)
        ok = _host().omc_inp_get(buf, key.encode(), out, 256)
        return out.value.decode() if ok else None
    return get


def test_inp_parser_reference_semantics(tmp_path):
    """omc_utilities.c: lines with '#' anywhere are dropped, keys match by
    substring (first match wins), values keep their leading blank."""
    get = _inp(tmp_path, "mono energy = 2.0\n# comment = 1\nncase = 100 # trailing\n"
                         "global ecut = 0.7\n\nspectrum file = a.spectrum\n")
    assert get("mono energy") == " 2.0"
    assert get("ncase") is None            # the line holds a '#'
    assert get("ecut") == " 0.7"           # substring match
    assert get("spectrum file") == " a.spectrum"
    assert get("absent") is None


def test_inp_single_key_file_reads_nothing(tmp_path):
    # the reference keeps count-1 items and treats 0 as "nothing parsed"
    get = _inp(tmp_path, "ncase = 100\n")
    assert get("ncase") is None


def test_egsphant_round_trip(tmp_path):
    P = omc.phantoms
    ph = P.small_test_phantom()
    path = str(tmp_path / "p.egsphant")
    P.write_egsphant(path, ph)
    back = P.read_egsphant(path)
    assert back.media == ph.media
    assert np.array_equal(back.med, ph.med)
    assert np.allclose(back.rho, ph.rho, rtol=0, atol=5e-7)   # written with %.6f
    assert np.allclose(back.xb, ph.xb)


def test_spectrum_inverse_cdf_matches_host_layer():
    """Python Spectrum.read builds the same inverse CDF as the C host layer
    used by the oracle and the CLI (omc_dosxyz.c:447-503)."""
    B = pytest.importorskip("oracle.bindings")
    s = omc.Spectrum.read(os.path.join(REPO, "data", "spectra", "mohan6.spectrum"))
    # the C layer is reached through the oracle library (same source file)
    L = B.lib()
    class Spec(C.Structure):
        _fields_ = [("spectrum", C.c_int), ("energy", C.c_double), ("deltak", C.c_double),
                    ("c1", C.c_double * 1000), ("c2", C.c_double * 1000)]
    L.omc_spectrum_read.argtypes = [C.POINTER(Spec), C.c_char_p]
    sp = Spec()
    assert L.omc_spectrum_read(C.byref(sp), os.path.join(REPO, "data", "spectra", "mohan6.spectrum").encode()) == 0
    assert np.allclose(np.array(sp.c1), s.cdfinv1, rtol=0, atol=1e-12)
    assert np.allclose(np.array(sp.c2), s.cdfinv2, rtol=0, atol=1e-12)
    # sampled bin frequencies follow the histogram up to the 1/1000 cell
    # quantisation of the reference's inverse CDF (each cell maps to one bin)
    rng = np.random.default_rng(5)
    n = 400000
    r1, r2 = rng.random(n), rng.random(n)
    k = np.minimum((1000 * r1).astype(int), 999)
    e = s.cdfinv1[k] + r2 * s.cdfinv2[k]
    edges = np.concatenate([[0.0], np.arange(0.25, 6.01, 0.25)])
    obs, _ = np.histogram(e, edges)
    w = np.loadtxt(os.path.join(REPO, "data", "spectra", "mohan6.spectrum"), skiprows=2)[:, 1] * 0.25
    p = w / w.sum()
    sigma = np.sqrt(p * (1 - p) / n)
    assert np.all(np.abs(obs / n - p) <= 1.0e-3 + 5 * sigma)


def test_python_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors in ompmc_amd/__init__.py have the size and field
    offsets of the C structs in include/omc_gpu.h (a probe compiled by gcc
    prints them)."""
    import ctypes as C
    import shutil
    import subprocess
    omc = pytest.importorskip("ompmc_amd")
    if not shutil.which("gcc"):
        pytest.skip("gcc absent")
    pairs = {"omc_spectrum_desc": omc.SpectrumDesc, "omc_phantom_desc": omc.PhantomDesc,
             "omc_dosxyz_source_desc": omc.DosxyzSourceDesc, "omc_run_opts": omc.RunOpts,
             "omc_run_stats": omc.RunStats, "omc_beamlets_desc": omc.BeamletsDesc, "omc_csc": omc.CSC,
             "omc_request": omc.Request, "omc_result": omc.Result}
    lines = ['#include <stdio.h>', '#include "omc_gpu.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I" + os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (l.split() for l in out if l)}
    for cname, py in pairs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
    # and the same members in the same order (a trailing int can hide in padding)
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(REPO, "include", "omc_gpu.h")).read(), flags=re.S)
    for cname, py in pairs.items():
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (cname, cname), hdr, re.S).group(1)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            parts = decl.split(",")
            names.append(re.findall(r"(\w+)\s*(?:\[[^\]]*\])?\s*$", parts[0])[0])
            names += [re.sub(r"[\s*]|\[.*\]", "", q) for q in parts[1:]]
        assert names == [f[0] for f in py._fields_], (cname, names)
