"""ctypes bindings of the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.orc_session_open.restype = vp
        L.orc_session_open.argtypes = [C.c_char_p]
        L.orc_session_close.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_session_nreg.argtypes = [vp]
        L.orc_session_table.restype = vp
        L.orc_session_table.argtypes = [vp, C.c_char_p, lp, ip]
        L.orc_session_scalars.argtypes = [vp, dp]
        L.orc_session_pegs.argtypes = [vp, dp, ip]
        L.orc_session_case.argtypes = [vp, C.c_int, dp, C.c_int, dp, C.c_long, dp, C.c_int, lp]
        L.orc_session_dose.argtypes = [vp, C.c_int, C.c_int, dp, dp, ip]
        L.orc_session_dose_omp.restype = C.c_double
        L.orc_session_dose_omp.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_session_finish.argtypes = [vp, dp, dp, C.c_int, C.c_int, dp, dp]
        L.orc_session_ranmar.argtypes = [vp, C.c_long, ip]
        L.orc_ranmar_stream.argtypes = [C.c_int, C.c_int, C.c_long, C.c_long, ip]
        L.orc_session_matrad.argtypes = [vp, C.c_int, C.c_double, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int,
                                         C.c_double, ip, C.c_int, dp, dp, ip]
        L.orc_session_matrad_finish.argtypes = [vp, dp, dp, C.c_int, dp, dp]
        L.orc_session_write_3ddose.argtypes = [vp, C.c_char_p, dp, dp]
        L.orc_session_matrad_source.argtypes = [vp, C.c_int, C.c_double, dp, dp, C.c_int, ip, dp, dp, dp, dp,
                                                C.c_int, C.c_double, C.c_int, C.c_int, dp]
        L.orc_session_matrad_omp.restype = C.c_double
        L.orc_session_matrad_omp.argtypes = [vp, C.c_int, C.c_double, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int,
                                             C.c_double, ip, C.c_int, C.c_int, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def ranmar_stream(ixx, jxx, skip, n):
    out = np.zeros(n, np.int32)
    lib().orc_ranmar_stream(ixx, jxx, skip, n, _ip(out))
    return out


class Session:
    """An oracle run set up from a dosxyz input file stem (path without .inp)."""

    def __init__(self, stem):
        self.h = lib().orc_session_open(stem.encode())
        if not self.h:
            raise RuntimeError(f"oracle session {stem}: {lib().orc_last_error().decode()}")
        self.nreg = lib().orc_session_nreg(self.h)

    def close(self):
        if self.h:
            lib().orc_session_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def table(self, name):
        n = C.c_long(0)
        isint = C.c_int(0)
        p = lib().orc_session_table(self.h, name.encode(), C.byref(n), C.byref(isint))
        if not p:
            raise KeyError(name)
        t = C.c_int if isint.value else C.c_double
        arr = (t * n.value).from_address(p)
        return np.array(arr, dtype=np.int32 if isint.value else np.float64)

    def scalars(self):
        out = np.zeros(7)
        lib().orc_session_scalars(self.h, _dp(out))
        return out

    def pegs(self, nmed):
        f = np.zeros(12 * nmed)
        i = np.zeros(8 * nmed, np.int32)
        lib().orc_session_pegs(self.h, _dp(f), _ip(i))
        return f, i

    def case(self, kind, inp, rn):
        inp = np.ascontiguousarray(inp, np.float64)
        rn = np.ascontiguousarray(rn, np.float64)
        out = np.zeros(16 * 11 + 4 * 4096 + 16)
        used = C.c_long(0)
        k = lib().orc_session_case(self.h, kind, _dp(inp), len(inp), _dp(rn), len(rn), _dp(out), len(out),
                                   C.byref(used))
        return (out[:k] if k >= 0 else None), used.value

    def dose(self, nperbatch, nbatch, want_counts=False):
        s = np.zeros(self.nreg)
        s2 = np.zeros(self.nreg)
        cnt = np.zeros(nperbatch * nbatch, np.int32)
        rc = lib().orc_session_dose(self.h, nperbatch, nbatch, _dp(s), _dp(s2), _ip(cnt) if want_counts else None)
        if rc:
            raise RuntimeError(f"oracle dose run failed: {rc}")
        return (s, s2, cnt) if want_counts else (s, s2)

    def dose_omp(self, nperbatch, nbatch, nthreads=0):
        s = np.zeros(self.nreg)
        s2 = np.zeros(self.nreg)
        secs = lib().orc_session_dose_omp(self.h, nperbatch, nbatch, nthreads, _dp(s), _dp(s2))
        return s, s2, secs

    def finish(self, s, s2, nperbatch, nbatch):
        nv = self.nreg - 1
        dose, unc = np.zeros(nv), np.zeros(nv)
        lib().orc_session_finish(self.h, _dp(s), _dp(s2), nperbatch, nbatch, _dp(dose), _dp(unc))
        return dose, unc

    def write_3ddose(self, path, dose, unc):
        rc = lib().orc_session_write_3ddose(self.h, path.encode(), _dp(np.ascontiguousarray(dose)),
                                            _dp(np.ascontiguousarray(unc)))
        if rc:
            raise RuntimeError(f"3ddose writer failed: {rc}")

    def ranmar(self, n):
        out = np.zeros(n, np.int32)
        lib().orc_session_ranmar(self.h, n, _ip(out))
        return out

    def matrad(self, spectrum, bix, nper, nbatch, want_counts=False):
        """Serial matRad Dij transport (reference loop order).  spectrum: an
        object with energy / cdfinv1 / cdfinv2 (cdfinv1 None for mono); bix:
        dict with ibeam, src, corner, side1, side2 (float arrays), gaussian,
        sigma; nper: histories per batch of each beamlet.  Returns per-beamlet
        sums [nbix, nreg] and sums of squares (no carried variance), and the
        draws per history when asked."""
        nb = len(nper)
        s = np.zeros((nb, self.nreg))
        s2 = np.zeros((nb, self.nreg))
        nper = np.ascontiguousarray(nper, np.int32)
        cnt = np.zeros(int(nper.sum()) * nbatch, np.int32)
        sp = spectrum.cdfinv1 is not None
        c1 = np.ascontiguousarray(spectrum.cdfinv1 if sp else np.zeros(1000))
        c2 = np.ascontiguousarray(spectrum.cdfinv2 if sp else np.zeros(1000))
        arr = {k: np.ascontiguousarray(bix[k], np.float64) for k in ("src", "corner", "side1", "side2")}
        ib = np.ascontiguousarray(bix["ibeam"], np.int32)
        rc = lib().orc_session_matrad(self.h, int(sp), float(spectrum.energy), _dp(c1), _dp(c2), nb, _ip(ib),
                                      _dp(arr["src"]), _dp(arr["corner"]), _dp(arr["side1"]), _dp(arr["side2"]),
                                      int(bix.get("gaussian", 0)), float(bix.get("sigma", 0.0)), _ip(nper), nbatch,
                                      _dp(s), _dp(s2), _ip(cnt) if want_counts else None)
        if rc:
            raise RuntimeError(f"oracle matrad run failed: {rc}")
        return (s, s2, cnt) if want_counts else (s, s2)

    def matrad_source(self, spectrum, bix, ib, n):
        """n primaries of beamlet ib: rows iq ir e x y z u v w wt dnear."""
        out = np.zeros((n, 11))
        sp = spectrum.cdfinv1 is not None
        c1 = np.ascontiguousarray(spectrum.cdfinv1 if sp else np.zeros(1000))
        c2 = np.ascontiguousarray(spectrum.cdfinv2 if sp else np.zeros(1000))
        arr = {k: np.ascontiguousarray(bix[k], np.float64) for k in ("src", "corner", "side1", "side2")}
        ibm = np.ascontiguousarray(bix["ibeam"], np.int32)
        rc = lib().orc_session_matrad_source(self.h, int(sp), float(spectrum.energy), _dp(c1), _dp(c2), len(ibm),
                                             _ip(ibm), _dp(arr["src"]), _dp(arr["corner"]), _dp(arr["side1"]),
                                             _dp(arr["side2"]), int(bix.get("gaussian", 0)),
                                             float(bix.get("sigma", 0.0)), int(ib), int(n), _dp(out))
        if rc:
            raise RuntimeError(f"oracle source sampling failed: {rc}")
        return out

    def matrad_omp(self, spectrum, bix, nper, nbatch, nthreads=0):
        """OpenMP matRad transport (CPU timing); returns (sum, sum2, seconds)."""
        nb = len(nper)
        s = np.zeros((nb, self.nreg))
        s2 = np.zeros((nb, self.nreg))
        nper = np.ascontiguousarray(nper, np.int32)
        sp = spectrum.cdfinv1 is not None
        c1 = np.ascontiguousarray(spectrum.cdfinv1 if sp else np.zeros(1000))
        c2 = np.ascontiguousarray(spectrum.cdfinv2 if sp else np.zeros(1000))
        arr = {k: np.ascontiguousarray(bix[k], np.float64) for k in ("src", "corner", "side1", "side2")}
        ib = np.ascontiguousarray(bix["ibeam"], np.int32)
        secs = lib().orc_session_matrad_omp(self.h, int(sp), float(spectrum.energy), _dp(c1), _dp(c2), nb, _ip(ib),
                                            _dp(arr["src"]), _dp(arr["corner"]), _dp(arr["side1"]),
                                            _dp(arr["side2"]), int(bix.get("gaussian", 0)),
                                            float(bix.get("sigma", 0.0)), _ip(nper), nbatch, nthreads, _dp(s), _dp(s2))
        return s, s2, secs

    def matrad_dij(self, s, s2, nbatch, rel):
        """The reference's per-beamlet finish with the variance carried from
        beamlet to beamlet (omc_matrad.c:1603-1709), thresholded strictly above
        rel * max: CSC (jc, ir, val, var) with 0-based voxel rows."""
        nb, nreg = s.shape
        nv = nreg - 1
        carry = np.zeros(nreg)
        jc, irs, vals, vars_ = [0], [], [], []
        for b in range(nb):
            dose, var = np.zeros(nv), np.zeros(nv)
            s2b = np.ascontiguousarray(s2[b] + carry)
            lib().orc_session_matrad_finish(self.h, _dp(np.ascontiguousarray(s[b])), _dp(s2b), nbatch, _dp(dose),
                                            _dp(var))
            carry = np.concatenate([[0.0], var])
            th = dose.max() * rel
            idx = np.nonzero(dose > th)[0]
            irs.append(idx)
            vals.append(dose[idx])
            vars_.append(var[idx])
            jc.append(jc[-1] + len(idx))
        return (np.array(jc, np.int64), np.concatenate(irs).astype(np.int64), np.concatenate(vals),
                np.concatenate(vars_))
