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
        L.orc_session_dose.argtypes = [vp, C.c_int, C.c_int, dp, dp, ip, dp]
        L.orc_session_dose_omp.restype = C.c_double
        L.orc_session_dose_omp.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_session_finish.argtypes = [vp, dp, dp, C.c_int, C.c_int, dp, dp]
        L.orc_session_ranmar.argtypes = [vp, C.c_long, ip]
        L.orc_ranmar_stream.argtypes = [C.c_int, C.c_int, C.c_long, C.c_long, ip]
        L.orc_session_matrad.argtypes = [vp, C.c_int, C.c_double, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int,
                                         C.c_double, ip, C.c_int, dp, dp, ip, dp]
        i64p = C.POINTER(C.c_int64)
        L.orc_mr_create.restype = vp
        L.orc_mr_create.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, C.c_int]
        L.orc_mr_destroy.argtypes = [vp]
        L.orc_mr_endep.restype = dp
        L.orc_mr_endep.argtypes = [vp]
        L.orc_mr_accum.restype = dp
        L.orc_mr_accum.argtypes = [vp, C.c_int]
        L.orc_mr_accum_endep.argtypes = [vp, C.c_int]
        L.orc_mr_accumulate.argtypes = [vp, C.c_int, C.c_int]
        L.orc_mr_beamlet_end.argtypes = [vp, C.c_int, C.c_double]
        L.orc_mr_nnz.restype = C.c_int64
        L.orc_mr_nnz.argtypes = [vp]
        L.orc_mr_copy.argtypes = [vp, i64p, i64p, dp, dp]
        L.orc_mr_run.restype = vp
        L.orc_mr_run.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, C.c_int, C.c_int, ip, dp, C.c_double]
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

    def dose(self, nperbatch, nbatch, want_counts=False, want_energy=False):
        """Serial dose run: (sum, sum2), then the draws per history and the
        energy balance [source, escape, lost, deposit] (MeV) when asked."""
        s = np.zeros(self.nreg)
        s2 = np.zeros(self.nreg)
        cnt = np.zeros(nperbatch * nbatch, np.int32)
        eb = np.zeros(4)
        rc = lib().orc_session_dose(self.h, nperbatch, nbatch, _dp(s), _dp(s2), _ip(cnt) if want_counts else None,
                                    _dp(eb))
        if rc:
            raise RuntimeError(f"oracle dose run failed: {rc}")
        out = (s, s2)
        if want_counts:
            out += (cnt,)
        if want_energy:
            out += (eb,)
        return out

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

    def matrad(self, spectrum, bix, nper, nbatch, want_counts=False, want_batches=False):
        """Serial matRad Dij transport (reference loop order).  spectrum: an
        object with energy / cdfinv1 / cdfinv2 (cdfinv1 None for mono); bix:
        dict with ibeam, src, corner, side1, side2 (float arrays), gaussian,
        sigma; nper: histories per batch of each beamlet.  Returns per-beamlet
        sums [nbix, nreg] and sums of squares (no carried variance), then the
        draws per history and the raw batch tallies [nbix, nbatch, nreg] when
        asked."""
        nb = len(nper)
        s = np.zeros((nb, self.nreg))
        s2 = np.zeros((nb, self.nreg))
        nper = np.ascontiguousarray(nper, np.int32)
        cnt = np.zeros(int(nper.sum()) * nbatch, np.int32)
        bat = np.zeros((nb, nbatch, self.nreg)) if want_batches else None
        sp = spectrum.cdfinv1 is not None
        c1 = np.ascontiguousarray(spectrum.cdfinv1 if sp else np.zeros(1000))
        c2 = np.ascontiguousarray(spectrum.cdfinv2 if sp else np.zeros(1000))
        arr = {k: np.ascontiguousarray(bix[k], np.float64) for k in ("src", "corner", "side1", "side2")}
        ib = np.ascontiguousarray(bix["ibeam"], np.int32)
        rc = lib().orc_session_matrad(self.h, int(sp), float(spectrum.energy), _dp(c1), _dp(c2), nb, _ip(ib),
                                      _dp(arr["src"]), _dp(arr["corner"]), _dp(arr["side1"]), _dp(arr["side2"]),
                                      int(bix.get("gaussian", 0)), float(bix.get("sigma", 0.0)), _ip(nper), nbatch,
                                      _dp(s), _dp(s2), _ip(cnt) if want_counts else None,
                                      _dp(bat) if want_batches else None)
        if rc:
            raise RuntimeError(f"oracle matrad run failed: {rc}")
        out = (s, s2)
        if want_counts:
            out += (cnt,)
        if want_batches:
            out += (bat,)
        return out

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


def _geom(ph):
    xb, yb, zb = (np.ascontiguousarray(a, np.float64) for a in (ph.xb, ph.yb, ph.zb))
    rho = np.ascontiguousarray(ph.rho, np.float64).reshape(-1)
    return len(xb) - 1, len(yb) - 1, len(zb) - 1, xb, yb, zb, rho


def matrad_dij(ph, batches, nper, nbatch, rel):
    """omc_matrad's beamlet loop (omc_matrad.c:1583-1702) over given batch
    tallies: accumEndep per batch, accumulateResults (Gy, the Q11 variance,
    rho < 0.044 zeroed), the strict threshold rel * max, the CSC append and
    the reference's carry of accum_endep2 into the next beamlet.  ph: an object
    with xb, yb, zb (boundaries) and rho (x fastest); batches [nbix, nbatch,
    nvox + 1] raw energy per batch (region 0 = outside); nper [nbix] histories
    per batch.  Returns (jc, ir, val, var) with 0-based voxel rows."""
    nx, ny, nz, xb, yb, zb, rho = _geom(ph)
    batches = np.ascontiguousarray(batches, np.float64)
    nb = batches.shape[0]
    assert batches.shape == (nb, nbatch, nx * ny * nz + 1)
    npr = np.ascontiguousarray(nper, np.int32)
    h = lib().orc_mr_run(nx, ny, nz, _dp(xb), _dp(yb), _dp(zb), _dp(rho), nb, nbatch, _ip(npr), _dp(batches),
                         float(rel))
    if not h:
        raise MemoryError("oracle Dij assembly")
    try:
        n = int(lib().orc_mr_nnz(h))
        jc, ir = np.zeros(nb + 1, np.int64), np.zeros(max(n, 1), np.int64)
        val, var = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        i64p = C.POINTER(C.c_int64)
        lib().orc_mr_copy(h, jc.ctypes.data_as(i64p), ir.ctypes.data_as(i64p), _dp(val), _dp(var))
    finally:
        lib().orc_mr_destroy(h)
    return jc, ir[:n], val[:n], var[:n]


class MatradScore:
    """Step-by-step access to the restated omc_matrad scoring (the closed-form
    tests drive accumEndep / accumulateResults / the beamlet end one call at a
    time and read the accumulators between them)."""

    def __init__(self, ph, nbeamlets):
        nx, ny, nz, xb, yb, zb, rho = _geom(ph)
        self.nreg = nx * ny * nz + 1
        self.nb = nbeamlets
        self.h = lib().orc_mr_create(nx, ny, nz, _dp(xb), _dp(yb), _dp(zb), _dp(rho), nbeamlets)
        if not self.h:
            raise MemoryError("oracle Dij scoring")

    def close(self):
        if self.h:
            lib().orc_mr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _view(self, p):
        return np.ctypeslib.as_array(p, shape=(self.nreg,))

    def set_endep(self, e):
        self._view(lib().orc_mr_endep(self.h))[:] = e

    def accum(self):
        """copies of (accum_endep, accum_endep2)"""
        return self._view(lib().orc_mr_accum(self.h, 0)).copy(), self._view(lib().orc_mr_accum(self.h, 1)).copy()

    def accum_endep(self, nperbatch):
        lib().orc_mr_accum_endep(self.h, int(nperbatch))

    def accumulate(self, nbatch, iout=1):
        lib().orc_mr_accumulate(self.h, int(iout), int(nbatch))

    def beamlet_end(self, ib, rel):
        if lib().orc_mr_beamlet_end(self.h, int(ib), float(rel)):
            raise MemoryError("oracle Dij assembly")

    def csc(self):
        n = int(lib().orc_mr_nnz(self.h))
        jc, ir = np.zeros(self.nb + 1, np.int64), np.zeros(max(n, 1), np.int64)
        val, var = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        i64p = C.POINTER(C.c_int64)
        lib().orc_mr_copy(self.h, jc.ctypes.data_as(i64p), ir.ctypes.data_as(i64p), _dp(val), _dp(var))
        return jc, ir[:n], val[:n], var[:n]
