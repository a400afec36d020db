"""MI355X transport engine for voxelized photon/electron Monte Carlo dose.

Python host mirror of the reference's two user codes over the C ABI in
include/omc_gpu.h (ompmc_amd/libompmc.so):

* ``Context``  -- physics tables for a set of PEGS4 media, resident on the GPU
                  (reference initMediaData, src/ompmc.c:5468-5505);
* ``Context.phantom`` -- a voxel phantom (reference initPhantom/initRegions);
* ``Context.run_dosxyz`` -- omc_dosxyz dose run (ucodes/omc_dosxyz/omc_dosxyz.c:1072-1307);
* ``Context.run_dij`` -- omc_matrad dose-influence matrix
                  (ucodes/omc_matrad/omc_matrad.c:1417-1789);
* ``read_inp`` / ``DosxyzJob`` -- the .inp front end of omc_dosxyz.

The library must be built (``__graft_entry__.build()``); importing never falls
back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import phantoms  # noqa: F401  (re-export)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libompmc.so")
REPO = os.path.dirname(HERE)
DATA = os.path.join(REPO, "data")

FP32, FP64 = 32, 64
NOALLOC_MAX_REGIONS = 4194304    # OMC_NOALLOC_MAX_REGIONS
FORM_HISTORY, FORM_EVENT = 0, 1
RNG_STREAMS, RNG_REPLAY = 0, 1
INVDIM = 1000

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_i64 = C.POINTER(C.c_int64)
_u32 = C.POINTER(C.c_uint32)


class SpectrumDesc(C.Structure):
    _fields_ = [("spectrum", C.c_int), ("energy", C.c_double), ("cdfinv1", _d), ("cdfinv2", _d)]


class PhantomDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("xb", _d), ("yb", _d), ("zb", _d),
                ("med", _i), ("rho", _d), ("ecut", C.c_double), ("pcut", C.c_double)]


class DosxyzSourceDesc(C.Structure):
    _fields_ = [("spec", SpectrumDesc), ("charge", C.c_int), ("ssd", C.c_double),
                ("collimator", C.c_double * 4)]


class RunOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("form", C.c_int), ("rng", C.c_int), ("seed", C.c_uint64),
                ("nsplit", C.c_int), ("replay", _u32), ("replay_offsets", _i64), ("replay_len", C.c_int64),
                ("grid", C.c_int), ("block", C.c_int), ("count_events", C.c_int), ("event_target", C.c_int64),
                ("electron_budget", C.c_int), ("group", C.c_int), ("hbm_budget", C.c_int64),
                ("voxel_walk", C.c_int), ("electron_turns", C.c_int), ("batches_per_job", C.c_int),
                ("streams", C.c_int), ("max_particle_steps", C.c_int64), ("csc_on_device", C.c_int),
                ("graph", C.c_int), ("lds_tables", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("gpu_seconds", C.c_double), ("wall_seconds", C.c_double), ("histories", C.c_double),
                ("n_howfar", C.c_uint64), ("n_hownear", C.c_uint64), ("n_cross", C.c_uint64),
                ("n_dep", C.c_uint64), ("n_rn", C.c_uint64), ("n_photon", C.c_uint64),
                ("n_electron", C.c_uint64), ("errors", C.c_int),
                ("kernel_seconds", C.c_double * 5), ("kernel_launches", C.c_uint64 * 5),
                ("ev_howfar", C.c_uint64 * 2), ("ev_hownear", C.c_uint64 * 2), ("ev_cross", C.c_uint64 * 2),
                ("ev_dep", C.c_uint64 * 2), ("events_photon", C.c_uint64), ("events_charged", C.c_uint64),
                ("rounds", C.c_uint64), ("peak_inflight", C.c_uint64), ("groups", C.c_int64),
                ("group_size", C.c_int64), ("nnz", C.c_int64), ("batches_per_job", C.c_int64),
                ("streams", C.c_int64), ("e_source", C.c_double), ("e_escape", C.c_double),
                ("e_lost", C.c_double), ("e_deposit", C.c_double),
                ("kernel_busy_seconds", C.c_double * 5), ("graph_replays", C.c_uint64),
                ("graph_fallbacks", C.c_uint64), ("lds_table_bytes", C.c_uint64)]

    def energy_residual(self):
        """(deposit + escape + lost - source) / source of the run"""
        return (self.e_deposit + self.e_escape + self.e_lost - self.e_source) / self.e_source

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if hasattr(v, "_length_") else v
        return out


class BeamletsDesc(C.Structure):
    _fields_ = [("nbixels", C.c_int), ("nbeams", C.c_int), ("ibeam", _i), ("src", _d), ("corner", _d),
                ("side1", _d), ("side2", _d), ("gaussian", C.c_int), ("sigma", C.c_double),
                ("nhist", _i64), ("nhist_each", C.c_int64)]


class CSC(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64), ("jc", _i64), ("ir", _i64),
                ("val", _d), ("device", C.c_int)]

    def to_numpy(self):
        """(jc, ir, val) copies on the host (downloaded when the matrix is in
        device memory); the library arrays stay owned by the struct."""
        jc = np.ctypeslib.as_array(self.jc, shape=(self.ncols + 1,)).copy()
        n = int(self.nnz)
        if self.device == 1:
            ir, val = np.zeros(n, np.int64), np.zeros(n)
            if n:
                _check(lib().omc_csc_download(C.byref(self), ir.ctypes.data_as(_i64), _dp(val)))
            return jc, ir, val
        ir = np.ctypeslib.as_array(self.ir, shape=(n,)).copy() if n else np.zeros(0, np.int64)
        val = np.ctypeslib.as_array(self.val, shape=(n,)).copy() if n else np.zeros(0)
        return jc, ir, val


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("phantom", PhantomDesc), ("dose_src", DosxyzSourceDesc),
                ("nperbatch", C.c_int64), ("spec", SpectrumDesc), ("bix", BeamletsDesc),
                ("rel_threshold", C.c_double), ("nbatch", C.c_int), ("seed", C.c_uint64), ("precision", C.c_int),
                ("nsplit", C.c_int)]


class Result(C.Structure):
    _fields_ = [("state", C.c_int), ("error", C.c_int), ("sum", _d), ("sum2", _d), ("nreg", C.c_int64),
                ("dij", CSC), ("submit_t", C.c_double), ("start_t", C.c_double), ("end_t", C.c_double),
                ("histories", C.c_double), ("message", C.c_char * 256)]


_lib = None


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"omc error {code}: {msg}")
        self.code = code


def lib():
    """Loads ompmc_amd/libompmc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.omc_last_error.restype = C.c_char_p
    L.omc_atomic_probe.argtypes = [C.c_int, C.c_int64, C.c_int64, C.POINTER(C.c_double)]
    L.omc_csc_download.argtypes = [C.POINTER(CSC), _i64, _d]
    L.omc_api_version.restype = C.c_int
    L.omc_ctx_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_char_p,
                                 C.c_char_p, C.c_int]
    L.omc_ctx_destroy.argtypes = [vp]
    L.omc_ctx_table.argtypes = [vp, C.c_char_p, C.POINTER(_d), C.POINTER(C.c_long)]
    L.omc_phantom_upload.argtypes = [vp, C.POINTER(PhantomDesc), _i]
    L.omc_phantom_release.argtypes = [vp, C.c_int]
    L.omc_run_opts_default.argtypes = [C.POINTER(RunOpts)]
    L.omc_run_dosxyz.argtypes = [vp, C.c_int, C.POINTER(DosxyzSourceDesc), C.c_int64, C.c_int,
                                 C.POINTER(RunOpts), _d, _d, C.POINTER(RunStats)]
    L.omc_run_dosxyz_noalloc.argtypes = [vp, C.c_int, C.POINTER(DosxyzSourceDesc), C.c_int64, C.c_int,
                                         C.POINTER(RunOpts), vp, vp, C.POINTER(RunStats)]
    L.omc_dosxyz_dose.argtypes = [vp, C.c_int, _d, _d, C.c_int64, C.c_int, _d, _d]
    L.omc_write_3ddose_file.argtypes = [vp, C.c_int, C.c_char_p, _d, _d]
    L.omc_sampler_eval.argtypes = [vp, C.c_int, C.c_int, C.c_int, _i, _d, C.c_int, _u32, _i64, _d, C.c_int, _i64]
    L.omc_run_dij.argtypes = [vp, C.c_int, C.POINTER(SpectrumDesc), C.POINTER(BeamletsDesc), C.c_int, C.c_double,
                              C.POINTER(RunOpts), C.POINTER(CSC), C.POINTER(CSC), C.POINTER(RunStats)]
    L.omc_run_dij_ex.argtypes = [vp, C.c_int, C.POINTER(SpectrumDesc), C.POINTER(BeamletsDesc), C.c_int,
                                 C.c_double, C.POINTER(RunOpts), C.POINTER(CSC), C.POINTER(CSC), _d, _d,
                                 C.POINTER(RunStats)]
    L.omc_csc_free.argtypes = [C.POINTER(CSC)]
    L.omc_dij_assemble.argtypes = [vp, C.c_int, C.c_int, C.c_int, _i64, _d, C.c_double, C.POINTER(RunOpts),
                                   C.POINTER(CSC), C.POINTER(CSC), C.POINTER(RunStats)]
    L.omc_server_create.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(vp)]
    L.omc_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
    L.omc_server_poll.argtypes = [vp, C.c_int64, _i]
    L.omc_server_wait.argtypes = [vp, C.c_int64, C.POINTER(Result)]
    L.omc_result_free.argtypes = [C.POINTER(Result)]
    L.omc_server_destroy.argtypes = [vp]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise EngineError(rc, lib().omc_last_error().decode(errors="replace"))


def _dp(a):
    return a.ctypes.data_as(_d)


def _ip(a):
    return a.ctypes.data_as(_i)


def default_paths():
    return (os.path.join(DATA, "pegs4", "521icru.pegs4dat"), os.path.join(DATA, "pegs4", "pgs4form.dat"),
            DATA + "/")


def run_opts(precision=FP32, form=FORM_EVENT, seed=1802, rng=RNG_STREAMS, replay=None, replay_offsets=None,
             grid=0, block=0, count_events=False, nsplit=1, event_target=0, electron_budget=0, group=0,
             hbm_budget=0, voxel_walk=False, electron_turns=0, batches_per_job=0, streams=0, max_particle_steps=0,
             csc_on_device=False, graph=True, lds_tables=True):
    o = RunOpts()
    lib().omc_run_opts_default(C.byref(o))
    o.precision, o.form, o.seed, o.rng, o.nsplit = precision, form, seed, rng, nsplit
    o.grid, o.block = grid, block
    o.count_events = 2 if count_events == "time" else int(bool(count_events))
    o.event_target, o.electron_budget, o.group, o.hbm_budget = event_target, electron_budget, group, hbm_budget
    o.voxel_walk = int(bool(voxel_walk))
    o.electron_turns, o.batches_per_job, o.streams = electron_turns, batches_per_job, streams
    o.max_particle_steps = max_particle_steps
    o.csc_on_device = int(bool(csc_on_device))
    o.graph = 0 if graph else -1
    o.lds_tables = 0 if lds_tables else -1
    keep = []
    if replay is not None:
        rp = np.ascontiguousarray(replay, np.uint32)
        ro = np.ascontiguousarray(replay_offsets, np.int64)
        keep = [rp, ro]
        o.replay = rp.ctypes.data_as(_u32)
        o.replay_offsets = ro.ctypes.data_as(_i64)
        o.replay_len = len(rp)
    return o, keep


@dataclass
class Spectrum:
    """Mono energy (``cdfinv1 is None``) or the 1000-bin inverse CDF of a
    spectrum file (reference omc_dosxyz.c:396-510)."""
    energy: float = 0.0
    cdfinv1: np.ndarray = None
    cdfinv2: np.ndarray = None

    @staticmethod
    def mono(e):
        return Spectrum(energy=float(e))

    @staticmethod
    def from_histogram(e_upper, weights, enmin=0.0, imode=0):
        """Inverse CDF exactly as the reference builds it."""
        e_upper = np.asarray(e_upper, float)
        pdf = np.asarray(weights, float).copy()
        if imode == 1:
            pdf[0] *= e_upper[0] - enmin
            pdf[1:] *= e_upper[1:] - e_upper[:-1]
        cdf = np.cumsum(pdf)
        cdf = cdf * (1.0 / cdf[-1])
        c1 = np.zeros(INVDIM)
        c2 = np.zeros(INVDIM)
        gridsz = 1.0 / INVDIM
        for k in range(INVDIM):
            ak = k * gridsz
            i = int(np.searchsorted(cdf, ak, side="left"))
            i = min(i, len(cdf) - 1)
            c1[k] = e_upper[i - 1] if i != 0 else enmin
            c2[k] = e_upper[i] - c1[k]
        return Spectrum(cdfinv1=c1, cdfinv2=c2)

    @staticmethod
    def read(path):
        with open(path) as f:
            f.readline()
            n, enmin, imode = f.readline().split()[:3]
            rows = [f.readline().split()[:2] for _ in range(int(n))]
        e = [float(r[0]) for r in rows]
        w = [float(r[1]) for r in rows]
        return Spectrum.from_histogram(e, w, float(enmin), int(imode))

    def desc(self):
        d = SpectrumDesc()
        if self.cdfinv1 is None:
            d.spectrum, d.energy = 0, self.energy
        else:
            self._c1 = np.ascontiguousarray(self.cdfinv1, float)
            self._c2 = np.ascontiguousarray(self.cdfinv2, float)
            d.spectrum = 1
            d.cdfinv1, d.cdfinv2 = _dp(self._c1), _dp(self._c2)
        return d


class Phantom:
    """A phantom resident on the context's device."""

    def __init__(self, ctx, ph: "phantoms.Phantom", ecut=0.7, pcut=0.01):
        self.ctx, self.ph = ctx, ph
        self.xb = np.ascontiguousarray(ph.xb, float)
        self.yb = np.ascontiguousarray(ph.yb, float)
        self.zb = np.ascontiguousarray(ph.zb, float)
        self.med = np.ascontiguousarray(ph.med.reshape(-1), np.int32)
        self.rho = np.ascontiguousarray(ph.rho.reshape(-1), float)
        d = PhantomDesc(len(self.xb) - 1, len(self.yb) - 1, len(self.zb) - 1, _dp(self.xb), _dp(self.yb),
                        _dp(self.zb), _ip(self.med), _dp(self.rho), ecut, pcut)
        pid = C.c_int(0)
        _check(lib().omc_phantom_upload(ctx.h, C.byref(d), C.byref(pid)))
        self.id = pid.value
        self.nreg = self.med.size + 1

    def release(self):
        if self.id is not None and self.ctx.h:
            lib().omc_phantom_release(self.ctx.h, self.id)
            self.id = None


class Context:
    """Physics tables for the given PEGS4 media, uploaded to GPU `device`."""

    def __init__(self, media, pegs_file=None, pgs4form_file=None, data_folder=None, device=0):
        pf, ff, df = default_paths()
        arr = (C.c_char_p * len(media))(*[m.encode() for m in media])
        h = C.c_void_p()
        _check(lib().omc_ctx_create(C.byref(h), len(media), arr, (pegs_file or pf).encode(),
                                    (pgs4form_file or ff).encode(), (data_folder or df).encode(), device))
        self.h = h
        self.media = list(media)

    def close(self):
        if self.h:
            lib().omc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def table(self, name):
        p = _d()
        n = C.c_long(0)
        _check(lib().omc_ctx_table(self.h, name.encode(), C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def phantom(self, ph, ecut=0.7, pcut=0.01):
        return Phantom(self, ph, ecut, pcut)

    def run_dosxyz(self, phantom: Phantom, spectrum: Spectrum, nperbatch, nbatch, collimator=(-5, 5, -5, 5),
                   ssd=1.0e10, charge=0, keep_diverged_replay=False, **opts):
        """Returns (sum, sum2, stats): per-region batch sums of deposited
        energy (MeV, region 0 = outside) and run statistics.  A replayed run in
        which a history needed more draws than it was given fails with
        OMC_ERR_ARG (stats.errors bit 1); keep_diverged_replay=True returns its
        outputs instead, for verification runs that count such histories."""
        src = DosxyzSourceDesc()
        src.spec = spectrum.desc()
        src.charge, src.ssd = charge, ssd
        for k in range(4):
            src.collimator[k] = collimator[k]
        o, keep = run_opts(**opts)
        s = np.zeros(phantom.nreg)
        s2 = np.zeros(phantom.nreg)
        st = RunStats()
        rc = lib().omc_run_dosxyz(self.h, phantom.id, C.byref(src), nperbatch, nbatch, C.byref(o), _dp(s), _dp(s2),
                                  C.byref(st))
        if not (keep_diverged_replay and rc == 5 and st.errors == 2):
            _check(rc)
        del keep
        return s, s2, st

    def run_dosxyz_noalloc(self, phantom: Phantom, spectrum: Spectrum, nperbatch, nbatch, d_sum, d_sum2,
                           collimator=(-5, 5, -5, 5), ssd=1.0e10, charge=0, **opts):
        """omc_run_dosxyz_noalloc: the history-form dose run with no device
        allocation; d_sum and d_sum2 are device addresses (int) of nreg
        doubles on the context's device that receive the batch sums.
        Returns the run statistics."""
        src = DosxyzSourceDesc()
        src.spec = spectrum.desc()
        src.charge, src.ssd = charge, ssd
        for k in range(4):
            src.collimator[k] = collimator[k]
        opts.setdefault("form", FORM_HISTORY)
        o, keep = run_opts(**opts)
        st = RunStats()
        _check(lib().omc_run_dosxyz_noalloc(self.h, phantom.id, C.byref(src), nperbatch, nbatch, C.byref(o),
                                            C.c_void_p(d_sum), C.c_void_p(d_sum2), C.byref(st)))
        del keep
        return st

    def dose(self, phantom: Phantom, s, s2, nperbatch, nbatch):
        nv = phantom.nreg - 1
        dose, unc = np.zeros(nv), np.zeros(nv)
        _check(lib().omc_dosxyz_dose(self.h, phantom.id, _dp(np.ascontiguousarray(s)), _dp(np.ascontiguousarray(s2)),
                                     nperbatch, nbatch, _dp(dose), _dp(unc)))
        return dose, unc

    def run_dij(self, phantom: Phantom, spectrum: Spectrum, bix: "phantoms.Beamlets", nbatch=10, rel_threshold=1e-4,
                nhist_each=0, want_var=False, want_plan=False, gaussian=False, sigma=0.0, download=True, **opts):
        """omc_matrad dose-influence matrix: returns (dij, var, plan, stats)
        with dij/var as (jc, ir, val) (var None unless asked; 0-based voxel
        rows) and plan = (beamlet-summed dose, its variance) when asked.
        With csc_on_device=True the matrices stay in device memory; they are
        copied back unless download=False, which returns (jc, None, None)."""
        keep = []
        b = BeamletsDesc()
        b.nbixels, b.nbeams = bix.n, len(bix.src)
        arrs = dict(ibeam=np.ascontiguousarray(bix.ibeam, np.int32),
                    src=np.ascontiguousarray(bix.src, float).reshape(-1),
                    corner=np.ascontiguousarray(bix.corner, float).reshape(-1),
                    side1=np.ascontiguousarray(bix.side1, float).reshape(-1),
                    side2=np.ascontiguousarray(bix.side2, float).reshape(-1))
        keep += list(arrs.values())
        b.ibeam = _ip(arrs["ibeam"])
        b.src, b.corner, b.side1, b.side2 = (_dp(arrs[k]) for k in ("src", "corner", "side1", "side2"))
        b.gaussian, b.sigma = int(bool(gaussian)), float(sigma)
        if nhist_each:
            b.nhist_each = int(nhist_each)
        else:
            nh = np.ascontiguousarray(bix.nhist, np.int64)
            keep.append(nh)
            b.nhist = nh.ctypes.data_as(_i64)
        o, k2 = run_opts(**opts)
        keep += k2
        sp = spectrum.desc()
        dij, var = CSC(), CSC()
        nv = phantom.nreg - 1
        pd = np.zeros(nv) if want_plan else None
        pv = np.zeros(nv) if want_plan else None
        st = RunStats()
        _check(lib().omc_run_dij_ex(self.h, phantom.id, C.byref(sp), C.byref(b), nbatch, rel_threshold, C.byref(o),
                                    C.byref(dij), C.byref(var) if want_var else None,
                                    _dp(pd) if want_plan else None, _dp(pv) if want_plan else None, C.byref(st)))
        def take(m):
            if m.device == 1 and not download:
                r = (np.ctypeslib.as_array(m.jc, shape=(m.ncols + 1,)).copy(), None, None)
            else:
                r = m.to_numpy()
            lib().omc_csc_free(C.byref(m))
            return r
        out = take(dij)
        vout = take(var) if want_var else None
        del keep
        return out, vout, ((pd, pv) if want_plan else None), st

    def dij_assemble(self, phantom: Phantom, endep, nperbatch, rel_threshold=1e-4, want_var=True, **opts):
        """omc_dij_assemble: the Dij scoring and assembly of the device kernels
        on given batch tallies endep [nbix, nbatch, nvox + 1] (MeV, region 0
        outside) with nperbatch [nbix] histories per batch.  Returns (dij, var,
        stats) with dij/var as (jc, ir, val)."""
        e = np.ascontiguousarray(endep, np.float64)
        nb, nbatch, nreg = e.shape
        if nreg != phantom.nreg:
            raise ValueError(f"tallies of {nreg} regions for a phantom of {phantom.nreg}")
        npb = np.ascontiguousarray(nperbatch, np.int64)
        o, keep = run_opts(**opts)
        dij, var, st = CSC(), CSC(), RunStats()
        _check(lib().omc_dij_assemble(self.h, phantom.id, nb, nbatch, npb.ctypes.data_as(_i64), _dp(e),
                                      float(rel_threshold), C.byref(o), C.byref(dij),
                                      C.byref(var) if want_var else None, C.byref(st)))
        del keep
        out = dij.to_numpy()
        lib().omc_csc_free(C.byref(dij))
        vout = None
        if want_var:
            vout = var.to_numpy()
            lib().omc_csc_free(C.byref(var))
        return out, vout, st

    def sampler_eval(self, phantom: Phantom, precision, kinds, inputs, rn_lists, in_stride=24, out_stride=64 + 2 * 512):
        """Evaluates golden-vector cases with injected random numbers on the
        GPU.  rn_lists: per case, the draws as floats k*2^-24.  Output row:
        [0] stack depth change (-1, 0, 1; 2 for scalar routines), [1:12] and
        [12:23] the particles (iq ir e x y z u v w wt -), [23] tally calls,
        [24] howfar and [25] hownear calls, [64:] (region, wt*edep) per tally."""
        n = len(kinds)
        kind = np.ascontiguousarray(kinds, np.int32)
        inp = np.zeros((n, in_stride))
        for i, v in enumerate(inputs):
            inp[i, :len(v)] = v
        off = np.zeros(n + 1, np.int64)
        off[1:] = np.cumsum([len(r) for r in rn_lists])
        rn = np.concatenate([np.rint(np.asarray(r) * 16777216.0) for r in rn_lists]).astype(np.uint32) \
            if off[-1] else np.zeros(1, np.uint32)
        out = np.zeros((n, out_stride))
        used = np.zeros(n, np.int64)
        _check(lib().omc_sampler_eval(self.h, phantom.id, precision, n, _ip(kind), _dp(inp), in_stride,
                                      rn.ctypes.data_as(_u32), off.ctypes.data_as(_i64), _dp(out), out_stride,
                                      used.ctypes.data_as(_i64)))
        return out, used


def atomic_probe(nslot, nops, device=0):
    """FP64 atomic adds per second at uniformly scattered addresses of a
    table of nslot doubles (omc_atomic_probe)."""
    r = C.c_double(0)
    _check(lib().omc_atomic_probe(int(device), int(nslot), int(nops), C.byref(r)))
    return r.value


class Server:
    """Resident request loop on a context's device (omc_server_*): requests are
    admitted in arrival order while their device memory fits the budget and
    fewer than ``max_inflight`` run; admitted requests run concurrently."""

    def __init__(self, ctx: Context, hbm_budget=0, max_inflight=4):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().omc_server_create(ctx.h, int(hbm_budget), int(max_inflight), C.byref(h)))
        self.h = h
        self._keep = {}

    def _phantom_desc(self, ph, ecut, pcut, keep):
        xb, yb, zb = (np.ascontiguousarray(a, float) for a in (ph.xb, ph.yb, ph.zb))
        med = np.ascontiguousarray(ph.med.reshape(-1), np.int32)
        rho = np.ascontiguousarray(ph.rho.reshape(-1), float)
        keep += [xb, yb, zb, med, rho]
        return PhantomDesc(len(xb) - 1, len(yb) - 1, len(zb) - 1, _dp(xb), _dp(yb), _dp(zb), _ip(med), _dp(rho),
                           ecut, pcut)

    def submit_dose(self, ph, spectrum, nperbatch, nbatch=10, collimator=(-5, 5, -5, 5), ssd=1.0e10, charge=0,
                    seed=1802, precision=FP32, ecut=0.7, pcut=0.01, nsplit=1):
        keep = []
        r = Request()
        r.kind = 0
        r.nsplit = int(nsplit)
        r.phantom = self._phantom_desc(ph, ecut, pcut, keep)
        r.dose_src.spec = spectrum.desc()
        keep.append(spectrum)
        r.dose_src.charge, r.dose_src.ssd = charge, ssd
        for k in range(4):
            r.dose_src.collimator[k] = collimator[k]
        r.nperbatch, r.nbatch, r.seed, r.precision = int(nperbatch), int(nbatch), int(seed), precision
        rid = C.c_int64(0)
        _check(lib().omc_server_submit(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def submit_dij(self, ph, spectrum, bix, nbatch=10, rel_threshold=1e-4, seed=1802, precision=FP32, ecut=0.7,
                   pcut=0.01, nhist_each=0, nsplit=1):
        keep = []
        r = Request()
        r.kind = 1
        r.nsplit = int(nsplit)
        r.phantom = self._phantom_desc(ph, ecut, pcut, keep)
        r.spec = spectrum.desc()
        keep.append(spectrum)
        arrs = [np.ascontiguousarray(bix.ibeam, np.int32), np.ascontiguousarray(bix.src, float).reshape(-1),
                np.ascontiguousarray(bix.corner, float).reshape(-1), np.ascontiguousarray(bix.side1, float).reshape(-1),
                np.ascontiguousarray(bix.side2, float).reshape(-1)]
        keep += arrs
        r.bix.nbixels, r.bix.nbeams = bix.n, len(bix.src)
        r.bix.ibeam = _ip(arrs[0])
        r.bix.src, r.bix.corner, r.bix.side1, r.bix.side2 = (_dp(a) for a in arrs[1:])
        if nhist_each:
            r.bix.nhist_each = int(nhist_each)
        else:
            nh = np.ascontiguousarray(bix.nhist, np.int64)
            keep.append(nh)
            r.bix.nhist = nh.ctypes.data_as(_i64)
        r.rel_threshold, r.nbatch, r.seed, r.precision = rel_threshold, int(nbatch), int(seed), precision
        rid = C.c_int64(0)
        _check(lib().omc_server_submit(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def poll(self, rid):
        st = C.c_int(0)
        _check(lib().omc_server_poll(self.h, rid, C.byref(st)))
        return st.value

    def wait(self, rid):
        """Returns a dict: state, error, times (submit/start/end on the server
        clock), histories, and sum/sum2 (dose requests) or dij=(jc, ir, val)."""
        res = Result()
        rc = lib().omc_server_wait(self.h, rid, C.byref(res))
        out = dict(state=res.state, error=res.error, submit_t=res.submit_t, start_t=res.start_t, end_t=res.end_t,
                   histories=res.histories)
        if rc == 0:
            if res.sum:
                n = int(res.nreg)
                out["sum"] = np.ctypeslib.as_array(res.sum, shape=(n,)).copy()
                out["sum2"] = np.ctypeslib.as_array(res.sum2, shape=(n,)).copy()
            if res.dij.jc:
                out["dij"] = res.dij.to_numpy()
        else:
            out["message"] = res.message.decode(errors="replace")
        lib().omc_result_free(C.byref(res))
        return out

    def close(self):
        if self.h:
            lib().omc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
