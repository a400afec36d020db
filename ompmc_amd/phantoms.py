"""Synthetic inputs in the reference's own file formats.

Writers for the egsphant phantom (reference ``ucodes/omc_dosxyz/omc_dosxyz.c:62-175``),
the spectrum file (``omc_dosxyz.c:396-503``) and the ``key = value`` input file
(``src/omc_utilities.c:63-161``), plus the seeded phantoms, spectra and beamlet
sets of the benchmark workloads:

* ``water_cube``       60^3 water at 0.5 cm (2 MeV parallel photon beam case)
* ``slab_phantom``     128^3 water/lung/bone/water slabs at 0.3 cm
* ``synthetic_ct``     elliptic water body with lungs, a bone cylinder and a
                       spherical target, any grid (the beamlet-matrix cases)
* ``beamlets``         matRad-style bixel sets (corner/side vectors per bixel)
* ``synthetic_6mv``    24-bin bremsstrahlung-like spectrum peaking near 1.5 MeV
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

WATER, LUNG, BONE, AIR = "H2O521ICRU", "LUNG521ICRU", "ICRPBONE521ICRU", "AIR521ICRU"

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(REPO, "data")


@dataclass
class Phantom:
    """Voxel phantom: media names, boundaries (cm) and per-voxel arrays in
    (z, y, x) order with x fastest, i.e. region ``1 + ix + iy*nx + iz*nx*ny``."""
    media: list
    xb: np.ndarray
    yb: np.ndarray
    zb: np.ndarray
    med: np.ndarray          # int32, 1-based medium index, 0 = vacuum
    rho: np.ndarray          # float64 g/cm^3
    name: str = "phantom"

    @property
    def shape(self):
        return (len(self.xb) - 1, len(self.yb) - 1, len(self.zb) - 1)

    @property
    def nvox(self):
        nx, ny, nz = self.shape
        return nx * ny * nz


def write_egsphant(path: str, ph: Phantom) -> None:
    """egsphant text: nmed, media names, a dummy line, ``I J K``, the three
    boundary lists, I single-digit media per line (blank line per slice), then
    densities (blank line per slice)."""
    nx, ny, nz = ph.shape
    if len(ph.media) > 9:
        raise ValueError("egsphant holds at most 9 media")
    lines = [f"{len(ph.media)}"] + list(ph.media) + ["  0.7000000     " * len(ph.media)]
    lines.append(f"  {nx}  {ny}  {nz}")
    for b in (ph.xb, ph.yb, ph.zb):
        lines.append(" ".join(f"{v:.6f}" for v in b))
    med = ph.med.reshape(nz, ny, nx)
    for k in range(nz):
        for j in range(ny):
            lines.append("".join(str(int(c)) for c in med[k, j]))
        lines.append("")
    rho = ph.rho.reshape(nz, ny, nx)
    for k in range(nz):
        for j in range(ny):
            lines.append(" ".join(f"{v:.6f}" for v in rho[k, j]))
        lines.append("")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_egsphant(path: str) -> Phantom:
    with open(path) as f:
        nmed = int(f.readline())
        media = [f.readline().strip() for _ in range(nmed)]
        f.readline()
        nx, ny, nz = (int(t) for t in f.readline().split())
        rest = f.read()
    toks = rest.split()
    pos = 0
    def take(n):
        nonlocal pos
        a = np.array([float(t) for t in toks[pos:pos + n]])
        pos += n
        return a
    xb, yb, zb = take(nx + 1), take(ny + 1), take(nz + 1)
    med = np.zeros((nz, ny, nx), dtype=np.int32)
    for k in range(nz):
        for j in range(ny):
            row = toks[pos]; pos += 1
            med[k, j] = [int(c) for c in row[:nx]]
    rho = take(nx * ny * nz).reshape(nz, ny, nx)
    return Phantom(media, xb, yb, zb, med, rho, os.path.basename(path))


def write_spectrum(path: str, title: str, e_upper, weights, enmin: float = 0.0, imode: int = 0) -> None:
    with open(path, "w") as f:
        f.write(title.rstrip("\n") + "\n")
        f.write(f"   {len(e_upper)} {enmin} {imode}\n")
        for e, w in zip(e_upper, weights):
            f.write(f"   {e:.6f} {w:.8e}\n")


def write_inp(stem: str, items: dict) -> str:
    """``key = value`` lines; the reader appends ``.inp`` to the stem."""
    path = stem + ".inp"
    with open(path, "w") as f:
        for k, v in items.items():
            f.write(f"{k} = {v}\n")
    return path


def default_inp(phantom_file: str, out_folder: str, *, energy=None, spectrum=None,
                collimator=(-5.0, 5.0, -5.0, 5.0), ssd=1.0e10, ncase=1000, nbatch=10,
                seeds=(1802, 9373), ecut=0.7, pcut=0.01, nsplit=1, charge=0,
                data_folder=None) -> dict:
    data_folder = data_folder or DATA
    d = {}
    if spectrum is not None:
        d["spectrum file"] = spectrum
    d["mono energy"] = energy if energy is not None else 2.0
    d["charge"] = charge
    d["collimator bounds"] = " ".join(f"{c}" for c in collimator)
    d["ssd"] = f"{ssd:.6E}"
    d["ncase"] = ncase
    d["nbatch"] = nbatch
    d["rng seeds"] = f"{seeds[0]} {seeds[1]}"
    d["phantom file"] = phantom_file
    d["global ecut"] = ecut
    d["global pcut"] = pcut
    d["pegs file"] = os.path.join(DATA, "pegs4", "521icru.pegs4dat")
    d["pgs4form file"] = os.path.join(DATA, "pegs4", "pgs4form.dat")
    d["nsplit"] = nsplit
    d["data folder"] = data_folder.rstrip("/") + "/"
    d["output folder"] = out_folder.rstrip("/") + "/"
    return d


def _bounds(n, vox, lo):
    return lo + vox * np.arange(n + 1, dtype=np.float64)


def water_cube(n: int = 60, vox: float = 0.5) -> Phantom:
    """All H2O521ICRU at 1.0 g/cm^3, x,y centred on 0, z from 0."""
    half = n * vox / 2.0
    return Phantom([WATER], _bounds(n, vox, -half), _bounds(n, vox, -half), _bounds(n, vox, 0.0),
                   np.ones((n, n, n), np.int32), np.ones((n, n, n)), f"water{n}")


def vacuum_gap_phantom(n: int = 12, vox: float = 0.5, gap=(2.0, 3.5)) -> Phantom:
    """Water cube with a vacuum layer (medium 0) across it for gap[0] < z <
    gap[1] (voxel centres decide): particles cross it without interacting."""
    ph = water_cube(n, vox)
    zc = 0.5 * (ph.zb[1:] + ph.zb[:-1])
    sel = (zc > gap[0]) & (zc < gap[1])
    ph.med[sel, :, :] = 0
    ph.rho[sel, :, :] = 0.0
    ph.name = f"vacgap{n}"
    return ph


def slab_phantom(n: int = 128, vox: float = 0.3) -> Phantom:
    """Layers along z: water 0-9 cm, lung 0.26 9-21 cm, bone 1.85 21-27 cm,
    water to the end (voxel centres decide the layer)."""
    half = n * vox / 2.0
    zb = _bounds(n, vox, 0.0)
    zc = 0.5 * (zb[1:] + zb[:-1])
    lay_med = np.where(zc < 9.0, 1, np.where(zc < 21.0, 2, np.where(zc < 27.0, 3, 1))).astype(np.int32)
    lay_rho = np.where(lay_med == 1, 1.0, np.where(lay_med == 2, 0.26, 1.85))
    med = np.broadcast_to(lay_med[:, None, None], (n, n, n)).copy()
    rho = np.broadcast_to(lay_rho[:, None, None], (n, n, n)).copy()
    return Phantom([WATER, LUNG, BONE], _bounds(n, vox, -half), _bounds(n, vox, -half), zb,
                   med, rho, f"slab{n}")


def synthetic_ct(nx=160, ny=160, nz=100, vx=0.25, vy=0.25, vz=0.3, seed=20261205,
                 target_radius=2.0) -> Phantom:
    """Seeded synthetic CT: air outside an elliptic water body (rho 0.95-1.05
    smooth noise), two lung ellipsoids (0.20-0.35), a bone cylinder (1.6-1.9)
    and a central spherical target of water.  Axes centred on 0."""
    rng = np.random.default_rng(seed)
    xb = _bounds(nx, vx, -nx * vx / 2); yb = _bounds(ny, vy, -ny * vy / 2); zb = _bounds(nz, vz, -nz * vz / 2)
    xc = 0.5 * (xb[1:] + xb[:-1]); yc = 0.5 * (yb[1:] + yb[:-1]); zc = 0.5 * (zb[1:] + zb[:-1])
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    ax, ay = 0.42 * nx * vx, 0.34 * ny * vy
    body = (X / ax) ** 2 + (Y / ay) ** 2 <= 1.0

    def smooth_noise(lo, hi):
        coarse = rng.random((max(2, nz // 8), max(2, ny // 8), max(2, nx // 8)))
        from scipy.ndimage import zoom
        f = zoom(coarse, (nz / coarse.shape[0], ny / coarse.shape[1], nx / coarse.shape[2]), order=1)
        f = f[:nz, :ny, :nx]
        return lo + (hi - lo) * f

    med = np.full((nz, ny, nx), 4, np.int32)       # AIR
    rho = np.full((nz, ny, nx), 0.0012)
    med[body] = 1
    rho_w = smooth_noise(0.95, 1.05)
    rho[body] = rho_w[body]
    lz = 0.35 * nz * vz
    for sx in (-1, 1):
        lung = ((X - sx * 0.45 * ax) / (0.3 * ax)) ** 2 + (Y / (0.55 * ay)) ** 2 + (Z / lz) ** 2 <= 1.0
        lung &= body
        med[lung] = 2
        rho_l = smooth_noise(0.20, 0.35)
        rho[lung] = rho_l[lung]
    bone = ((X / (0.12 * ax)) ** 2 + ((Y + 0.7 * ay) / (0.15 * ay)) ** 2 <= 1.0) & body
    med[bone] = 3
    rho_b = smooth_noise(1.6, 1.9)
    rho[bone] = rho_b[bone]
    target = X ** 2 + Y ** 2 + Z ** 2 <= target_radius ** 2
    med[target] = 1
    rho[target] = 1.0
    return Phantom([WATER, LUNG, BONE, AIR], xb, yb, zb, med, rho, f"ct{nx}x{ny}x{nz}")


@dataclass
class Beamlets:
    """matRad-style bixel source (``omc_matrad.c:780-829``): per beam a source
    point, per bixel a corner and two side vectors spanning the bixel at the
    isocentre plane, and the 0-based beam index of each bixel."""
    ibeam: np.ndarray
    src: np.ndarray          # (nbeams, 3)
    corner: np.ndarray       # (nbix, 3)
    side1: np.ndarray        # (nbix, 3)
    side2: np.ndarray        # (nbix, 3)
    nhist: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))

    @property
    def n(self):
        return len(self.ibeam)


def beamlets(nbeams=5, sad=100.0, bixel=0.5, radius=5.6, iso=(0.0, 0.0, 0.0),
             hist_median=7e3, hist_sigma_log10=0.35, hist_range=(1e3, 5e4), seed=20261205,
             plane_axis="z") -> Beamlets:
    """Coplanar beams at equispaced gantry angles (rotation about z), bixels on
    a ``bixel`` grid covering a disc of ``radius`` cm at the isocentre plane.
    Per-bixel history counts: log10 N ~ Normal(log10 median, sigma), clipped."""
    rng = np.random.default_rng(seed)
    iso = np.asarray(iso, float)
    ib, src, cor, s1, s2 = [], [], [], [], []
    grid = np.arange(-radius, radius, bixel) + 0.5 * bixel
    for b in range(nbeams):
        g = 2.0 * math.pi * b / nbeams
        d = np.array([math.sin(g), -math.cos(g), 0.0])    # beam direction (source -> iso)
        src.append(iso - sad * d)
        e1 = np.array([math.cos(g), math.sin(g), 0.0])    # in-plane, perpendicular to d
        e2 = np.array([0.0, 0.0, 1.0])
        for a in grid:
            for c in grid:
                if a * a + c * c > radius * radius:
                    continue
                centre = iso + a * e1 + c * e2
                cor.append(centre - 0.5 * bixel * (e1 + e2))
                s1.append(bixel * e1)
                s2.append(bixel * e2)
                ib.append(b)
    n = len(ib)
    lg = rng.normal(math.log10(hist_median), hist_sigma_log10, n)
    nh = np.clip(np.rint(10.0 ** lg), hist_range[0], hist_range[1]).astype(np.int64)
    return Beamlets(np.array(ib, np.int32), np.array(src), np.array(cor), np.array(s1), np.array(s2), nh)


def large_plan(total_histories=1e9, seed=20261207):
    """The large-plan workload: a 256x256x200 synthetic CT at 0.2 cm with the
    tissue model of ``synthetic_ct`` (its target a 5 cm sphere), 9 coplanar
    beams and 10044 bixels of 0.5 cm on a 5 mm grid (a disc of 9.4 cm at the
    isocentre), lognormal histories per bixel rescaled so that the plan
    carries ``total_histories``.  Returns (phantom, beamlets)."""
    ph = synthetic_ct(nx=256, ny=256, nz=200, vx=0.2, vy=0.2, vz=0.2, seed=seed, target_radius=5.0)
    ph.name = "ct256x256x200"
    bix = beamlets(nbeams=9, radius=9.4, seed=seed)
    w = bix.nhist / bix.nhist.sum()
    bix.nhist = np.maximum(np.rint(w * total_histories), 10).astype(np.int64)
    return ph, bix


def synthetic_6mv(seed=20261204, nbins=24, emin=0.25, emax=6.0):
    """24 bins with upper edges 0.25..6.0 MeV; weights E*exp(-E/1.5) times
    seeded +-5% noise, smoothed with a 3-point kernel; counts per bin."""
    rng = np.random.default_rng(seed)
    e = np.linspace(emin, emax, nbins)
    w = e * np.exp(-e / 1.5) * (1.0 + 0.05 * (2.0 * rng.random(nbins) - 1.0))
    w = np.convolve(np.pad(w, 1, mode="edge"), [0.25, 0.5, 0.25], mode="valid")
    return e, w / w.sum()


def small_test_phantom(nx=4, ny=3, nz=3, vox=0.5) -> Phantom:
    """Tiny four-medium phantom for sampler and single-history tests: media
    cycle with the voxel index, densities scaled away from the PEGS values."""
    media = [WATER, LUNG, BONE, AIR]
    n = nx * ny * nz
    idx = np.arange(n)
    med = (idx % 4 + 1).astype(np.int32).reshape(nz, ny, nx)
    base = np.array([1.0, 0.26, 1.85, 0.0012])
    scale = 1.0 + 0.1 * np.sin(idx).reshape(nz, ny, nx)
    rho = base[med - 1] * scale
    return Phantom(media, _bounds(nx, vox, -nx * vox / 2), _bounds(ny, vox, -ny * vox / 2),
                   _bounds(nz, vox, 0.0), med, rho, "small4")
