#!/usr/bin/env python3
"""Write a synthetic ``spinms.data`` in the binary layout the engine reads.

The EGSnrc spin-correction file (Mott/PWA rejection surfaces for Z = 1..100) is
not distributed with the reference (it is listed as a missing large blob), yet
``initSpinData`` (reference ``src/ompmc.c:2394-2965``) exits without it.  This
script writes a stand-in with the exact record layout that reader expects
(``src/ompmc.c:2424-2625``):

* record size 1104 bytes (276 float32 = 552 int16), little endian;
* record 0: 32-byte version string, 4-byte endianness tag, float32
  ``espin_min, espin_max`` (keV) and ``b2spin_min, b2spin_max``, zero padding;
* data record ``r = (Z-1)*64 + 32*iq + i + 1`` (0-based; iq 0 = e-, 1 = e+,
  i = 0..31 energy index): float32 ``dum1, dum2, dum3, aux_o``, float32
  ``fmax[16]``, then uint16 ``q[16][32]`` starting at byte 80.

Two variants:

``neutral``  (default, used by every dose configuration): q = 65535 and
    fmax = 1, so the spin rejection function is identically 1; dum1 = dum3 = 1,
    aux_o = 1.13 and dum2 = gamma(eta) of the element so that the q1/q2
    moment corrections reduce to ~1.  Physically this is "no spin correction".
``textured`` (used only by sampler-parity tests): q falls smoothly with the
    angle index k, differently per (Z, iq, energy, q1) row, so the rejection
    loops in ``mscat``/``sscat`` actually reject; dum/aux values vary with Z.

Usage: ``python data/gen_spinms.py [--variant neutral|textured] OUT``
"""
import argparse
import math
import struct

import numpy as np

RM = 0.5109989461
NREC = 1 + 100 * 64
RECLEN = 1104
ESPIN_MIN = 1.0       # keV
ESPIN_MAX = 100.0     # keV
B2SPIN_MAX = 0.9999


def _beta2(ekin_kev):
    tau = ekin_kev / (1000.0 * RM)
    return tau * (tau + 2.0) / ((tau + 1.0) * (tau + 1.0))


def _energy_grid(b2min):
    """The 32-point kinetic-energy grid (keV) the reader rebuilds from record 0
    (16 log-spaced points, then 16 points uniform in beta^2)."""
    f32 = np.float32
    emin, emax = f32(ESPIN_MIN), f32(ESPIN_MAX)
    dloge = math.log(float(f32(emax / emin))) / 15.0
    e = [float(emin)]
    el = math.log(float(emin))
    for _ in range(15):
        el += dloge
        e.append(math.exp(el))
    e.append(float(emax))
    db = float(f32((f32(B2SPIN_MAX) - f32(b2min)) / f32(15)))
    b = float(f32(b2min))
    for _ in range(15):
        b += db
        e.append(RM * 1000.0 * (1.0 / math.sqrt(1.0 - b) - 1.0) if b < 0.999 else 50585.1)
    return e


def _gamma_of_eta(z, aux_o, ekin_kev):
    tau = ekin_kev / (1000.0 * RM)
    eta = z ** (2.0 / 3.0) / ((137.03604 * 0.88534138) ** 2) * aux_o / 4.0 / tau / (tau + 2.0)
    lg = math.log(1.0 + 1.0 / eta)
    return 3.0 * (1.0 + eta) * (lg * (1.0 + 2.0 * eta) - 2.0) / (lg * (1.0 + eta) - 1.0)


def build(variant="neutral"):
    b2min = _beta2(ESPIN_MAX)
    egrid = _energy_grid(b2min)
    buf = bytearray(NREC * RECLEN)
    head = b"synthetic spinms (ompmc_amd)    "[:32].ljust(32, b" ")
    buf[0:32] = head
    buf[32:36] = b"1234"
    struct.pack_into("<4f", buf, 36, ESPIN_MIN, ESPIN_MAX, b2min, B2SPIN_MAX)
    k = np.arange(32, dtype=np.float64)
    for z in range(1, 101):
        for iq in range(2):
            for i in range(32):
                r = (z - 1) * 64 + 32 * iq + i + 1
                off = r * RECLEN
                if variant == "neutral":
                    aux_o, dum1, dum3 = 1.13, 1.0, 1.0
                    fmax = np.ones(16)
                    q = np.full((16, 32), 65535, dtype=np.uint16)
                else:
                    aux_o = 1.13 + 0.02 * math.sin(0.7 * z)
                    dum1 = 1.0 + 0.03 * math.cos(0.3 * z + iq)
                    dum3 = 1.0 - 0.02 * math.sin(0.5 * z + i / 7.0)
                    fmax = 1.0 + 0.05 * np.arange(16)
                    slope = 0.35 + 0.25 * (i / 31.0) + 0.1 * iq + 0.002 * z
                    j = np.arange(16, dtype=np.float64)[:, None]
                    shape = 1.0 - slope * (k[None, :] / 31.0) ** 1.5 * (0.4 + 0.6 * j / 15.0)
                    q = np.clip(np.rint(65535.0 * shape), 1, 65535).astype(np.uint16)
                dum2 = _gamma_of_eta(z, aux_o, egrid[i])
                struct.pack_into("<4f", buf, off, dum1, dum2, dum3, aux_o)
                struct.pack_into("<16f", buf, off + 16, *[float(x) for x in fmax])
                buf[off + 80: off + 80 + 1024] = q.astype("<u2").tobytes()
    return bytes(buf)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("out")
    ap.add_argument("--variant", choices=("neutral", "textured"), default="neutral")
    a = ap.parse_args()
    data = build(a.variant)
    with open(a.out, "wb") as f:
        f.write(data)


if __name__ == "__main__":
    main()
