"""Readers for the golden-vector files written by oracle/ref_harness.c.

File = b"OMCG" followed by records:
  b"OMCA" name[32] int32 dtype(0 f64, 1 i32) int64 n  data      -> named array
  b"OMCC" int32 kind, n_in, n_rn, n_out  f64 in[] rn[] out[]     -> sampler case
"""
import gzip
import struct

import numpy as np

KINDS = {1: "compton", 2: "pair", 3: "photo", 4: "rayleigh", 5: "brems", 6: "moller", 7: "bhabha",
         8: "annih", 9: "rannih", 10: "msdist", 11: "mscat", 12: "sscat", 13: "drange", 14: "eloss",
         15: "photon", 16: "electron"}
KIND_ID = {v: k for k, v in KINDS.items()}


def read(path):
    """Returns (arrays: dict name -> ndarray, cases: list of (kind, in, rn, out))."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        buf = f.read()
    if buf[:4] != b"OMCG":
        raise ValueError(f"{path}: not a golden file")
    pos, arrays, cases = 4, {}, []
    while pos < len(buf):
        tag = buf[pos:pos + 4]
        pos += 4
        if tag == b"OMCA":
            name = buf[pos:pos + 32].split(b"\0", 1)[0].decode()
            dtype, n = struct.unpack_from("<iq", buf, pos + 32)
            pos += 44
            dt = np.float64 if dtype == 0 else np.int32
            size = n * (8 if dtype == 0 else 4)
            arrays[name] = np.frombuffer(buf, dtype=dt, count=n, offset=pos).copy()
            pos += size
        elif tag == b"OMCC":
            kind, nin, nrn, nout = struct.unpack_from("<4i", buf, pos)
            pos += 16
            vals = np.frombuffer(buf, dtype=np.float64, count=nin + nrn + nout, offset=pos)
            pos += 8 * (nin + nrn + nout)
            cases.append((kind, vals[:nin].copy(), vals[nin:nin + nrn].copy(), vals[nin + nrn:].copy()))
        else:
            raise ValueError(f"{path}: bad record tag {tag!r} at {pos - 4}")
    return arrays, cases
