"""Reader for the binary state dumps written by oracle/ref_driver.F90.

TEST INFRASTRUCTURE ONLY.  Layout: 8-byte magic ``FWDUMP01``, then records
``name[32] kind ndim d1 d2 d3`` (int32) followed by the data in Fortran
(column-major) order; ``kind`` is the byte width of a real (4 or 8), 104 for
int32 or 200 for an 80-character string; trailer ``FWDUMPEN``.
Arrays are returned as numpy arrays indexed ``[j, i]`` (row = y index), i.e.
the transpose view of the Fortran ``(I, J)`` layout, with ghost cells.
"""
from __future__ import annotations

import struct

import numpy as np


def read_dump(path: str) -> dict:
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] != b"FWDUMP01":
        raise ValueError(f"{path}: not a state dump")
    pos = 8
    n = len(buf)
    while pos < n:
        if buf[pos:pos + 8] == b"FWDUMPEN":
            break
        name = buf[pos:pos + 32].decode().strip()
        kind, ndim, d1, d2, d3 = struct.unpack_from("<5i", buf, pos + 32)
        pos += 52
        dims = (d1, d2, d3)[:max(ndim, 1)]
        count = int(np.prod(dims))
        if kind == 200:
            out[name] = buf[pos:pos + d1].decode().strip()
            pos += d1
            continue
        dt = {4: np.float32, 8: np.float64, 104: np.int32}[kind]
        nbytes = count * np.dtype(dt).itemsize
        a = np.frombuffer(buf, dtype=dt, count=count, offset=pos).copy()
        pos += nbytes
        if ndim == 0:
            out[name] = a[0].item()
        elif ndim == 1:
            out[name] = a
        else:
            # Fortran order (I fastest) -> numpy [.., J, I]
            out[name] = a.reshape(tuple(reversed(dims)))
    return out
