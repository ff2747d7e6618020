"""Restart records in the reference's on-disk layout.

`output` (PF:2460-2504) writes, and `initial(isign=1)` (PF:1861-1924) reads,
a Fortran sequential unformatted file of two records:

    record 1: clamb1n, clamb1nm1, clamb3n, clamb3nm1, au0, w0     each real*8 (2, 0:nz)
    record 2: vnp1, omegaynp1, rhsvn, rhsomegayn, unp1, wnp1     each real*8 (0:nx, 0:ny, 0:nz)

Each record is framed by a 4-byte little-endian length marker before and
after it (the record length in bytes), as the Fortran runtimes of gfortran
and flang write it.  In memory the arrays are identical to the C-ABI state:
(2, 0:nz) is complex128[Ny] and (0:nx, 0:ny, 0:nz) with Re/Im interleaved
along dim 1 is complex128[Ny][Mz][Mx].  The file written by the reference at
rank id is named data00000002_instant.<id:03d>.dat (PF:2466).

Records longer than 2**31 - 1 bytes (record 2 above about 512 x 257 x 512)
would need the runtime's subrecord continuation markers; they are refused.
"""
from __future__ import annotations

import struct

import numpy as np

RECORD1 = ("clamb1n", "clamb1nm1", "clamb3n", "clamb3nm1", "au0", "w0")
RECORD2 = ("vnp1", "omegaynp1", "rhsvn", "rhsomegayn", "unp1", "wnp1")
MAX_RECORD = 2 ** 31 - 1


def restart_name(rank: int = 0) -> str:
    """File name of the restart record written by rank `rank` (PF:2466)."""
    return "data%08d_instant.%03d.dat" % (2, rank)


def _write_record(f, arrays):
    payload = b"".join(np.ascontiguousarray(a, dtype=np.complex128).tobytes() for a in arrays)
    n = len(payload)
    if n > MAX_RECORD:
        raise ValueError(f"record of {n} bytes exceeds the single-marker limit of {MAX_RECORD}")
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def _read_record(f, nbytes: int) -> bytes:
    head = f.read(4)
    if len(head) != 4:
        raise ValueError("truncated restart file: missing record marker")
    (n,) = struct.unpack("<i", head)
    if n != nbytes:
        raise ValueError(f"record length {n} does not match the expected {nbytes} bytes "
                         f"(wrong grid size for this file?)")
    data = f.read(n)
    tail = f.read(4)
    if len(data) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
        raise ValueError("truncated or corrupt restart record")
    return data


def write_restart(path: str, state: dict) -> None:
    """Writes the 12-array pre-solve state as the reference's two records."""
    with open(path, "wb") as f:
        _write_record(f, [state[k] for k in RECORD1])
        _write_record(f, [state[k] for k in RECORD2])


def read_restart(path: str, nnx: int, nny: int, nnz: int) -> dict:
    """Reads a restart file for the grid (nnx, nny, nnz) of PC:5-7 (one rank)."""
    mx, mz, ny = nnx // 2, nny, nnz + 1
    n1, n2 = ny, ny * mz * mx
    out = {}
    with open(path, "rb") as f:
        r1 = np.frombuffer(_read_record(f, 6 * n1 * 16), dtype=np.complex128)
        r2 = np.frombuffer(_read_record(f, 6 * n2 * 16), dtype=np.complex128)
        if f.read(1):
            raise ValueError("trailing bytes after the second restart record")
    for i, k in enumerate(RECORD1):
        out[k] = r1[i * n1:(i + 1) * n1].copy()
    for i, k in enumerate(RECORD2):
        out[k] = r2[i * n2:(i + 1) * n2].reshape(ny, mz, mx).copy()
    return out
