"""Restart records in the reference's layout (PF:1861-1924, 2460-2504).

The Python reader/writer is checked against the Fortran runtime itself: a
small Fortran program reads a Python-written file with the reference's read
statements and writes it back with its write statements.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from channel_flow_amd import restart
from channel_flow_amd.build import flang

HERE = os.path.dirname(os.path.abspath(__file__))


def random_state(nnx, nny, nnz, seed=3):
    rng = np.random.default_rng(seed)
    mx, mz, ny = nnx // 2, nny, nnz + 1
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    st = {k: c(ny) for k in restart.RECORD1}
    st.update({k: c(ny, mz, mx) for k in restart.RECORD2})
    return st


def test_roundtrip_and_framing(tmp_path):
    st = random_state(8, 6, 5)
    p = tmp_path / restart.restart_name(0)
    restart.write_restart(str(p), st)
    raw = p.read_bytes()
    n1, n2 = 6 * 6 * 16, 6 * 6 * 6 * 4 * 16
    assert len(raw) == n1 + n2 + 16
    assert int.from_bytes(raw[:4], "little") == n1 == int.from_bytes(raw[4 + n1:8 + n1], "little")
    back = restart.read_restart(str(p), 8, 6, 5)
    for k in st:
        np.testing.assert_array_equal(back[k], st[k])
    with pytest.raises(ValueError):
        restart.read_restart(str(p), 8, 6, 6)


def test_layout_matches_fortran_unformatted_io(tmp_path):
    fc = flang()
    if fc is None:
        pytest.skip("no Fortran compiler")
    exe = tmp_path / "restart_echo"
    subprocess.run([fc, "-O1", "-J", str(tmp_path), os.path.join(HERE, "fortran", "restart_echo.f90"),
                    "-o", str(exe)], check=True, capture_output=True)
    nnx, nny, nnz = 8, 6, 5
    st = random_state(nnx, nny, nnz)
    fin, fout = tmp_path / "in.dat", tmp_path / "out.dat"
    restart.write_restart(str(fin), st)
    subprocess.run([str(exe), str(nnx), str(nny), str(nnz), str(fin), str(fout)], check=True,
                   capture_output=True, timeout=60)
    back = restart.read_restart(str(fout), nnx, nny, nnz)
    ref = {k: 2 * v for k, v in st.items()}
    # Fortran vnp1(2, 2, 3) = Re of streamwise mode 1, span mode 2, degree 3
    ref["vnp1"][3, 2, 1] = 12345.0 + 1j * ref["vnp1"][3, 2, 1].imag
    for k in st:
        np.testing.assert_array_equal(back[k], ref[k])
