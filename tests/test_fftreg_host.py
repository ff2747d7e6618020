"""The register DFTs of csrc/fftreg.h (the small transforms inside every FFT
kernel: span, streamwise and wall-normal), compiled for the host and checked
against numpy's FFT in both directions and both precisions, for every radix
the kernels use.  The composite radices (6 ... 32) exercise the twiddle paths,
including the twiddles that are whole quarter turns (applied as swaps and
sign changes)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RADICES = [2, 3, 4, 6, 8, 12, 16, 24, 32]


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("no hipcc")
    out = tmp_path_factory.mktemp("dft") / "libdft.so"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-shared", "-fPIC", "--offload-arch=gfx950",
                    "-Wno-unused-command-line-argument", os.path.join(HERE, "native", "dft_harness.cpp"),
                    "-o", str(out)], check=True, capture_output=True)
    l = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    l.dft_c.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
    l.dft_c.restype = ctypes.c_int
    return l


def run(lib, x, sign, fp32):
    r = len(x)
    inp = np.ascontiguousarray(np.stack([x.real, x.imag], axis=1).ravel())
    out = np.zeros(2 * r)
    dp = ctypes.POINTER(ctypes.c_double)
    assert lib.dft_c(r, sign, int(fp32), inp.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    return out[0::2] + 1j * out[1::2]


@pytest.mark.parametrize("r", RADICES)
@pytest.mark.parametrize("sign", [-1, 1])
def test_dft_matches_numpy(lib, r, sign):
    rng = np.random.default_rng(100 * r + sign)
    for _ in range(20):
        x = rng.standard_normal(r) + 1j * rng.standard_normal(r)
        ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * r
        scale = np.max(np.abs(ref))
        got = run(lib, x, sign, False)
        assert np.max(np.abs(got - ref)) <= 1e-14 * scale * np.log2(r + 1), (r, sign)
        # the input rounded to float, transformed in float: float rounding only
        xf = x.real.astype(np.float32).astype(np.float64) + 1j * x.imag.astype(np.float32).astype(np.float64)
        reff = np.fft.fft(xf) if sign < 0 else np.fft.ifft(xf) * r
        gotf = run(lib, x, sign, True)
        assert np.max(np.abs(gotf - reff)) <= 2e-7 * scale * np.log2(r + 1), (r, sign)


def test_unit_impulses_give_the_exponentials(lib):
    """A unit impulse at n gives X[k] = e^{s 2 pi i n k / r}: every input
    position, hence every twiddle path of the composite transforms, within
    rounding of the exact exponentials."""
    for r in RADICES:
        for sign in (-1, 1):
            for n in range(r):
                x = np.zeros(r, dtype=complex)
                x[n] = 1.0
                got = run(lib, x, sign, False)
                # the phase reduced mod r first, so the reference carries no argument rounding
                exact = np.exp(sign * 2j * np.pi * ((n * np.arange(r)) % r) / r)
                assert np.max(np.abs(got - exact)) < 1e-15 * r, (r, sign, n)
