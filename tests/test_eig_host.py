"""Host linear algebra of the operator set-up (csrc/eig.h), compiled with g++
and checked against numpy: eigen-decomposition of real matrices with real
spectra, including the Dirichlet second-derivative pencil of the omega_y
half systems that the diagonalised omega_y solve uses, and the inverse."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import cf_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    out = tmp_path_factory.mktemp("eig") / "libeig.so"
    subprocess.run([cxx, "-O2", "-std=c++17", "-shared", "-fPIC", os.path.join(HERE, "native", "eig_harness.cpp"),
                    "-o", str(out)], check=True, capture_output=True)
    l = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    l.eig_real_c.argtypes = [ctypes.c_int, dp, dp, dp]
    l.inverse_c.argtypes = [ctypes.c_int, dp, dp]
    return l


def eig(lib, a):
    n = a.shape[0]
    a = np.ascontiguousarray(a, dtype=np.float64)
    lam = np.zeros(n)
    v = np.zeros((n, n))
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.eig_real_c(n, a.ctypes.data_as(dp), lam.ctypes.data_as(dp), v.ctypes.data_as(dp))
    assert rc == 0
    return lam, v


def omega_pencil(N, par):
    """G = M0^-1 M2 of one parity half of the omega_y collocation system:
    columns l = 2j + par, rows = interior points p = 1..n-1, null space of
    the all-ones wall row through x = Q z, Q = [-1^T; I]."""
    t0, t1, t2, t4 = O.scheb(N)
    n = N // 2 + 1 if par == 0 else (N + 1) // 2
    cols = [2 * j + par for j in range(n)]
    T0 = t0[1:n][:, cols]
    T2 = t2[1:n][:, cols]
    Q = np.vstack([-np.ones((1, n - 1)), np.eye(n - 1)])
    M0, M2 = T0 @ Q, T2 @ Q
    return np.linalg.solve(M0, M2)


def test_random_real_spectra(lib):
    rng = np.random.default_rng(5)
    for n in (2, 5, 17, 64, 97):
        lam0 = np.sort(rng.uniform(-10, 10, n))
        S = rng.standard_normal((n, n)) + 3 * np.eye(n)
        A = S @ np.diag(lam0) @ np.linalg.inv(S)
        lam, V = eig(lib, A)
        np.testing.assert_allclose(np.sort(lam), lam0, rtol=0, atol=1e-9 * np.abs(lam0).max())
        res = np.abs(A @ V - V * lam).max() / (np.abs(A).max() * n)
        assert res < 1e-13, (n, res)


@pytest.mark.parametrize("N,par", [(16, 0), (16, 1), (192, 0), (192, 1), (21, 1)])
def test_omega_pencil_decomposition(lib, N, par):
    G = omega_pencil(N, par)
    lam, V = eig(lib, G)
    ref = np.sort(np.linalg.eigvals(G).real)
    assert np.all(np.sort(lam) < 0)
    np.testing.assert_allclose(np.sort(lam), ref, rtol=1e-9, atol=1e-9)
    res = np.abs(G @ V - V * lam).max() / (np.abs(G).max())
    assert res < 1e-12, res
    # the reconstruction the solve relies on: (sigma - G)^-1 = V diag(1/(sigma - lam)) V^-1
    sigma = 1.0 / 1.6e-8
    Vi = np.linalg.inv(V)
    recon = (V / (sigma - lam)) @ Vi
    direct = np.linalg.inv(sigma * np.eye(len(lam)) - G)
    assert np.abs(recon - direct).max() / np.abs(direct).max() < 1e-12


def test_inverse(lib):
    rng = np.random.default_rng(6)
    a = rng.standard_normal((40, 40)) + 5 * np.eye(40)
    out = np.zeros_like(a)
    dp = ctypes.POINTER(ctypes.c_double)
    assert lib.inverse_c(40, np.ascontiguousarray(a).ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    np.testing.assert_allclose(out @ a, np.eye(40), atol=1e-12)
