"""Seeded initial conditions in the reference's spectral layout.

The reference builds its cold start from physical-space noise drawn with the
compiler's RNG (input, PF:2086-2339); that stream is not reproducible outside
the reference binary.  This module instead draws divergence-free
perturbations directly in spectral space:

* wall-normal velocity  v-hat(y) = A (1 - y^2)^2   (v = v' = 0 at the walls)
* wall-normal vorticity w-hat(y) = B (1 - y^2)     (omega_y = 0 at the walls)

for every mode 0 < |k| <= kmax (|k| = sqrt(i^2 + j^2) in mode-index units,
j signed), with A, B complex normal from PCG64(seed) scaled by |k|^slope.
The kx = 0 plane is made Hermitian, the spanwise Nyquist column and the
(0,0) mode of v and omega_y are zero.  The perturbation is rescaled to a
prescribed volume rms of (u', v', w') (u', w' from the kinematic relations
of PF:2367-2375), and the mean flow is the parabola U = U_c (1 - y^2).

All arrays use the reference memory layout: complex128[Ny][Mz][Mx] for
fields and complex128[Ny] for the mean-mode arrays.
"""
from __future__ import annotations

import numpy as np


def cheb_deriv(a: np.ndarray) -> np.ndarray:
    """Chebyshev coefficient derivative along axis 0 (backward recurrence)."""
    n = a.shape[0] - 1
    b = np.zeros_like(a)
    if n >= 1:
        b[n - 1] = 2.0 * n * a[n]
    for k in range(n - 2, -1, -1):
        b[k] = b[k + 2] + 2.0 * (k + 1) * a[k + 1]
    b[0] = b[0] * 0.5
    return b


def wavenumbers(nnx, nny, aphi, beta):
    mx, mz = nnx // 2, nny
    alpha = np.arange(mx, dtype=np.float64) / aphi
    j = np.arange(mz)
    bet = np.where(j <= mz // 2, j, j - mz).astype(np.float64) / beta
    return alpha, bet


def perturbation_rms(v, oy, nnx, nny, aphi, beta):
    """Volume rms of (u', v', w') excluding the (0,0) mode.

    Parseval in x and z (streamwise modes i > 0 count twice, their conjugates
    are not stored) and Gauss-Legendre quadrature of the Chebyshev series in y.
    """
    ny, mz, mx = v.shape
    alpha, bet = wavenumbers(nnx, nny, aphi, beta)
    k2 = alpha[None, :] ** 2 + bet[:, None] ** 2
    k2s = np.where(k2 == 0.0, 1.0, k2)
    dv = cheb_deriv(v)
    u = (1j * alpha[None, None, :] * dv - 1j * bet[None, :, None] * oy) / k2s
    w = (1j * bet[None, :, None] * dv + 1j * alpha[None, None, :] * oy) / k2s
    mask = np.ones((mz, mx))
    mask[0, 0] = 0.0
    mask[mz // 2, :] = 0.0
    wx = np.where(np.arange(mx) == 0, 1.0, 2.0)[None, :] * mask
    xg, wg = np.polynomial.legendre.leggauss(ny + 4)
    tk = np.cos(np.arange(ny)[None, :] * np.arccos(xg)[:, None])      # (q, ny)
    tot = 0.0
    for f in (u, v, w):
        vals = np.tensordot(tk, f.reshape(ny, -1), axes=(1, 0)).reshape(len(xg), mz, mx)
        tot += float(np.sum(wg[:, None, None] * np.abs(vals) ** 2 * wx[None]))
    return np.sqrt(tot / 2.0)


def seeded_ic(nnx, nny, nnz, aphi, beta, u_c, rms, kmax, slope, seed):
    """Return (v, omega_y, au0, w0) for the seeded recipe described above."""
    mx, mz, ny = nnx // 2, nny, nnz + 1
    if nnz < 4:
        raise ValueError("nnz must be >= 4 for the (1-y^2)^2 shape function")
    rng = np.random.Generator(np.random.PCG64(seed))
    amp_a = rng.standard_normal((mz, mx)) + 1j * rng.standard_normal((mz, mx))
    amp_b = rng.standard_normal((mz, mx)) + 1j * rng.standard_normal((mz, mx))
    jj = np.arange(mz)
    kz = np.where(jj <= mz // 2, jj, jj - mz)
    kk = np.sqrt(np.arange(mx)[None, :] ** 2 + kz[:, None] ** 2.0)
    active = (kk > 0) & (kk <= kmax)
    active[mz // 2, :] = False
    scale = np.where(active, np.where(kk > 0, kk, 1.0) ** slope, 0.0)
    amp_a *= scale
    amp_b *= scale
    # Hermitian kx = 0 plane: the -kz entry is the conjugate of +kz
    for j in range(1, mz // 2):
        amp_a[mz - j, 0] = np.conj(amp_a[j, 0])
        amp_b[mz - j, 0] = np.conj(amp_b[j, 0])
    amp_a[0, 0] = 0.0
    amp_b[0, 0] = 0.0
    sv = np.zeros(ny)
    sv[0], sv[2], sv[4] = 3.0 / 8.0, -0.5, 1.0 / 8.0        # (1-y^2)^2
    so = np.zeros(ny)
    so[0], so[2] = 0.5, -0.5                                # (1-y^2)
    v = sv[:, None, None] * amp_a[None]
    oy = so[:, None, None] * amp_b[None]
    cur = perturbation_rms(v, oy, nnx, nny, aphi, beta)
    if cur > 0.0 and rms > 0.0:
        v *= rms / cur
        oy *= rms / cur
    au0 = np.zeros(ny, dtype=np.complex128)
    au0[0] = 0.5 * u_c
    au0[2] = -0.5 * u_c
    w0 = np.zeros(ny, dtype=np.complex128)
    return (np.ascontiguousarray(v, dtype=np.complex128),
            np.ascontiguousarray(oy, dtype=np.complex128), au0, w0)


# Workload table: the five configurations of the measurement plan.
# All are in centreline units (U_c = 1, re = centreline Reynolds number), the
# convention the 32x33x32 case states explicitly (Re_centerline = 3000).  The
# Re_tau classes map to the centreline Reynolds number of the turbulent flow at
# that Re_tau (U_c+ = 18.3, 20.1, 21.2 for Re_tau = 180, 395, 590).  re_p = 2/re
# makes the laminar profile U_c (1 - y^2) an exact steady state of the mean
# equation.  The noise is 10% of centreline (rms 0.1).  With these units the
# explicit AB2 advection stays well inside its stability range:
# kx_max U_c dt = 0.035 at 256x193x256 (wall units, U_c = 20, would give 0.64,
# where AB2 amplifies the highest streamwise modes by ~30% per step).
# (nnx, nny, nnz, aphi, beta, re, re_p, dt, U_c, rms, kmax, slope, seed)
CONFIGS = {
    "re3000_32x33x32": dict(nnx=32, nny=32, nnz=32, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0,
                          dt=1.0e-3, u_c=1.0, rms=1.0e-3, kmax=4.0, slope=0.0, seed=1001, steps=100),
    "retau180_128x129x128": dict(nnx=128, nny=128, nnz=128, aphi=2.0, beta=2.0 / 3.0, re=3300.0,
                             re_p=2.0 / 3300.0, dt=5.0e-4, u_c=1.0, rms=0.1, kmax=16.0,
                             slope=-4.0 / 3.0, seed=1002, steps=1000),
    "retau395_256x193x256": dict(nnx=256, nny=256, nnz=192, aphi=1.0, beta=0.5, re=7940.0,
                             re_p=2.0 / 7940.0, dt=2.5e-4, u_c=1.0, rms=0.1, kmax=16.0,
                             slope=-4.0 / 3.0, seed=1003, steps=100),
    "retau590_512x257x512": dict(nnx=512, nny=512, nnz=256, aphi=1.0, beta=0.5, re=12500.0,
                             re_p=2.0 / 12500.0, dt=2.5e-4, u_c=1.0, rms=0.1, kmax=16.0,
                             slope=-4.0 / 3.0, seed=1005, steps=50),
}


# Batched ensemble workload: 24 independent channels, resolutions drawn
# uniformly from {64x65x64, 96x97x96, 128x97x128} (nnx x (N+1) x nny), centreline
# Reynolds number uniform in [2000, 6000] with re_p = 2/re (the parabola is the
# steady state), dt = 1e-3, mixed precision, 500 steps; member m uses IC seed
# 2000 + m.
ENSEMBLE_SIZES = ((64, 64, 64), (96, 96, 96), (128, 128, 96))  # (nnx, nny, nnz)


def ensemble_members(n=24, seed=4000, precision="mixed"):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for m in range(n):
        nnx, nny, nnz = ENSEMBLE_SIZES[int(rng.integers(0, len(ENSEMBLE_SIZES)))]
        re = float(rng.uniform(2000.0, 6000.0))
        out.append(dict(nnx=nnx, nny=nny, nnz=nnz, aphi=1.0, beta=0.5, re=re, re_p=2.0 / re, dt=1.0e-3,
                        u_c=1.0, rms=0.1, kmax=16.0, slope=-4.0 / 3.0, seed=2000 + m,
                        precision=precision))
    return out
