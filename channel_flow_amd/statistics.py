"""The flow diagnostics of the reference's statistics (PF:2816-3190),
computed on the host from the plane statistics that the GPU reduces
(`Channel.plane_statistics`, C ABI chf_statistics).

Input rows (per wall-normal point y_k = cos(pi k/N), k = 0..N):
ua va wa uu vv ww uv uw vw max|u| max|v| max|w|.
"""
from __future__ import annotations

import math

import numpy as np


def _cheb_derivative_at_points(f):
    """d/dy of the degree-N interpolant of f(y_k), evaluated at y_k: the
    reference's chebyshev(-1) -> dz -> chebyshev(1) sequence (PF:3083-3085,
    analysis PF:990-994, dz PF:570-599, synthesis PF:951-957)."""
    N = len(f) - 1
    k = np.arange(N + 1)
    T = np.cos(np.pi * np.outer(k, k) / N)          # T[j, l] = T_l(y_j)
    w = np.ones(N + 1)
    w[0] = w[N] = 0.5
    c = np.ones(N + 1)
    c[0] = c[N] = 2.0
    a = (2.0 / (N * c)) * (T.T @ (w * f))           # a_l = 2/(N c_l) sum'' f_j T_l(y_j)
    b = np.zeros(N + 1)
    for j in range(N - 1, -1, -1):                  # b_j = b_{j+2} + 2 (j+1) a_{j+1}
        b[j] = (b[j + 2] if j + 2 <= N else 0.0) + 2.0 * (j + 1) * a[j + 1]
    b[0] *= 0.5
    return T @ b


def diagnostics(ps, nnx, nny, nnz, aphi, beta, dt, re, re_p) -> dict:
    ps = np.asarray(ps, dtype=np.float64)
    ua, va, wa, uu, vv, ww, uv, uw, vw, umax, vmax, wmax = ps
    N = nnz
    z = np.cos(np.pi * np.arange(N + 1) / N)
    pi2 = 2.0 * math.pi
    dx = aphi * pi2 / nnx                            # PF:2817-2820
    ddz = beta * pi2 / nny
    ddy = math.pi / N
    dy = np.zeros(N + 1)
    for k in range(1, N // 2 + 1):                   # PF:2821-2824
        dy[k] = -(math.cos(k * ddy) - math.cos((k - 1) * ddy))
        dy[N - k] = dy[k]
    cfl_x = float(np.max(umax)) * dt / dx            # PF:2825-2838
    cfl_z = float(np.max(wmax)) * dt / ddz           # PF:2839-2852
    cfl_y = float(max(vmax[k] * dt / dy[k] for k in range(1, N)))  # PF:2853-2868
    dudy = _cheb_derivative_at_points(ua)
    tau = -uv + dudy / re                            # total shear stress, PF:3086-3088
    ut_bot, ut_top = math.sqrt(abs(tau[N])), math.sqrt(abs(tau[0]))
    ubulk = float(sum(ua[k] * abs(z[k + 1] - z[k]) for k in range(N // 2, N)))  # PF:3163-3166
    plane_points = nnx * nny
    energy = 0.5 * plane_points * float(np.sum(uu + ua ** 2 + vv + va ** 2 + ww + wa ** 2))  # PF:2766-2774
    return dict(
        y=z, mean_u=ua, mean_v=va, mean_w=wa,
        uu=uu, vv=vv, ww=ww, uv=uv, uw=uw, vw=vw,
        total_shear_stress=tau, viscous_stress=dudy / re,
        cfl_x=cfl_x, cfl_y=cfl_y, cfl_z=cfl_z,
        wall_shear_velocity=(ut_bot, ut_top),            # sqrt|tau| at y = -1 and y = +1 (PF:3168)
        mean_pressure_gradient=-re_p,
        re_tau=(ut_bot * re, ut_top * re),               # PF:3171
        re_centre=float(ua[N // 2] * re),                # PF:3172
        re_bulk=2.0 * ubulk * re,                        # PF:3173
        energy=energy,
    )
