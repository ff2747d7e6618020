"""Shared helpers for the parity tests (oracle on the CPU, HIP path through the C ABI)."""
import numpy as np

from channel_flow_amd import ic
from oracle import cf_oracle as O

STATE_KEYS = ("vnp1", "omegaynp1", "rhsvn", "rhsomegayn", "unp1", "wnp1",
              "clamb1n", "clamb1nm1", "clamb3n", "clamb3nm1", "au0", "w0")


def params_of(cfg):
    return O.Params(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"],
                    cfg["re"], cfg["re_p"])


def ic_of(cfg):
    return ic.seeded_ic(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["u_c"],
                        cfg["rms"], cfg["kmax"], cfg["slope"], cfg["seed"])


def rel_err(a, b):
    """max |a - b| / max |b| (absolute when b is identically zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    diff = float(np.max(np.abs(a - b))) if b.size else 0.0
    return diff / scale if scale > 0 else diff


def nyquist_masked(a, mz):
    """Copy with the spanwise Nyquist column zeroed (it is never solved or used)."""
    b = np.array(a, copy=True)
    if b.ndim == 3:
        b[:, mz // 2, :] = 0
    return b


MEAN_PAIRS = (("au0", "w0"), ("clamb1n", "clamb3n"), ("clamb1nm1", "clamb3nm1"))


def state_errors(gpu, ref):
    """Per-array relative errors of a pre-solve state.

    Field arrays are measured against their own max modulus.  The mean-mode
    arrays come in vector pairs, the mean velocity (U, W) and the mean Lamb
    vector (L_x, L_z)(0,0), and each member is measured against the max
    modulus of its pair: W and L_z(0,0) are second order in the perturbation
    and sit far below the rounding floor of the O(1) first-order products
    whose (0,0) modes cancel, so their size alone is not a meaningful scale.
    """
    errs = {}
    paired = {}
    for a, b in MEAN_PAIRS:
        scale = max(float(np.max(np.abs(ref[a]))), float(np.max(np.abs(ref[b]))))
        paired[a] = paired[b] = scale
    for k in STATE_KEYS:
        if k in paired and paired[k] > 0:
            errs[k] = float(np.max(np.abs(np.asarray(gpu[k]) - np.asarray(ref[k])))) / paired[k]
        else:
            errs[k] = rel_err(gpu[k], ref[k])
    return errs


SMALL = dict(nnx=16, nny=16, nnz=16, aphi=1.0, beta=0.5, re=500.0, re_p=2.0 / 500.0, dt=1.0e-3,
             u_c=1.0, rms=5.0e-2, kmax=4.0, slope=0.0, seed=7)


def solved_state_invariants(p, st, fl, before=None):
    """Relative residuals of the properties every solved state has: continuity
    iαu + Dv + iβw = 0 and ω_y = iβu − iαw for k² ≠ 0 (PF:2357-2403), the
    Hermitian kx = 0 plane, the mean-mode vorticity ω_x(0,0) = DW,
    ω_z(0,0) = −DU with u(0,0) = U, w(0,0) = W (PF:423-472), the unsolved
    spanwise Nyquist column and the frozen v(0,0), ω_y(0,0) (PF:247, 2365).
    `st` holds v, ω_y, U, W; `fl` the post-solve u, w, ω_x, ω_z.  Without
    `before` the Nyquist column and the (0,0) modes are measured as they are
    (zero for a seeded start); with `before` (the state the run started from)
    as their change, which a frozen entry never has."""
    alpha, beta = O.wavenumbers(p)
    a3 = alpha[None, None, :]
    b3 = beta[None, :, None]
    v, oy, u, w = st["vnp1"], st["omegaynp1"], fl["unp1"], fl["wnp1"]
    mask = np.ones((1, p.Mz, p.Mx), dtype=bool)
    mask[0, p.Mz // 2, :] = False
    mask[0, 0, 0] = False
    dv = O.dz(v)
    vs = max(float(np.max(np.abs(dv))), 1e-300)
    us = max(float(np.max(np.abs(a3 * u))), float(np.max(np.abs(b3 * w))))
    div = np.where(mask, 1j * a3 * u + dv + 1j * b3 * w, 0.0)
    curl = np.where(mask, 1j * b3 * u - 1j * a3 * w - oy, 0.0)
    mz = p.Mz
    jr = (mz - np.arange(1, mz // 2)) % mz
    herm = max(float(np.max(np.abs(f[:, jr, 0] - np.conj(f[:, 1:mz // 2, 0]))))
               / float(np.max(np.abs(f))) for f in (v, oy))
    dU, dW = O.dz(st["au0"]), O.dz(st["w0"])
    mscale = max(float(np.max(np.abs(dU))), float(np.max(np.abs(dW))), 1e-300)
    mean_curl = max(float(np.max(np.abs(fl["omegaxnp1"][:, 0, 0] - dW))),
                    float(np.max(np.abs(fl["omegaznp1"][:, 0, 0] + dU)))) / mscale
    mean_vel = max(float(np.max(np.abs(u[:, 0, 0] - st["au0"]))), float(np.max(np.abs(w[:, 0, 0] - st["w0"]))))
    ref = before if before is not None else {k: np.zeros_like(st[k]) for k in ("vnp1", "omegaynp1")}
    return {
        "divergence": float(np.max(np.abs(div))) / max(vs, us),
        "omega_y_curl": float(np.max(np.abs(curl))) / float(np.max(np.abs(oy))),
        "hermitian_kx0": herm,
        "mean_curl": mean_curl,
        "mean_velocity": mean_vel,
        "nyquist": max(float(np.max(np.abs(v[:, mz // 2, :] - ref["vnp1"][:, mz // 2, :]))),
                       float(np.max(np.abs(oy[:, mz // 2, :] - ref["omegaynp1"][:, mz // 2, :])))),
        "frozen_00": max(float(np.max(np.abs(v[:, 0, 0] - ref["vnp1"][:, 0, 0]))),
                         float(np.max(np.abs(oy[:, 0, 0] - ref["omegaynp1"][:, 0, 0])))),
    }
