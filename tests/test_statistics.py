"""Flow statistics (statistics, PF:2816-3190).

CPU: the host diagnostics on the laminar profile U = U_c (1 - y^2) reproduce
the closed forms (wall shear velocity sqrt(re_p) with re_p = 2 U_c / re,
centreline Reynolds number U_c re, CFL_x = U_c dt / dx, CFL_y = 0).
GPU: chf_physical (transform(1, .)) and chf_statistics against the oracle's
restatement on the fields after a few steps.
"""
import math

import numpy as np
import pytest

from channel_flow_amd import statistics
from oracle import cf_oracle as O

from .helpers import SMALL, ic_of, params_of, rel_err


def test_laminar_diagnostics_closed_form():
    nnx, nny, N = 16, 8, 24
    re, uc, dt, aphi, beta = 3000.0, 1.0, 1e-3, 1.0, 0.5
    re_p = 2 * uc / re
    y = np.cos(np.pi * np.arange(N + 1) / N)
    u = np.broadcast_to((uc * (1 - y ** 2))[:, None, None], (N + 1, nny, nnx))
    zero = np.zeros_like(u)
    ps = O.plane_statistics(u, zero, zero)
    d = statistics.diagnostics(ps, nnx, nny, N, aphi, beta, dt, re, re_p)
    assert abs(d["wall_shear_velocity"][0] - math.sqrt(re_p)) < 1e-12
    assert abs(d["wall_shear_velocity"][1] - math.sqrt(re_p)) < 1e-12
    assert abs(d["re_centre"] - uc * re) < 1e-9
    assert abs(d["cfl_x"] - uc * dt / (2 * math.pi * aphi / nnx)) < 1e-15
    assert d["cfl_y"] == 0.0 and d["cfl_z"] == 0.0
    np.testing.assert_allclose(d["total_shear_stress"], -2 * uc * y / re, rtol=0, atol=1e-14)


@pytest.mark.gpu
def test_physical_fields_and_statistics_match_oracle(require_gpu):
    from channel_flow_amd.channel import Channel
    cfg = SMALL
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    with Channel(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"],
                 cfg["re_p"]) as ch:
        ch.initial(v, oy, au0, w0)
        ch.step(3)
        st, fl = ch.get_state(), ch.fields()
        spec = {"u": fl["unp1"], "v": st["vnp1"], "w": fl["wnp1"], "omega_x": fl["omegaxnp1"],
                "omega_y": st["omegaynp1"], "omega_z": fl["omegaznp1"]}
        phys = {k: ch.physical(k) for k in spec}
        ps = ch.plane_statistics()
        diag = ch.statistics()
    ref = {k: O.physical_field(f, p) for k, f in spec.items()}
    for k in spec:
        e = rel_err(phys[k], ref[k])
        print(f"physical {k}: {e:.2e}", flush=True)
        assert e < 1e-12, (k, e)
    rps = O.plane_statistics(ref["u"], ref["v"], ref["w"])
    # each statistic against the size of the fields it is made of: the plane
    # means of v and w are zero up to rounding, so a relative error is no measure
    mx = [float(np.max(rps[9 + i])) for i in range(3)]
    scale = mx + [mx[0] ** 2, mx[1] ** 2, mx[2] ** 2, mx[0] * mx[1], mx[0] * mx[2], mx[1] * mx[2]] + mx
    for q in range(12):
        e = float(np.max(np.abs(ps[q] - rps[q]))) / scale[q]
        assert e < 1e-12, (q, e)
    assert all(np.isfinite([diag["cfl_x"], diag["cfl_y"], diag["cfl_z"], diag["re_bulk"]]))
    print("diagnostics:", {k: diag[k] for k in ("cfl_x", "cfl_y", "cfl_z", "wall_shear_velocity", "re_tau",
                                                 "re_centre", "re_bulk")}, flush=True)
