"""Timestep throughput of the MI355X channel-flow core (one GPU).

    python bench.py [--steps K] [--warmup W] [--config retau395_256x193x256]
                    [--precision fp64] [--no-cpu-baseline]

A step is one pass of the per-timestep core (PF:220-532): per-mode v/omega_y
solves, kinematics and mean mode, then getnonlinear (six 3/2-dealiased
products, Lamb vector, h assembly, AB2 rotation).  Inputs are resident in HBM
when the timed region starts.  The default workload is the 256 x 193 x 256
channel (nnx = nny = 256, N = 192, Lx = 2 pi, Lz = pi, Re_tau ~ 395 class, in
centreline units: U_c = 1, re = 7940, dt = 2.5e-4) from a seeded
divergence-free initial condition.  Prints one JSON line: `value` is the
mixed form and `fp64_value` the FP64 form (wall-normal transforms as FFTs
along y in both); `mixed_matrix_value` and `fp64_matrix_value` are the two
forms with those transforms as cos-table products on the matrix cores;
`mixed_over_fp64` and `mixed_over_fp64_matrix` are the ratios.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, GB/s
FP64_PEAK_TFS = 78.6         # MI355X dense FP64 (matrix and vector), TFLOP/s
FP32_PEAK_TFS = 157.3        # MI355X dense FP32 matrix, TFLOP/s


def dims(c):
    N = c["nnz"]
    d = dict(Mx=c["nnx"] // 2, Mz=c["nny"], N=N, Ny=N + 1, Lx=3 * c["nnx"] // 2, Lz=3 * c["nny"] // 2,
             Py=2 * N + 1, Ph=N + 1, ne=N // 2 + 1, no=(N + 1) // 2, Pn=N // 2 + 1)
    d["ncol"] = d["Mz"] * d["Mx"]
    d["ngroups"] = d["Mx"] * (d["Mz"] // 2)
    # operator classes: groups with equal k^2 share one set of inverses
    al = [i / c["aphi"] for i in range(d["Mx"])]
    be = [(j if j <= d["Mz"] // 2 else j - d["Mz"]) / c["beta"] for j in range(d["Mz"])]
    d["nclass"] = len({al[g % d["Mx"]] ** 2 + be[g // d["Mx"]] ** 2 for g in range(d["ngroups"])})
    return d


def kernel_work(d, precision="fp64", yfft=False):
    """Algorithmic HBM bytes and flops per launch of each kernel class.

    Each stage reads its inputs once and writes its outputs once at the
    storage precision: spectral fields and the solve are FP64 (16 B per
    complex value); the wall-normal/span point arrays of the nonlinear term are
    FP64 in the FP64 form and FP32 (8 B) in the mixed form.  Flops count the
    useful multiply-adds (2 flops each) of the parity-split contractions.
    With yfft (the wall-normal transforms run as FFTs along y) the padded
    synthesis, the analysis, the point values and the right-hand sides have
    O(N log N) work per column and are priced by their bytes alone, and the
    right-hand-side kernels form h^n's point values themselves (they read
    h^n's coefficients and h^(n-1)'s point values and write h^n's): there is
    no separate point-value pass.  In the mixed form with yfft the kinematics
    writes u, w, omega_x, omega_z in FP32 and the padded synthesis reads them
    so (v and omega_y stay FP64), and the analysis writes L^n's coefficients
    in FP32, which the h assembly reads so.
    """
    w = 16 if precision == "fp64" else 8
    S = d["Ny"] * d["ncol"] * 16            # spectral field (FP64)
    Y = d["Py"] * d["ncol"] * w             # wall-normal points, Fourier modes
    Z = d["Py"] * d["Lz"] * d["Mx"] * w     # wall-normal + span points
    col = 2 * d["ncol"]                     # real columns
    ne, no, Ny, Ph, Pn = d["ne"], d["no"], d["Ny"], d["Ph"], d["Pn"]
    rv_rows = (ne - 2) + (no - 2)
    ro_rows = (ne - 1) + (no - 1)
    inv_v_bytes = (d["nclass"] - 1) * (ne * ne + no * no) * 8   # one per k^2 class; k^2 = 0 has no v system
    ws_flops = 2.0 * col * ((ne - 1) * (ne - 1) + ne * (ne - 1) + (no - 1) * (no - 1) + no * (no - 1))
    yf = 0.0 if yfft else 1.0
    f32f = precision != "fp64" and yfft      # u, w, omega_x, omega_z held in FP32
    S_in = (4 * S / 2 + 2 * S) if f32f else 6 * S
    return {
        "synth_padded": (S_in + 6 * Y, yf * 2.0 * Ph * Ny * col * 6),
        "zfft_inv": (6 * (Y + Z), 0.0),
        "xpass": (9 * Z, 0.0),
        "zfft_fwd": (3 * (Z + Y), 0.0),
        "analysis": (3 * (Y + (S / 2 if f32f else S)), yf * 2.0 * Ny * Ph * col * 3),
        "hassemble": ((3 * S / 2 if f32f else 3 * S) + 2 * S + 4 * Ny * 16, 0.0),   # plus the mean-mode point block
        "synth_points": (4 * S, yf * 2.0 * Pn * Ny * col * 2),
        "rhs_v": ((4 * S if yfft else S + 2 * (Ny - 4) / Ny * S) + rv_rows / Ny * S,
                  yf * 2.0 * ((ne - 2) * ne + (no - 2) * no) * col),
        "rhs_o": ((4 * S if yfft else S + 2 * (Ny - 2) / Ny * S) + ro_rows / Ny * S,
                  yf * 2.0 * ((ne - 1) * ne + (no - 1) * no) * col),
        # v solve (omega_y of the mean group only) and the diagonalised omega_y solve
        "solve": (inv_v_bytes + rv_rows / Ny * S + S, 0.0),
        "wsolve": (ro_rows / Ny * S + S, ws_flops),
        # reads v, omega_y; writes u, w, omega_x, omega_z and the two brackets
        "kinematics": ((2 * S + 4 * S / 2 + 2 * S) if f32f else 8 * S, 0.0),
    }


def roofline_of(name, avg_ms, work, precision="fp64"):
    """Roofline entry of one kernel class from its average launch time."""
    nbytes, flops = work[name]
    sec = avg_ms * 1e-3
    gbs = nbytes / sec / 1e9
    tfs = flops / sec / 1e12 if flops else 0.0
    fp32_mm = precision == "mixed" and name in ("synth_padded", "analysis")
    peak_tf = FP32_PEAK_TFS if fp32_mm else FP64_PEAK_TFS
    # bound = the resource whose time at peak is larger
    if flops and flops / (peak_tf * 1e12) > nbytes / (HBM_PEAK_GBS * 1e9):
        return dict(kernel=name, bound="mfma", achieved=round(tfs, 3), peak=peak_tf,
                    unit="TFLOP/s", frac=round(tfs / peak_tf, 4), avg_ms=round(avg_ms, 4),
                    algorithmic_bytes=int(nbytes), flops=flops, achieved_GBs=round(gbs, 1))
    return dict(kernel=name, bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(gbs / HBM_PEAK_GBS, 4), avg_ms=round(avg_ms, 4),
                algorithmic_bytes=int(nbytes))


# HIP kernel symbol behind each timed kernel class (the FP32 forms share symbols
# with the FP64 ones; template arguments differ).
KERNEL_SYMBOL = {"synth_padded": "k_mfma_cos", "zfft_inv": "k_zfft", "xpass": "k_xpass_st",
                 "zfft_fwd": "k_zfft", "analysis": "k_mfma_cos", "hassemble": "k_hassemble",
                 "synth_points": "k_mfma_cos",
                 "rhs_v": "k_mfma_cos", "rhs_o": "k_mfma_cos",
                 "solve": "k_solve", "kinematics": "k_kinematics", "wsolve": "k_wsolve"}


def measured_traffic(config, form, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a device copy of known
    size), or None when no summary for this workload and form is present or
    the kernel's symbol is shared by several kernel classes."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{config}_{form}.json")
    sym = KERNEL_SYMBOL.get(kernel)
    if not os.path.exists(path) or sym is None or list(KERNEL_SYMBOL.values()).count(sym) != 1:
        return None
    with open(path) as f:
        data = json.load(f)["kernels"]
    hits = [v for k, v in data.items() if k.split("<")[0].split()[-1] == sym]
    if len(hits) != 1 or hits[0]["fetch_bytes"] is None or hits[0]["write_bytes"] is None:
        return None
    return int(hits[0]["fetch_bytes"] + hits[0]["write_bytes"])


def copy_probe(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of `nbytes`, run under the FETCH_SIZE and
    WRITE_SIZE counters to calibrate them (the copy moves exactly nbytes each
    way)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)
    return 0


def run_ensemble(steps, warmup):
    """The batched-ensemble workload: 24 channels of mixed sizes and Reynolds
    numbers, one hipGraph launch per step, against the same members stepped by
    eager launches on their own streams.  Prints its own JSON line."""
    from channel_flow_amd import ic
    from channel_flow_amd.ensemble import Ensemble
    members = ic.ensemble_members()
    with Ensemble(members) as ens:
        for i, m in enumerate(members):
            ens.member(i).initial(*ic.seeded_ic(m["nnx"], m["nny"], m["nnz"], m["aphi"], m["beta"], m["u_c"],
                                                m["rms"], m["kmax"], m["slope"], m["seed"]))
        ens.step(warmup)
        t0 = time.perf_counter()
        ens.step(steps)
        t_graph = time.perf_counter() - t0
        # same members, eager launches: each member's steps queued on its own stream
        for i in range(len(members)):
            ens.member(i).step(warmup)
        for i in range(len(members)):
            ens.member(i).sync()
        t0 = time.perf_counter()
        for i in range(len(members)):
            ens.member(i).step(steps)
        for i in range(len(members)):
            ens.member(i).sync()
        t_eager = time.perf_counter() - t0
        finite = all(np.all(np.isfinite(ens.member(i).get_state()["vnp1"])) for i in range(len(members)))
    sizes = {}
    step_bytes = 0.0
    for m in members:
        key = f"{m['nnx']}x{m['nnz'] + 1}x{m['nny']}"
        sizes[key] = sizes.get(key, 0) + 1
        work = kernel_work(dims(m), m.get("precision", "mixed"))
        step_bytes += sum(nbytes for nbytes, _ in work.values())
    agg_gbs = step_bytes / (t_graph / steps) / 1e9
    out = {
        "metric": "ensemble timesteps/s (24 channels advanced together, one hipGraph launch per step)",
        "value": round(steps / t_graph, 3), "unit": "ensemble steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(t_graph * 1e3 / steps, 4), "higher_is_better": True,
        "member_steps_per_s": round(len(members) * steps / t_graph, 1),
        # algorithmic bytes of one step of every member (each kernel class reading
        # its inputs and writing its outputs once) over the measured step time
        "aggregate_hbm": {"achieved": round(agg_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(agg_gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_step": int(step_bytes)},
        "eager_streams_ms_per_step": round(t_eager * 1e3 / steps, 4),
        "dtype": "f32 transforms and products, f64 solves and state", "data": "synthetic",
        "config": {"workload": "ensemble24: re_c in [2000, 6000], dt=1e-3, mixed precision",
                   "members": sizes},
        "state_finite": bool(finite),
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(cfg, seconds_budget=30.0):
    """Oracle (numpy restatement of the reference) on one core, bounded sample.

    Sample: one of the six dealiased products (getlamb) at the full workload
    size, the h assembly of getnonlinear on the full field, and the per-mode
    solve of 256 randomly chosen modes; scaled to one full step as
    6 * t_product + t_h + (modes / 256) * t_solve256.
    """
    from threadpoolctl import threadpool_limits
    from oracle import cf_oracle as O
    from channel_flow_amd import ic
    p = O.Params(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"],
                 cfg["re_p"])
    v, oy, au0, w0 = ic.seeded_ic(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"],
                                  cfg["u_c"], cfg["rms"], cfg["kmax"], cfg["slope"], cfg["seed"])
    with threadpool_limits(1):
        t = time.perf_counter()
        O.getlamb(v, oy, p)
        t_prod = time.perf_counter() - t
        t = time.perf_counter()
        alpha, beta = O.wavenumbers(p)
        k2 = alpha[None, None, :] ** 2 + beta[None, :, None] ** 2
        dlx = O.dz(v)
        dlz = O.dz(oy)
        hv = k2 * v + 1j * alpha[None, None, :] * dlx + 1j * beta[None, :, None] * dlz
        hg = 1j * alpha[None, None, :] * oy - 1j * beta[None, :, None] * v
        O.points_from_coeffs(hv.real, p.N)
        O.points_from_coeffs(hg.real, p.N)
        t_h = time.perf_counter() - t
        st = O.zero_state(p)
        st["vnp1"], st["omegaynp1"] = v.copy(), oy.copy()
        stp = O.Stepper(p)
        rng = np.random.default_rng(0)
        pick = rng.choice(len(stp.solve_modes), size=256, replace=False)
        stp.solve_modes = [stp.solve_modes[i] for i in sorted(pick)]
        for (_, _, kmn) in stp.solve_modes:
            stp._inverses(kmn)                       # set-up, not timed
        t = time.perf_counter()
        stp.solve(st)
        t_solve = time.perf_counter() - t
    nmodes = p.Mx * p.Mz - p.Mx - 1      # all modes but the Nyquist column and (0,0) (PF:247)
    sec_per_step = 6 * t_prod + t_h + nmodes / 256.0 * t_solve
    return {"value": round(1.0 / sec_per_step, 6), "unit": "timesteps/s", "cores": 1,
            "kind": "port",
            "sample": (f"numpy oracle, 1 thread: one getlamb product at full size ({t_prod:.2f} s) x6, "
                       f"h assembly ({t_h:.2f} s), 256-mode solve ({t_solve:.2f} s) x{nmodes}/256 "
                       f"-> {sec_per_step:.1f} s/step")}


def run_form(cfg, ic_arrays, precision, steps, warmup, total, log, y_matrix=False):
    """Set-up, warm-up, the timed region (wall clock, no per-kernel events),
    the same number of steps again with per-kernel HIP events (the launches
    then run one at a time), then untimed steps up to `total` for the error
    report.  y_matrix: the wall-normal transforms as cos-table products on
    the matrix cores (CHF_Y_TRANSFORMS=matrix) instead of FFTs."""
    from channel_flow_amd.channel import Channel
    v, oy, au0, w0 = ic_arrays
    t0 = time.perf_counter()
    if y_matrix:
        os.environ["CHF_Y_TRANSFORMS"] = "matrix"
    try:
        ch = Channel(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"],
                     cfg["re_p"], precision=precision)
    finally:
        os.environ.pop("CHF_Y_TRANSFORMS", None)
    t_setup = time.perf_counter() - t0
    log(f"{precision}: setup {t_setup:.2f}s, device bytes {ch.device_bytes() / 1e9:.2f} GB")
    tiles = ch.tile_choices()
    ch.initial(v, oy, au0, w0)
    ch.step(warmup)
    ch.sync()
    t0 = time.perf_counter()
    ch.step(steps)
    ch.sync()
    wall = time.perf_counter() - t0
    ch.set_timing(True)
    ch.step(steps)
    kt = ch.kernel_times()
    ch.set_timing(False)
    rest = total - warmup - 2 * steps
    if rest > 0:
        ch.step(rest)
    st = ch.get_state()
    fl = ch.fields()
    ch.close()
    per_kernel = {k: dict(total_ms=round(ms, 3), launches=n, avg_ms=round(ms / n, 4) if n else None)
                  for k, (ms, n) in kt.items() if n}
    # the transforms ran as FFTs when they have no tile choice
    yfft = "synth_padded" not in tiles
    return dict(ms_per_step=wall * 1e3 / steps, per_kernel=per_kernel, state=st, fields=fl,
                setup_s=t_setup, tiles=tiles, yfft=yfft)


def field_errors(cfg, mixed, ref):
    """Errors of the mixed form against the FP64 form after the same steps."""
    from channel_flow_amd import ic as icmod
    def l2(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out = {}
    out["v"] = l2(mixed["state"]["vnp1"], ref["state"]["vnp1"])
    out["omega_y"] = l2(mixed["state"]["omegaynp1"], ref["state"]["omegaynp1"])
    for k, n in (("unp1", "u"), ("wnp1", "w"), ("omegaxnp1", "omega_x"), ("omegaznp1", "omega_z")):
        out[n] = l2(mixed["fields"][k], ref["fields"][k])
    N = cfg["nnz"]
    kk = np.arange(N + 1)
    # wall shear dU/dy / re at y = +1 and -1 (T_k'(+-1) = (+-1)^(k+1) k^2)
    def shear(a):
        return (np.sum(a * kk ** 2).real / cfg["re"], np.sum(a * kk ** 2 * (-1.0) ** (kk + 1)).real / cfg["re"])
    sm, sr = shear(mixed["state"]["au0"]), shear(ref["state"]["au0"])
    out["wall_shear_top"] = abs(sm[0] - sr[0]) / abs(sr[0])
    out["wall_shear_bottom"] = abs(sm[1] - sr[1]) / abs(sr[1])
    y = np.cos(np.pi * kk / N)
    Um = np.polynomial.chebyshev.chebval(y, mixed["state"]["au0"].real)
    Ur = np.polynomial.chebyshev.chebval(y, ref["state"]["au0"].real)
    out["mean_profile"] = float(np.max(np.abs(Um - Ur)) / np.max(np.abs(Ur)))
    # divergence i alpha u + Dv + i beta w of the mixed fields, relative to |Dv|
    alpha, beta = icmod.wavenumbers(cfg["nnx"], cfg["nny"], cfg["aphi"], cfg["beta"])
    dv = icmod.cheb_deriv(mixed["state"]["vnp1"])
    div = 1j * alpha[None, None, :] * mixed["fields"]["unp1"] + dv + \
        1j * beta[None, :, None] * mixed["fields"]["wnp1"]
    div[:, cfg["nny"] // 2, :] = 0
    div[:, 0, 0] = 0
    out["max_divergence_rel"] = float(np.max(np.abs(div)) / max(np.max(np.abs(dv)), 1e-300))
    return {k: float("%.3e" % v) for k, v in out.items()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="retau395_256x193x256")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--error-steps", type=int, default=100,
                    help="steps after which the mixed form is compared with the FP64 form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble", action="store_true",
                    help="run the 24-channel batched-ensemble workload instead (its own JSON line)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a 1 GiB device copy (counter calibration)")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    if a.copy_probe:
        return copy_probe()
    if a.ensemble and a.gpus == 1:
        return run_ensemble(a.steps, a.warmup)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the whole channel lives on one device "
              "(multi-GPU sharding is a design note, not a run)", file=sys.stderr)
        return 2

    from channel_flow_amd import ic

    cfg = dict(ic.CONFIGS[a.config])
    d = dims(cfg)
    log = (lambda *x: print(*x, file=sys.stderr, flush=True)) if a.verbose else (lambda *x: None)
    t0 = time.perf_counter()
    ic_arrays = ic.seeded_ic(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["u_c"],
                             cfg["rms"], cfg["kmax"], cfg["slope"], cfg["seed"])
    log(f"ic {time.perf_counter() - t0:.2f}s")
    forms = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    total = max(a.error_steps, a.warmup + 2 * a.steps) if a.precision == "both" else a.warmup + 2 * a.steps
    runs = {f: run_form(cfg, ic_arrays, f, a.steps, a.warmup, total, log) for f in forms}
    if a.precision == "both":
        # both forms with their transforms on the matrix cores, for comparison
        for f in forms:
            if runs[f]["yfft"]:
                runs[f + "_matrix"] = run_form(cfg, ic_arrays, f, a.steps, a.warmup, a.warmup + 2 * a.steps, log,
                                               y_matrix=True)
    head = "mixed" if "mixed" in runs else "fp64"
    r = runs[head]
    finite = all(np.all(np.isfinite(r["state"][k])) for k in ("vnp1", "omegaynp1", "au0"))
    work = kernel_work(d, head, r["yfft"])
    pk = r["per_kernel"]
    dom = max(pk, key=lambda k: pk[k]["total_ms"])
    roof = roofline_of(dom, pk[dom]["avg_ms"], work, head)
    roof["traffic"] = measured_traffic(a.config, head, dom)
    for f, rr in runs.items():
        pf = "mixed" if f.startswith("mixed") else "fp64"
        w = kernel_work(d, pf, rr["yfft"])
        log(f, json.dumps({k: roofline_of(k, rr["per_kernel"][k]["avg_ms"], w, pf)
                           for k in rr["per_kernel"]}, indent=1))
    cpu = None if a.no_cpu_baseline else cpu_baseline(cfg)
    out = {
        "metric": "DNS timesteps/s at 256\u00d7193\u00d7256, mixed vs FP64 GPU form, %HBM/MFMA roofline",
        "value": round(1e3 / r["ms_per_step"], 3),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 transforms and products, f64 solves and state" if head == "mixed" else "f64"),
        "data": "synthetic: seeded divergence-free Fourier-Chebyshev perturbation on a parabola",
        "config": {"workload": f"{a.config}: nnx={cfg['nnx']} nny={cfg['nny']} N={cfg['nnz']} "
                               f"aphi={cfg['aphi']} beta={cfg['beta']} re={cfg['re']} "
                               f"re_p={cfg['re_p']} dt={cfg['dt']}",
                   "form": head, "padded_grid": f"{d['Lx']}x{d['Py']}x{d['Lz']}",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": pk,
        "setup_s": round(r["setup_s"], 2),
        "autotuned_tiles": r["tiles"],
        "state_finite": bool(finite),
    }
    if "fp64" in runs and "mixed" in runs:
        out["fp64_value"] = round(1e3 / runs["fp64"]["ms_per_step"], 3)
        out["mixed_over_fp64"] = round(runs["fp64"]["ms_per_step"] / runs["mixed"]["ms_per_step"], 3)
        out["fp64_kernels"] = runs["fp64"]["per_kernel"]
        out["y_transforms"] = "fft" if runs["fp64"]["yfft"] else "matrix"
        if "fp64_matrix" in runs and "mixed_matrix" in runs:
            fm, mm = runs["fp64_matrix"]["ms_per_step"], runs["mixed_matrix"]["ms_per_step"]
            out["fp64_matrix_value"] = round(1e3 / fm, 3)
            out["mixed_matrix_value"] = round(1e3 / mm, 3)
            out["mixed_over_fp64_matrix"] = round(fm / mm, 3)
            out["matrix_kernels"] = {f: {k: runs[f + "_matrix"]["per_kernel"][k]["avg_ms"]
                                         for k in ("synth_padded", "analysis", "synth_points")}
                                     for f in ("fp64", "mixed")}
        out["errors_vs_fp64"] = dict(field_errors(cfg, runs["mixed"], runs["fp64"]), steps=total)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
