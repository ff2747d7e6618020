"""Timestep throughput of the MI355X channel-flow core (one GPU).

    python bench.py [--steps K] [--warmup W] [--config cfg3_256x193x256]
                    [--precision fp64] [--no-cpu-baseline]

A step is one pass of the per-timestep core (PF:220-532): per-mode v/omega_y
solves, kinematics and mean mode, then getnonlinear (six 3/2-dealiased
products, Lamb vector, h assembly, AB2 rotation).  Inputs are resident in HBM
when the timed region starts.  The default workload is the 256 x 193 x 256
channel (nnx = nny = 256, N = 192, Lx = 2 pi, Lz = pi, Re_tau ~ 395 class) from
a seeded divergence-free initial condition.  Prints one JSON line.
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


def dims(c):
    N = c["nnz"]
    d = dict(Mx=c["nnx"] // 2, Mz=c["nny"], N=N, Ny=N + 1, Lx=3 * c["nnx"] // 2, Lz=3 * c["nny"] // 2,
             Py=2 * N + 1, Ph=N + 1, ne=N // 2 + 1, no=(N + 1) // 2, Pn=N // 2 + 1)
    d["ncol"] = d["Mz"] * d["Mx"]
    d["ngroups"] = d["Mx"] * (d["Mz"] // 2)
    return d


def kernel_work(d, word=16):
    """Algorithmic HBM bytes and flops per launch of each kernel class.

    Each stage reads its inputs once and writes its outputs once at the
    storage precision (complex = `word` bytes); flops count the useful
    multiply-adds of the parity-split contractions (2 flops each).
    """
    S = d["Ny"] * d["ncol"] * word           # spectral field
    Y = d["Py"] * d["ncol"] * word           # wall-normal points, Fourier modes
    Z = d["Py"] * d["Lz"] * d["Mx"] * word   # wall-normal + span points
    col = 2 * d["ncol"]                      # real columns
    ne, no, Ny, Ph, Pn = d["ne"], d["no"], d["Ny"], d["Ph"], d["Pn"]
    rv_rows = (ne - 2) + (no - 2)
    ro_rows = (ne - 1) + (no - 1)
    inv_bytes = 2 * d["ngroups"] * (ne * ne + no * no) * 8 - (ne * ne + no * no) * 8
    return {
        "synth_padded": (6 * (S + Y), 2.0 * Ph * Ny * col * 6),
        "zfft_inv": (6 * (Y + Z), 0.0),
        "xpass": (9 * Z, 0.0),
        "zfft_fwd": (3 * (Z + Y), 0.0),
        "analysis": (3 * (Y + S), 2.0 * Ny * Ph * col * 3),
        "hassemble": (5 * S, 0.0),
        "synth_points": (4 * S, 2.0 * Pn * Ny * col * 2),
        "mean_points": (4 * Ny * 16, 0.0),
        "rhs_v": (S + 2 * (Ny - 4) / Ny * S + rv_rows / Ny * S, 3 * 2.0 * (rv_rows * ne) * col),
        "rhs_o": (S + 2 * (Ny - 2) / Ny * S + ro_rows / Ny * S, 2 * 2.0 * (ro_rows * ne) * col),
        "solve": (inv_bytes + (rv_rows + ro_rows) / Ny * S + 6 * S, 0.0),
    }


def roofline_of(name, avg_ms, work):
    nbytes, flops = work[name]
    sec = avg_ms * 1e-3
    gbs = nbytes / sec / 1e9
    tfs = flops / sec / 1e12 if flops else 0.0
    # bound = the resource whose time at peak is larger
    if flops and flops / (FP64_PEAK_TFS * 1e12) > nbytes / (HBM_PEAK_GBS * 1e9):
        return dict(kernel=name, bound="mfma", achieved=round(tfs, 3), peak=FP64_PEAK_TFS,
                    unit="TFLOP/s", frac=round(tfs / FP64_PEAK_TFS, 4), avg_ms=round(avg_ms, 4),
                    algorithmic_bytes=int(nbytes), flops=flops, achieved_GBs=round(gbs, 1))
    return dict(kernel=name, bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(gbs / HBM_PEAK_GBS, 4), avg_ms=round(avg_ms, 4),
                algorithmic_bytes=int(nbytes))


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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="cfg3_256x193x256")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the whole channel lives on one device "
              "(multi-GPU sharding is a design note, not a run)", file=sys.stderr)
        return 2

    from channel_flow_amd import ic
    from channel_flow_amd.channel import Channel

    cfg = dict(ic.CONFIGS[a.config])
    d = dims(cfg)
    log = (lambda *x: print(*x, file=sys.stderr, flush=True)) if a.verbose else (lambda *x: None)
    t0 = time.perf_counter()
    v, oy, au0, w0 = ic.seeded_ic(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"],
                                  cfg["u_c"], cfg["rms"], cfg["kmax"], cfg["slope"], cfg["seed"])
    log(f"ic {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    ch = Channel(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"],
                 cfg["re_p"], precision=a.precision)
    t_setup = time.perf_counter() - t0
    log(f"setup {t_setup:.2f}s, device bytes {ch.device_bytes() / 1e9:.2f} GB")
    ch.initial(v, oy, au0, w0)
    ch.step(a.warmup)
    ch.sync()
    log("warmup done")
    ch.set_timing(True)
    ch.sync()
    t0 = time.perf_counter()
    ch.step(a.steps)
    ch.sync()
    wall = time.perf_counter() - t0
    kt = ch.kernel_times()
    ch.set_timing(False)
    st = ch.get_state()
    finite = all(np.all(np.isfinite(st[k])) for k in ("vnp1", "omegaynp1", "au0"))
    ms_per_step = wall * 1e3 / a.steps
    work = kernel_work(d, 16)
    per_kernel = {k: dict(total_ms=round(ms, 3), launches=n, avg_ms=round(ms / n, 4) if n else None)
                  for k, (ms, n) in kt.items() if n}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["total_ms"])
    roof = roofline_of(dom, per_kernel[dom]["avg_ms"], work)
    roof["traffic"] = None
    kernel_roofs = {k: roofline_of(k, per_kernel[k]["avg_ms"], work) for k in per_kernel}
    log(json.dumps(kernel_roofs, indent=1))
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(cfg)
    out = {
        "metric": "DNS timesteps/s at 256x193x256 (mixed vs FP64 GPU form, %HBM/MFMA roofline)",
        "value": round(1e3 / ms_per_step, 3),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32/bf16-split transforms, f64 solves",
        "data": "synthetic: seeded divergence-free Fourier-Chebyshev perturbation on a parabola",
        "config": {"workload": f"{a.config}: nnx={cfg['nnx']} nny={cfg['nny']} N={cfg['nnz']} "
                               f"aphi={cfg['aphi']} beta={cfg['beta']} re={cfg['re']} "
                               f"re_p={cfg['re_p']} dt={cfg['dt']}",
                   "precision": a.precision, "padded_grid": f"{d['Lx']}x{d['Py']}x{d['Lz']}",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": per_kernel,
        "setup_s": round(t_setup, 2),
        "state_finite": bool(finite),
    }
    print(json.dumps(out), flush=True)
    ch.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
