"""The Fortran driver (PF's time loop over ISO_C_BINDING) end to end.

GPU: a restart file holding the oracle's cold-start state is advanced 10 steps
by `poiseuilleflow_gpu`; the restart record it writes at step 10 (output,
PF:2460-2504) is the oracle's pre-solve state after 10 steps.
CPU: without a device the driver stops at chf_create with the library's
message and a non-zero exit status.
"""
import os
import subprocess

import numpy as np
import pytest

from channel_flow_amd import restart
from channel_flow_amd.build import DRIVER
from oracle import cf_oracle as O

from .helpers import SMALL, ic_of, params_of, state_errors

TRAJ_TOL = 1e-9


def namelist(path, cfg, **extra):
    keys = dict(nnx=cfg["nnx"], nny=cfg["nny"], nnz=cfg["nnz"], aphi=cfg["aphi"], beta=cfg["beta"],
                dt=cfg["dt"], re=cfg["re"], re_p=cfg["re_p"])
    keys.update(extra)
    lines = ["&channel"]
    for k, v in keys.items():
        if isinstance(v, str):
            lines.append(f"  {k} = '{v}'")
        elif isinstance(v, float):
            lines.append(f"  {k} = " + ("%.17e" % v).replace("e", "d"))
        else:
            lines.append(f"  {k} = {v}")
    lines.append("/")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (python -m channel_flow_amd.build)")


def test_driver_reports_create_failure_without_a_device(tmp_path):
    need_driver()
    from channel_flow_amd import runtime
    if runtime.device_count() > 0:
        pytest.skip("a HIP device is present")
    nml = tmp_path / "run.nml"
    namelist(nml, SMALL, nstep=2, interupt=1, isign=-1, output_dir=str(tmp_path))
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "chf_create failed" in r.stdout + r.stderr


@pytest.mark.gpu
def test_driver_ten_steps_match_oracle(require_gpu, tmp_path):
    need_driver()
    cfg = SMALL
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    st = O.initial_presolve_state(v, oy, au0, w0, p)
    rin = tmp_path / "start.dat"
    restart.write_restart(str(rin), st)
    outdir = tmp_path / "out"
    outdir.mkdir()
    nml = tmp_path / "run.nml"
    namelist(nml, cfg, istepbegin=0, nstep=10, interupt=10, outputstep=5, isign=1, precision=0,
             restart_in=str(rin), output_dir=str(outdir))
    print("driver run:", DRIVER, nml, flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert r.stdout.count("wall shear velocity") == 2   # statistics at steps 5 and 10
    got = restart.read_restart(str(outdir / restart.restart_name(0)), cfg["nnx"], cfg["nny"], cfg["nnz"])
    O.Stepper(p).step(st, 10)
    errs = state_errors(got, st)
    print("fortran driver 10 steps", " ".join(f"{k}={e:.2e}" for k, e in errs.items()), flush=True)
    for k, e in errs.items():
        assert e < TRAJ_TOL, (k, e)
