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
    # statistics at steps 5 and 10 (outputstep), the profiles at 10 only (interupt, PF:2793)
    assert r.stdout.count("engery=") == 2 and r.stdout.count("wall shear velocity") == 1
    got = restart.read_restart(str(outdir / restart.restart_name(0)), cfg["nnx"], cfg["nny"], cfg["nnz"])
    O.Stepper(p).step(st, 10)
    errs = state_errors(got, st)
    print("fortran driver 10 steps", " ".join(f"{k}={e:.2e}" for k, e in errs.items()), flush=True)
    for k, e in errs.items():
        assert e < TRAJ_TOL, (k, e)


def read_profile(path):
    return np.loadtxt(str(path), ndmin=2)


@pytest.mark.gpu
def test_driver_writes_the_reference_output_files(require_gpu, tmp_path):
    """statistics (PF:2666-3198) from the driver: with outputstep = 5 and
    interupt = 10 over 10 steps it writes 3Ddata/velocity0000000{5,10}.000.dat
    (one record of real*8 (0:nx,0:ny,0:nz,3) = (u + u0, w, v)), two energy.plt
    lines, and at step 10 the meanu/, stress/, rms/, uplus/ profiles and a
    diagnostics.plt line.  Each is checked against the same channel stepped
    through the Python binding: the dump bit for bit against chf_physical, the
    rest against the reference's formulas evaluated here in numpy on those
    fields.  The dump also reads back through the post-processing read
    statement (tests/fortran/velocity_reader.f90)."""
    from channel_flow_amd.build import flang
    from channel_flow_amd.channel import Channel
    from channel_flow_amd import statistics as S
    need_driver()
    cfg = SMALL
    p = params_of(cfg)
    st = O.initial_presolve_state(*ic_of(cfg), p)
    rin = tmp_path / "start.dat"
    restart.write_restart(str(rin), st)
    outdir = tmp_path / "out"
    nml = tmp_path / "run.nml"
    namelist(nml, cfg, istepbegin=0, nstep=10, interupt=10, outputstep=5, isign=1, precision=0,
             restart_in=str(rin), output_dir=str(outdir))
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0

    nnx, nny, N = cfg["nnx"], cfg["nny"], cfg["nnz"]
    phys = {}
    stats = {}
    for istep in (5, 10):
        with Channel(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"],
                     cfg["re_p"]) as ch:
            ch.set_state(st)
            ch.step(istep - 1)           # the driver's first iteration only solves (PF:222)
            ch.solve()
            phys[istep] = np.stack([ch.physical("u"), ch.physical("w"), ch.physical("v")])
            stats[istep] = ch.plane_statistics()
    energy = read_profile(outdir / "energy.plt")
    assert energy.shape == (2, 2) and list(energy[:, 0]) == [5, 10]
    for row, istep in zip(energy, (5, 10)):
        raw = (outdir / "3Ddata" / f"velocity{istep:08d}.000.dat").read_bytes()
        n = 3 * (N + 1) * nny * nnx * 8
        assert len(raw) == n + 8
        assert int.from_bytes(raw[:4], "little") == n == int.from_bytes(raw[-4:], "little")
        dump = np.frombuffer(raw[4:4 + n], dtype="<f8").reshape(3, N + 1, nny, nnx)
        np.testing.assert_array_equal(dump, phys[istep])
        e = 0.5 * float(np.sum(phys[istep] ** 2))
        assert abs(row[1] - e) <= 1e-12 * e, (istep, row[1], e)

    fc = flang()
    exe = tmp_path / "velocity_reader"
    subprocess.run([fc, "-O1", "-J", str(tmp_path), os.path.join(os.path.dirname(__file__), "fortran",
                    "velocity_reader.f90"), "-o", str(exe)], check=True, capture_output=True)
    rr = subprocess.run([str(exe), str(nnx), str(nny), str(N), str(outdir / "3Ddata" / "velocity00000010.000.dat")],
                        check=True, capture_output=True, text=True, timeout=60)
    for line, comp in zip(rr.stdout.strip().splitlines(), phys[10]):
        m, s1, s2, v123 = line.split()
        assert abs(float(s1) - comp.sum()) <= 1e-12 * np.abs(comp).sum()
        assert abs(float(s2) - (comp ** 2).sum()) <= 1e-12 * (comp ** 2).sum()
        assert float(v123) == comp[3, 2, 1]

    d = S.diagnostics(stats[10], nnx, nny, N, cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"], cfg["re_p"])
    z = np.cos(np.pi * np.arange(N + 1) / N)
    tau = d["total_shear_stress"]
    tn = tau[N]
    ua, uu, vv, ww, uv = stats[10][0], stats[10][3], stats[10][4], stats[10][5], stats[10][6]
    expect = {
        "meanu/average_velocity": ua,
        "stress/Reynolds_shear_stress": -uv / tn,
        "stress/viscous_stress": d["viscous_stress"] / tn,
        "stress/total_shear_stress": tau / tn,
        "rms/urms": np.sqrt(uu) / np.sqrt(tn),
        "rms/vrms": np.sqrt(vv) / np.sqrt(tn),
        "rms/wrms": np.sqrt(ww) / np.sqrt(tn),
    }
    for name, f in expect.items():
        prof = read_profile(outdir / f"{name}00000010.plt")
        np.testing.assert_allclose(prof[:, 0], z, rtol=0, atol=1e-15, err_msg=name)
        np.testing.assert_allclose(prof[:, 1], f, rtol=1e-12, atol=1e-12 * np.max(np.abs(f)), err_msg=name)
    assert not (outdir / "meanu" / "average_velocity00000005.plt").exists()
    h = N // 2
    yplus = cfg["re"] * ((1 - z[:h + 1]) + (1 + z[N - np.arange(h + 1)])) / 2 * np.sqrt(tn)
    uplus = (ua[:h + 1] + ua[N - np.arange(h + 1)]) / 2 / np.sqrt(tn)
    prof = read_profile(outdir / "uplus" / "uplus00000010.plt")
    np.testing.assert_allclose(prof[:, 0], yplus, rtol=1e-12)
    np.testing.assert_allclose(prof[:, 1], uplus, rtol=1e-12)
    # list-directed output may wrap a long record over several lines: read the numbers
    diag = np.array([float(t) for t in (outdir / "diagnostics.plt").read_text().split()]).reshape(1, -1)
    assert diag.shape == (1, 10) and diag[0, 0] == 10
    ut_b, ut_t = d["wall_shear_velocity"]
    want = [ut_b, ut_t, -cfg["re_p"], -tn, tau[0], d["re_tau"][0], d["re_tau"][1], d["re_centre"], d["re_bulk"]]
    np.testing.assert_allclose(diag[0, 1:], want, rtol=1e-12)
