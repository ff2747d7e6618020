"""Parity of the HIP timestep (through the C ABI) with the numpy oracle.

Floating-point tolerance: the oracle is the reference's algorithm in FP64
with explicit per-mode inverses; the GPU path is FP64 with parity-split
inverses, FFT-free wall-normal contractions and different summation orders.
Single operations agree to ~1e-13 relative; multi-step trajectories are
held to 1e-9 relative per array (the field-error budget of the mixed form).
"""
import numpy as np
import pytest

from channel_flow_amd import ic
from oracle import cf_oracle as O

from .helpers import SMALL, STATE_KEYS, ic_of, params_of, rel_err, state_errors

pytestmark = pytest.mark.gpu

ONE_OP_TOL = 1e-11
TRAJ_TOL = 1e-9


def make_channel(cfg, precision="fp64"):
    from channel_flow_amd.channel import Channel
    return Channel(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"],
                   cfg["re"], cfg["re_p"], precision=precision)


def report(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


@pytest.fixture(scope="module")
def small_case(require_gpu):
    cfg = SMALL
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    ref = O.initial_presolve_state(v, oy, au0, w0, p)
    return cfg, p, (v, oy, au0, w0), ref


def test_state_roundtrip(small_case):
    cfg, p, _, ref = small_case
    with make_channel(cfg) as ch:
        ch.set_state(ref)
        got = ch.get_state()
    for k in STATE_KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_initial_matches_oracle(small_case):
    """initial(): kinematics of every mode plus one getnonlinear (PF:1932-1946)."""
    cfg, p, (v, oy, au0, w0), ref = small_case
    with make_channel(cfg) as ch:
        ch.initial(v, oy, au0, w0)
        got = ch.get_state()
    errs = state_errors(got, ref)
    report("initial", errs)
    for k, e in errs.items():
        assert e < ONE_OP_TOL, (k, e)


def test_solve_matches_oracle(small_case):
    """One per-mode solve + mean mode (PF:242-474) from the same pre-solve state."""
    cfg, p, _, ref = small_case
    st = O.copy_state(ref)
    O.Stepper(p).solve(st)
    with make_channel(cfg) as ch:
        ch.set_state(ref)
        ch.solve()
        got = ch.get_state()
        fl = ch.fields()
    mz = p.Mz
    errs = {k: rel_err(got[k], st[k]) for k in ("vnp1", "omegaynp1", "au0", "w0")}
    for k in ("unp1", "wnp1", "omegaxnp1", "omegaznp1"):
        g = fl[k].copy()
        r = st[k].copy()
        g[:, mz // 2, :] = 0     # the Nyquist column is never solved (PF:247)
        r[:, mz // 2, :] = 0
        errs[k] = rel_err(g, r)
    report("solve", errs)
    for k, e in errs.items():
        assert e < ONE_OP_TOL, (k, e)


def test_nonlinear_matches_oracle(small_case):
    """getnonlinear (PF:1380-1612) on the fields left by one solve."""
    cfg, p, _, ref = small_case
    st = O.copy_state(ref)
    O.Stepper(p).solve(st)
    O.getnonlinear(st, p)
    with make_channel(cfg) as ch:
        ch.set_state(ref)
        ch.solve()
        ch.getnonlinear()
        got = ch.get_state()
    errs = state_errors(got, st)
    report("nonlinear", errs)
    for k, e in errs.items():
        assert e < ONE_OP_TOL, (k, e)


@pytest.mark.parametrize("nsteps", [10, 100])
def test_32x33x32_trajectory(require_gpu, nsteps):
    """Config 1 (32x33x32, Re_c = 3000, dt = 1e-3): n steps against the oracle."""
    cfg = ic.CONFIGS["re3000_32x33x32"]
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    st = O.initial_presolve_state(v, oy, au0, w0, p)
    O.Stepper(p).step(st, nsteps)
    with make_channel(cfg) as ch:
        ch.initial(v, oy, au0, w0)
        ch.step(nsteps)
        got = ch.get_state()
    errs = state_errors(got, st)
    report(f"32x33x32 {nsteps} steps", errs)
    for k, e in errs.items():
        assert e < TRAJ_TOL, (k, e)


def test_odd_degree_and_rectangular_grid(require_gpu):
    """N odd, nnx != nny, aphi/beta != 1: 3 steps against the oracle."""
    cfg = dict(nnx=24, nny=16, nnz=21, aphi=2.0, beta=2.0 / 3.0, re=180.0, re_p=1.0, dt=5.0e-4,
               u_c=18.0, rms=1.8, kmax=6.0, slope=-4.0 / 3.0, seed=11)
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    st = O.initial_presolve_state(v, oy, au0, w0, p)
    O.Stepper(p).step(st, 3)
    with make_channel(cfg) as ch:
        ch.initial(v, oy, au0, w0)
        ch.step(3)
        got = ch.get_state()
    errs = state_errors(got, st)
    report("odd N", errs)
    for k, e in errs.items():
        assert e < TRAJ_TOL, (k, e)


def test_laminar_fixed_point_gpu(require_gpu):
    """U = (re re_p/2)(1-y^2) with steady AB2 history is unchanged after 20 steps."""
    cfg = dict(SMALL)
    p = params_of(cfg)
    z = np.zeros((p.Ny, p.Mz, p.Mx), complex)
    au0 = np.zeros(p.Ny, complex)
    au0[0], au0[2] = 0.5, -0.5
    ref = O.initial_presolve_state(z, z, au0, np.zeros(p.Ny, complex), p)
    ref["clamb1nm1"] = ref["clamb1n"].copy()
    with make_channel(cfg) as ch:
        ch.set_state(ref)
        ch.step(20)
        got = ch.get_state()
    assert rel_err(got["au0"], au0) < 1e-13
    assert np.max(np.abs(got["vnp1"])) == 0.0


def test_mixed_form_against_fp64_and_oracle(require_gpu):
    """Mixed form (FP32 transforms/products, FP64 solves/state) on config 1.

    Tolerances: the Lamb vector and h_g carry FP32 rounding (~1e-7 relative);
    h_v = k^2 L_y + i alpha D L_x + i beta D L_z differentiates the Lamb
    coefficients, which multiplies the FP32 noise floor of the highest
    Chebyshev degrees by up to ~N^2 (max-norm ~1e-3 at N = 32).  The FP64
    solve damps those degrees, so v and omega_y after 10 steps agree with the
    FP64 form to ~1e-7 and the mean flow to ~1e-13."""
    cfg = ic.CONFIGS["re3000_32x33x32"]
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    ref = O.initial_presolve_state(v, oy, au0, w0, p)
    with make_channel(cfg, "mixed") as mx, make_channel(cfg, "fp64") as fp:
        mx.initial(v, oy, au0, w0)
        fp.initial(v, oy, au0, w0)
        e0 = state_errors(mx.get_state(), ref)
        report("mixed initial vs oracle", e0)
        mx.step(10)
        fp.step(10)
        gm, gf = mx.get_state(), fp.get_state()
    e1 = state_errors(gm, gf)
    report("mixed vs fp64 after 10 steps", e1)
    assert e0["rhsvn"] < 5e-3, e0["rhsvn"]
    for k in ("rhsomegayn", "clamb1n", "clamb3n"):
        assert e0[k] < 1e-5, (k, e0[k])
    for k in ("vnp1", "omegaynp1"):
        assert e1[k] < 1e-6, (k, e1[k])
    for k in ("au0", "w0"):
        assert e1[k] < 1e-10, (k, e1[k])


@pytest.mark.parametrize("nnz", [48, 64, 96, 100, 128, 192, 256])
def test_large_degree_small_box(require_gpu, nnz):
    """initial() + 3 steps vs the oracle on a 16x16 box, FP64 form.  N = 48,
    64, 96, 128, 192 and 256 run the wall-normal transforms as FFTs, one of
    each 2N split (96 ... 512 points); N = 100 (2N = 200 has no split) runs
    them as cos-table products on the matrix cores.  N = 192 is the
    256x193x256 degree, N = 256 the 512x257x512 one (its diagonalised omega_y
    solve takes two row blocks per wave)."""
    cfg = dict(nnx=16, nny=16, nnz=nnz, aphi=1.0, beta=0.5, re=7940.0, re_p=2.0 / 7940.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=21)
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    st = O.initial_presolve_state(v, oy, au0, w0, p)
    with make_channel(cfg) as ch:
        ch.initial(v, oy, au0, w0)
        e0 = state_errors(ch.get_state(), st)
        report(f"N={nnz} initial", e0)
        O.Stepper(p).step(st, 3)
        ch.step(3)
        e1 = state_errors(ch.get_state(), st)
    report(f"N={nnz} 3 steps", e1)
    for k, e in e0.items():
        assert e < ONE_OP_TOL, (k, e)
    for k, e in e1.items():
        assert e < TRAJ_TOL, (k, e)


def test_mixed_large_degree(require_gpu):
    """Mixed form at N = 192 on a 32x32 box (the wide contraction tiles and
    FP32 paths of the 256x193x256 workload): 5 steps against the FP64 form."""
    cfg = dict(nnx=32, nny=32, nnz=192, aphi=1.0, beta=0.5, re=7940.0, re_p=2.0 / 7940.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=22)
    v, oy, au0, w0 = ic_of(cfg)
    with make_channel(cfg, "mixed") as mx, make_channel(cfg, "fp64") as fp:
        mx.initial(v, oy, au0, w0)
        fp.initial(v, oy, au0, w0)
        mx.step(5)
        fp.step(5)
        gm, gf = mx.get_state(), fp.get_state()
    e = state_errors(gm, gf)
    report("mixed N=192 vs fp64, 5 steps", e)
    for k in ("vnp1", "omegaynp1"):
        assert e[k] < 1e-5, (k, e[k])
    for k in ("au0", "w0"):
        assert e[k] < 1e-9, (k, e[k])


@pytest.mark.parametrize("nnx", [128, 256, 512])
def test_long_streamwise_lines(require_gpu, nnx):
    """Padded streamwise lines of 192, 384 and 768 points (the one-wavefront-per-line
    Stockham streamwise pass of the 128-, 256- and 512-mode workloads; at 768 a
    wavefront runs two radix-8 butterflies per lane) on a short span and low
    degree: FP64 initial() + 3 steps against the oracle, and the mixed form
    against the FP64 form."""
    cfg = dict(nnx=nnx, nny=16, nnz=12, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=48.0, slope=-4.0 / 3.0, seed=23)
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    st = O.initial_presolve_state(v, oy, au0, w0, p)
    with make_channel(cfg) as ch, make_channel(cfg, "mixed") as mx:
        ch.initial(v, oy, au0, w0)
        mx.initial(v, oy, au0, w0)
        e0 = state_errors(ch.get_state(), st)
        report(f"nnx={nnx} initial", e0)
        O.Stepper(p).step(st, 3)
        ch.step(3)
        mx.step(3)
        g, gm = ch.get_state(), mx.get_state()
    e1 = state_errors(g, st)
    report(f"nnx={nnx} 3 steps", e1)
    em = state_errors(gm, g)
    report(f"nnx={nnx} mixed vs fp64", em)
    for k, e in e0.items():
        assert e < ONE_OP_TOL, (k, e)
    for k, e in e1.items():
        assert e < TRAJ_TOL, (k, e)
    for k in ("vnp1", "omegaynp1"):
        assert em[k] < 1e-5, (k, em[k])


def test_autotuner_reports_choices(require_gpu):
    """chf_create tunes the column blocks per wavefront of the five cos-table
    contractions; each has a choice of 1, 2 or 4 and positive timings (the
    parity tests above run with these tuned choices)."""
    with make_channel(SMALL) as ch:
        tiles = ch.tile_choices()
    assert set(tiles) == {"synth_padded", "analysis", "synth_points", "rhs_v", "rhs_o"}, tiles
    for name, t in tiles.items():
        assert t["choice"] in (1, 2, 4), (name, t)
        assert all(ms > 0 for ms in t["ms"]), (name, t)


def test_fft_forms_have_no_tile_choice(require_gpu):
    """At N = 192 both forms run every wall-normal transform (padded
    synthesis, analysis, point values, right-hand sides) as FFTs, which have
    no tile choice."""
    cfg = dict(nnx=32, nny=32, nnz=192, aphi=1.0, beta=0.5, re=7940.0, re_p=2.0 / 7940.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=21)
    with make_channel(cfg, "fp64") as ch:
        assert ch.tile_choices() == {}, ch.tile_choices()
    with make_channel(cfg, "mixed") as ch:
        assert ch.tile_choices() == {}, ch.tile_choices()


def test_fft_and_matrix_transforms_agree(require_gpu, monkeypatch):
    """The FP64 form's wall-normal transforms as FFTs along y (the default at
    N = 192) and as cos-table products on the FP64 matrix cores
    (CHF_Y_TRANSFORMS=matrix) are two independent evaluations of the same
    sums: after initial() and 3 steps every state array agrees to 1e-11."""
    cfg = dict(nnx=32, nny=32, nnz=192, aphi=1.0, beta=0.5, re=7940.0, re_p=2.0 / 7940.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=25)
    v, oy, au0, w0 = ic_of(cfg)
    states = {}
    for mode in ("fft", "matrix"):
        if mode == "matrix":
            monkeypatch.setenv("CHF_Y_TRANSFORMS", "matrix")
        with make_channel(cfg) as ch:
            assert ("synth_padded" in ch.tile_choices()) == (mode == "matrix"), mode
            ch.initial(v, oy, au0, w0)
            ch.step(3)
            states[mode] = ch.get_state()
        monkeypatch.delenv("CHF_Y_TRANSFORMS", raising=False)
    e = state_errors(states["fft"], states["matrix"])
    report("fft vs matrix transforms, 3 steps", e)
    for k, err in e.items():
        assert err < ONE_OP_TOL, (k, err)


def test_fft_and_matrix_transforms_agree_mixed(require_gpu, monkeypatch):
    """The mixed form's FP32 transforms as FFTs and as cos-table products on the
    FP32 matrix cores differ by FP32 rounding only: after initial() and 5 steps
    at N = 192 they agree as closely as the mixed form agrees with the FP64
    form (the bounds of test_mixed_large_degree)."""
    cfg = dict(nnx=32, nny=32, nnz=192, aphi=1.0, beta=0.5, re=7940.0, re_p=2.0 / 7940.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=26)
    v, oy, au0, w0 = ic_of(cfg)
    states = {}
    for mode in ("fft", "matrix"):
        if mode == "matrix":
            monkeypatch.setenv("CHF_Y_TRANSFORMS", "matrix")
        with make_channel(cfg, "mixed") as ch:
            assert ("synth_padded" in ch.tile_choices()) == (mode == "matrix"), mode
            ch.initial(v, oy, au0, w0)
            ch.step(5)
            states[mode] = ch.get_state()
        monkeypatch.delenv("CHF_Y_TRANSFORMS", raising=False)
    e = state_errors(states["fft"], states["matrix"])
    report("mixed fft vs matrix transforms, 5 steps", e)
    for k in ("vnp1", "omegaynp1"):
        assert e[k] < 1e-5, (k, e[k])
    for k in ("au0", "w0"):
        assert e[k] < 1e-9, (k, e[k])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tile_choices_bit_identical(require_gpu, precision, monkeypatch):
    """Every column-block choice of the tuned contractions (forced through
    chf_set_tile_choice) gives bit-identical states after initial() and two
    steps, at N = 192 (row blocks that admit 1, 2 and 4 column blocks, the
    centre rows summed outside the matrix cores), with the wall-normal
    transforms as cos-table products (CHF_Y_TRANSFORMS=matrix)."""
    monkeypatch.setenv("CHF_Y_TRANSFORMS", "matrix")
    cfg = dict(nnx=32, nny=32, nnz=192, aphi=1.0, beta=0.5, re=7940.0, re_p=2.0 / 7940.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=24)
    v, oy, au0, w0 = ic_of(cfg)
    states = {}
    for cb in (1, 2, 4):
        with make_channel(cfg, precision) as ch:
            assert len(ch.tile_choices()) == 5, ch.tile_choices()
            for name in ch.tile_choices():
                ch.set_tile_choice(name, cb)
            ch.initial(v, oy, au0, w0)
            ch.step(2)
            states[cb] = ch.get_state()
    for cb in (2, 4):
        for k in STATE_KEYS:
            np.testing.assert_array_equal(states[cb][k], states[1][k], err_msg=f"{precision} cb={cb} {k}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_timing_does_not_change_results(require_gpu, precision):
    """With per-kernel timing on, the launches of a step run one after another
    on the context stream (without it the omega_y path overlaps the v path on
    a second stream): both give bit-identical states, and every kernel class
    of the step is timed."""
    cfg = dict(nnx=32, nny=32, nnz=48, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0,
               dt=5.0e-4, u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=26)
    v, oy, au0, w0 = ic_of(cfg)
    with make_channel(cfg, precision) as a, make_channel(cfg, precision) as b:
        a.initial(v, oy, au0, w0)
        b.initial(v, oy, au0, w0)
        a.step(3)
        b.set_timing(True)
        b.step(3)
        kt = b.kernel_times()
        b.set_timing(False)
        sa, sb = a.get_state(), b.get_state()
    for k in STATE_KEYS:
        np.testing.assert_array_equal(sb[k], sa[k], err_msg=f"{precision} {k}")
    for name in ("rhs_v", "rhs_o", "solve", "wsolve", "kinematics", "synth_padded", "analysis"):
        assert kt[name][1] == 3, (name, kt[name])


@pytest.mark.parametrize("nny", [256, 512])
def test_long_span_lines(require_gpu, nny):
    """Padded span lines of 384 and 768 points (the register span FFTs of the
    256- and 512-mode workloads, split 16 x 24 and 24 x 32) on a short
    streamwise extent and low degree: FP64 initial() + 3 steps against the
    oracle, and the mixed form against the FP64 form."""
    cfg = dict(nnx=32, nny=nny, nnz=12, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0,
               dt=2.5e-4, u_c=1.0, rms=0.1, kmax=48.0, slope=-4.0 / 3.0, seed=27)
    p = params_of(cfg)
    v, oy, au0, w0 = ic_of(cfg)
    st = O.initial_presolve_state(v, oy, au0, w0, p)
    with make_channel(cfg) as ch, make_channel(cfg, "mixed") as mx:
        ch.initial(v, oy, au0, w0)
        mx.initial(v, oy, au0, w0)
        e0 = state_errors(ch.get_state(), st)
        report(f"nny={nny} initial", e0)
        O.Stepper(p).step(st, 3)
        ch.step(3)
        mx.step(3)
        g, gm = ch.get_state(), mx.get_state()
    e1 = state_errors(g, st)
    report(f"nny={nny} 3 steps", e1)
    em = state_errors(gm, g)
    report(f"nny={nny} mixed vs fp64", em)
    for k, e in e0.items():
        assert e < ONE_OP_TOL, (k, e)
    for k, e in e1.items():
        assert e < TRAJ_TOL, (k, e)
    for k in ("vnp1", "omegaynp1"):
        assert em[k] < 1e-5, (k, em[k])
