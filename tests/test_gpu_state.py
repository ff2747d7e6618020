"""State handling at the C ABI on the GPU: states that carry the reference's
quirks, resume from a state record, blow-up reporting, and the stored
omega_y inverses.

* A reference restart file (PF:1913-1917) can hold nonzero v, omega_y at the
  (0,0) mode and on the spanwise Nyquist column, and a kx = 0 plane that is not
  Hermitian.  The reference never solves those modes (PF:247, 2365): they stay
  as loaded, the products see the Nyquist column zeroed (PF:730) and the kx = 0
  plane through Re() of its inverse transform (PF:1116-1118).  The GPU must
  carry the same state forward as the oracle does.
* A run resumed from chf_get_state in a new context continues bit for bit as
  the uninterrupted run (the record holds the full AB2 history, PF:2476-2479).
* A NaN or an infinity in the state is reported as CHF_NONFINITE.
"""
import numpy as np
import pytest

from oracle import cf_oracle as O

from .helpers import SMALL, STATE_KEYS, ic_of, params_of, solved_state_invariants, state_errors

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-9


def make_channel(cfg, precision="fp64"):
    from channel_flow_amd.channel import Channel
    return Channel(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"],
                   cfg["re_p"], precision=precision)


def report(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def quirky_state(cfg, seed=99):
    """A pre-solve state whose v, omega_y carry nonzero (0,0) modes, a nonzero
    spanwise Nyquist column and a non-Hermitian kx = 0 plane, as a restart
    file of the reference may."""
    p = params_of(cfg)
    st = O.initial_presolve_state(*ic_of(cfg), p)
    rng = np.random.default_rng(seed)
    h = p.Mz // 2
    for key in ("vnp1", "omegaynp1"):
        f = st[key]
        scale = float(np.max(np.abs(f)))

        def noise(shape):
            return scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))

        f[:, 0, 0] = noise(p.Ny)
        f[:, h, :] = noise((p.Ny, p.Mx))
        f[:, 1:h, 0] += 0.5 * noise((p.Ny, h - 1))           # +kz of kx = 0 only: not Hermitian
    return p, st


def frozen_entries(st, mz):
    return {k: (st[k][:, 0, 0].copy(), st[k][:, mz // 2, :].copy()) for k in ("vnp1", "omegaynp1")}


def test_quirky_state_matches_oracle(require_gpu):
    """FP64: 3 steps from the quirky state against the oracle at 1e-9; the
    (0,0) modes and the Nyquist column of v, omega_y come out bit for bit as
    they went in.  Mixed: 1 step, the frozen entries equally untouched."""
    cfg = dict(SMALL, nnx=32)          # the mixed form needs nnx % 32 == 0
    p, st0 = quirky_state(cfg)
    ref = O.copy_state(st0)
    O.Stepper(p).step(ref, 3)
    frozen = frozen_entries(st0, p.Mz)
    for k, (m00, nyq) in frozen_entries(ref, p.Mz).items():      # the oracle freezes them too
        assert np.array_equal(m00, frozen[k][0]) and np.array_equal(nyq, frozen[k][1]), k
    with make_channel(cfg) as ch:
        ch.set_state(st0)
        ch.step(3)
        got = ch.get_state()
        fl = ch.fields()
    errs = state_errors(got, ref)
    report("quirky state, 3 steps", errs)
    for k, e in errs.items():
        assert e < TRAJ_TOL, (k, e)
    for k, (m00, nyq) in frozen_entries(got, p.Mz).items():
        assert np.array_equal(m00, frozen[k][0]), f"{k}(0,0) changed"
        assert np.array_equal(nyq, frozen[k][1]), f"{k} Nyquist column changed"
    inv = solved_state_invariants(p, got, fl, before=st0)
    report("quirky state invariants", inv)
    assert inv["divergence"] < 1e-12 and inv["omega_y_curl"] < 1e-12 and inv["mean_curl"] < 1e-12, inv
    assert inv["nyquist"] == 0.0 and inv["frozen_00"] == 0.0, inv

    ref1 = O.copy_state(st0)
    O.Stepper(p).step(ref1, 1)
    with make_channel(cfg, "mixed") as mx:
        mx.set_state(st0)
        mx.step(1)
        gm = mx.get_state()
    em = state_errors(gm, ref1)
    report("quirky state, mixed, 1 step", em)
    for k in ("vnp1", "omegaynp1", "au0", "w0"):
        assert em[k] < 1e-6, (k, em[k])
    for k, (m00, nyq) in frozen_entries(gm, p.Mz).items():
        assert np.array_equal(m00, frozen[k][0]) and np.array_equal(nyq, frozen[k][1]), k


def test_set_state_derives_the_fields(require_gpu):
    """After chf_set_state the velocity and vorticity (chf_get_fields) are the
    kinematics of the loaded v, omega_y, U, W, not those of an earlier state."""
    cfg = dict(SMALL)
    p = params_of(cfg)
    st = O.initial_presolve_state(*ic_of(cfg), p)
    ref = O.copy_state(st)
    O.kinematics_all(ref, p)
    with make_channel(cfg) as ch:
        ch.initial(*ic_of(dict(cfg, seed=cfg["seed"] + 1)))     # some other state first
        ch.set_state(st)
        fl = ch.fields()
    mz = p.Mz
    for k in ("unp1", "wnp1", "omegaxnp1", "omegaznp1"):
        g, r = fl[k].copy(), ref[k].copy()
        g[:, mz // 2, :] = 0
        r[:, mz // 2, :] = 0
        e = float(np.max(np.abs(g - r))) / float(np.max(np.abs(r)))
        assert e < 1e-13, (k, e)


def test_mixed_fields_read_after_steps_are_the_state_kinematics(require_gpu, monkeypatch):
    """The mixed form with the FFT transforms keeps u, w, omega_x, omega_z in
    FP32 between its steps and forms the FP64 fields only when they are read.
    After steps (and again after more steps and a second read) chf_get_fields,
    chf_physical and chf_statistics equal, bit for bit, what a new context
    derives from the same state with chf_set_state, and the fields are the
    oracle's kinematics of that state (1e-13).  The reads rewrite the next
    solve's brackets with the values they hold: the run continues bit for bit
    as one without reads."""
    cfg = dict(nnx=32, nny=32, nnz=32, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0, dt=1.0e-3,
               u_c=1.0, rms=0.1, kmax=8.0, slope=-4.0 / 3.0, seed=47)
    p = params_of(cfg)
    monkeypatch.delenv("CHF_Y_TRANSFORMS", raising=False)

    def reads(ch):
        return ch.fields(), {f: ch.physical(f) for f in ("u", "omega_z")}, ch.plane_statistics()

    with make_channel(cfg, "mixed") as a:
        assert "synth_padded" not in a.tile_choices(), "the FFT transforms are expected at this size"
        a.initial(*ic_of(cfg))
        runs = []
        for n in (5, 3):
            a.step(n)
            st = a.get_state()
            first = reads(a)
            again = reads(a)
            runs.append((st, first, again))
    with make_channel(cfg, "mixed") as s:
        s.initial(*ic_of(cfg))
        s.step(8)
        straight = s.get_state()
    for k in STATE_KEYS:
        np.testing.assert_array_equal(runs[-1][0][k], straight[k], err_msg=f"reads changed the run: {k}")
    mz = p.Mz
    for st, (fl, ph, ps), (fl2, ph2, ps2) in runs:
        with make_channel(cfg, "mixed") as b:
            b.set_state(st)
            rf, rph, rps = reads(b)
        for k in fl:
            np.testing.assert_array_equal(fl[k], rf[k], err_msg=k)
            np.testing.assert_array_equal(fl2[k], rf[k], err_msg=k)
        for f in ph:
            np.testing.assert_array_equal(ph[f], rph[f], err_msg=f)
            np.testing.assert_array_equal(ph2[f], rph[f], err_msg=f)
        np.testing.assert_array_equal(ps, rps)
        np.testing.assert_array_equal(ps2, rps)
        ref = {k: np.array(v, copy=True) for k, v in st.items()}
        for k in fl:                       # the record holds no fields; every mode is set below
            ref[k] = np.zeros_like(st["vnp1"])
        O.kinematics_all(ref, p)
        for k in ("unp1", "wnp1", "omegaxnp1", "omegaznp1"):
            g, r = fl[k].copy(), ref[k].copy()
            g[:, mz // 2, :] = 0
            r[:, mz // 2, :] = 0
            e = float(np.max(np.abs(g - r))) / float(np.max(np.abs(r)))
            assert e < 1e-13, (k, e)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nnz", [24, 32])
def test_resume_is_bit_identical(require_gpu, precision, nnz):
    """20 straight steps against 10 steps, chf_get_state, a new context,
    chf_set_state and 10 more steps: every state array and field identical.
    N = 24 runs the wall-normal transforms as cos-table products, N = 32 as
    FFTs (the resumed run's first right-hand sides then read h^n's point
    values from the record, where the straight run forms them)."""
    cfg = dict(nnx=32, nny=32, nnz=nnz, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0, dt=1.0e-3,
               u_c=1.0, rms=0.1, kmax=8.0, slope=-4.0 / 3.0, seed=41)
    arrays = ic_of(cfg)
    with make_channel(cfg, precision) as a:
        a.initial(*arrays)
        a.step(20)
        straight, fs = a.get_state(), a.fields()
    with make_channel(cfg, precision) as b:
        b.initial(*arrays)
        b.step(10)
        mid = b.get_state()
    with make_channel(cfg, precision) as c:
        c.set_state(mid)
        c.step(10)
        resumed, fr = c.get_state(), c.fields()
    for k in STATE_KEYS:
        np.testing.assert_array_equal(resumed[k], straight[k], err_msg=f"{precision} {k}")
    for k in fs:
        np.testing.assert_array_equal(fr[k], fs[k], err_msg=f"{precision} {k}")


@pytest.mark.parametrize("nnz", [32, 96])
def test_h_point_values_whatever_the_call_order(require_gpu, monkeypatch, nnz):
    """With the FFT transforms the next solve's right-hand-side kernels form
    h^n's point values, or chf_get_state or a second nonlinear pass forms them
    first.  (1) A run split by state reads and separate solve / nonlinear calls
    is bit-identical to chf_step straight through.  (2) Two nonlinear passes in
    a row (the AB2 history then holds the same h twice) leave the state that
    the cos-table transforms, which form the point values in the nonlinear
    pass, leave (FP64, 1e-11).  N = 32 and 96: FFT splits 8 x 8 and 12 x 16."""
    cfg = dict(nnx=32, nny=32, nnz=nnz, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0, dt=1.0e-3,
               u_c=1.0, rms=0.1, kmax=8.0, slope=-4.0 / 3.0, seed=43)
    arrays = ic_of(cfg)
    monkeypatch.delenv("CHF_Y_TRANSFORMS", raising=False)
    with make_channel(cfg) as a:
        assert a.tile_choices() == {}, "the FFT transforms are expected at this size"
        a.initial(*arrays)
        a.step(6)
        straight, fs = a.get_state(), a.fields()
    with make_channel(cfg) as b:
        b.initial(*arrays)
        b.get_state()                                   # before the first solve
        b.step(2)
        b.solve()
        b.get_state()                                   # between a solve and its nonlinear pass
        b.getnonlinear()
        b.step(1)
        b.solve()
        b.getnonlinear()
        b.get_state()
        b.step(1)
        split, fsp = b.get_state(), b.fields()
    for k in STATE_KEYS:
        np.testing.assert_array_equal(split[k], straight[k], err_msg=k)
    for k in fs:
        np.testing.assert_array_equal(fsp[k], fs[k], err_msg=k)

    def twice(matrix):
        if matrix:
            monkeypatch.setenv("CHF_Y_TRANSFORMS", "matrix")
        else:
            monkeypatch.delenv("CHF_Y_TRANSFORMS", raising=False)
        with make_channel(cfg) as c:
            c.initial(*arrays)
            c.step(2)
            c.getnonlinear()
            c.step(1)
            return c.get_state()

    errs = state_errors(twice(False), twice(True))
    report("two nonlinear passes, FFT vs cos-table transforms", errs)
    for k, e in errs.items():
        assert e < 1e-11, (k, e)


def test_nonfinite_state_is_reported(require_gpu):
    """A NaN loaded into v is reported by chf_set_state; a NaN in h^n (which
    the loaded fields do not show) by the chf_step whose solve spreads it; a
    finite state loaded afterwards steps normally again."""
    from channel_flow_amd import runtime as rt
    cfg = dict(SMALL)
    p = params_of(cfg)
    st = O.initial_presolve_state(*ic_of(cfg), p)
    with make_channel(cfg) as ch:
        bad = O.copy_state(st)
        bad["vnp1"][3, 2, 1] = np.nan
        with pytest.raises(rt.ChfNonFinite):
            ch.set_state(bad)
        bad = O.copy_state(st)
        bad["rhsvn"][5, 3, 2] = np.inf
        ch.set_state(bad)
        with pytest.raises(rt.ChfNonFinite) as ei:
            ch.step(1)
        assert ei.value.code == rt.NONFINITE
        ch.set_state(st)
        ch.step(2)
        assert np.all(np.isfinite(ch.get_state()["vnp1"]))


def test_blow_up_is_reported(require_gpu):
    """A time step far beyond the explicit advection limit (kx_max U dt ~ 14)
    makes AB2 amplify the highest streamwise modes every step; the state
    overflows and chf_step reports CHF_NONFINITE within a bounded number of
    steps instead of returning 0."""
    from channel_flow_amd import runtime as rt
    cfg = dict(SMALL, dt=2.0)
    steps = 0
    with make_channel(cfg) as ch:
        ch.initial(*ic_of(cfg))
        with pytest.raises(rt.ChfNonFinite):
            while steps < 5000:
                ch.step(50)
                steps += 50
    print(f"blow-up reported after at most {steps + 50} steps", flush=True)
    assert steps < 5000


@pytest.mark.parametrize("nnz", [32, 192])
def test_stored_omega_inverses_match_oracle(require_gpu, nnz, monkeypatch):
    """The stored omega_y half-system inverses (CHF_OMEGA_INVERSES=stored; the
    path taken when the diagonalised operator does not verify) against the
    oracle: initial() + 3 steps at 1e-9, with no diagonalised omega_y solve
    launched."""
    monkeypatch.setenv("CHF_OMEGA_INVERSES", "stored")
    cfg = dict(nnx=16, nny=16, nnz=nnz, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0, dt=1.0e-3,
               u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=43)
    p = params_of(cfg)
    arrays = ic_of(cfg)
    st = O.initial_presolve_state(*arrays, p)
    O.Stepper(p).step(st, 3)
    with make_channel(cfg) as ch:
        ch.initial(*arrays)
        ch.set_timing(True)
        ch.step(3)
        kt = ch.kernel_times()
        ch.set_timing(False)
        got = ch.get_state()
    assert kt["wsolve"][1] == 0 and kt["solve"][1] == 3, kt
    errs = state_errors(got, st)
    report(f"stored omega_y inverses N={nnz}, 3 steps", errs)
    for k, e in errs.items():
        assert e < TRAJ_TOL, (k, e)
