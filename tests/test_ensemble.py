"""Batched ensemble (chf_ensemble_*): members of different sizes, Reynolds
numbers and precisions advanced by hipGraph launches must reproduce, bit for
bit, the same channels advanced one by one (chf_step), including across the
AB2 history phase (odd and even step counts per call)."""
import numpy as np
import pytest

from channel_flow_amd import ic

from .helpers import STATE_KEYS

MEMBERS = [
    dict(nnx=16, nny=16, nnz=16, aphi=1.0, beta=0.5, re=500.0, re_p=2.0 / 500.0, dt=1.0e-3, u_c=1.0,
         rms=5.0e-2, kmax=4.0, slope=0.0, seed=31, precision="fp64"),
    dict(nnx=32, nny=32, nnz=24, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0, dt=1.0e-3, u_c=1.0,
         rms=0.1, kmax=8.0, slope=-4.0 / 3.0, seed=32, precision="mixed"),
    dict(nnx=24, nny=16, nnz=21, aphi=2.0, beta=2.0 / 3.0, re=2000.0, re_p=2.0 / 2000.0, dt=5.0e-4,
         u_c=1.0, rms=0.1, kmax=6.0, slope=-4.0 / 3.0, seed=33, precision="fp64"),
]


def seeded(m):
    return ic.seeded_ic(m["nnx"], m["nny"], m["nnz"], m["aphi"], m["beta"], m["u_c"], m["rms"], m["kmax"],
                        m["slope"], m["seed"])


@pytest.mark.gpu
def test_ensemble_equals_independent_channels(require_gpu):
    from channel_flow_amd.channel import Channel
    from channel_flow_amd.ensemble import Ensemble
    with Ensemble(MEMBERS) as ens:
        for i, m in enumerate(MEMBERS):
            ens.member(i).initial(*seeded(m))
        ens.step(2)
        ens.step(3)
        got = [ens.member(i).get_state() for i in range(len(MEMBERS))]
        fields = [ens.member(i).fields() for i in range(len(MEMBERS))]
    for i, m in enumerate(MEMBERS):
        with Channel(m["nnx"], m["nny"], m["nnz"], m["aphi"], m["beta"], m["dt"], m["re"], m["re_p"],
                     precision=m["precision"]) as ch:
            ch.initial(*seeded(m))
            ch.step(5)
            ref = ch.get_state()
            reff = ch.fields()
        for k in STATE_KEYS:
            assert np.array_equal(got[i][k], ref[k]), (i, k)
        for k in reff:
            assert np.array_equal(fields[i][k], reff[k]), (i, k)
        print(f"member {i} ({m['nnx']}x{m['nnz'] + 1}x{m['nny']}, {m['precision']}): identical after 5 steps",
              flush=True)


def test_ensemble_workload_definition():
    ms = ic.ensemble_members()
    assert len(ms) == 24
    assert {(m["nnx"], m["nny"], m["nnz"]) for m in ms} <= set(ic.ENSEMBLE_SIZES)
    assert all(2000.0 <= m["re"] <= 6000.0 for m in ms)


def _member_channel(m):
    from channel_flow_amd.channel import Channel
    return Channel(m["nnx"], m["nny"], m["nnz"], m["aphi"], m["beta"], m["dt"], m["re"], m["re_p"],
                   precision=m["precision"])


@pytest.mark.gpu
def test_ensemble_workload_24_members(require_gpu):
    """The batched-ensemble workload itself: the 24 members of
    ic.ensemble_members() (64x65x64, 96x97x96 and 128x97x128, Re_c in
    [2000, 6000], mixed form) advanced 3 steps by hipGraph launches.  Each
    member is bit-identical to the same channel stepped alone, and each solved
    state has the invariants of a solved state (continuity, omega_y identity,
    mean-mode curl, frozen Nyquist column and (0,0) modes)."""
    from oracle import cf_oracle as O
    from channel_flow_amd.ensemble import Ensemble
    from .helpers import solved_state_invariants
    members = ic.ensemble_members()
    with Ensemble(members) as ens:
        for i, m in enumerate(members):
            ens.member(i).initial(*seeded(m))
        ens.step(1)
        ens.step(2)
        got = [ens.member(i).get_state() for i in range(len(members))]
        fields = [ens.member(i).fields() for i in range(len(members))]
    for i, m in enumerate(members):
        with _member_channel(m) as ch:
            ch.initial(*seeded(m))
            ch.step(3)
            ref = ch.get_state()
            reff = ch.fields()
        for k in STATE_KEYS:
            assert np.array_equal(got[i][k], ref[k]), (i, k)
        for k in reff:
            assert np.array_equal(fields[i][k], reff[k]), (i, k)
        p = O.Params(m["nnx"], m["nny"], m["nnz"], m["aphi"], m["beta"], m["dt"], m["re"], m["re_p"])
        v0, oy0, _, _ = seeded(m)
        inv = solved_state_invariants(p, got[i], fields[i], before={"vnp1": v0, "omegaynp1": oy0})
        assert inv["divergence"] < 1e-12 and inv["omega_y_curl"] < 1e-12 and inv["mean_curl"] < 1e-12, (i, inv)
        assert inv["hermitian_kx0"] < 1e-5 and inv["nyquist"] == 0.0 and inv["frozen_00"] == 0.0, (i, inv)
        print(f"member {i} ({m['nnx']}x{m['nnz'] + 1}x{m['nny']}, re={m['re']:.0f}): identical to a lone "
              f"channel, divergence {inv['divergence']:.1e}", flush=True)


@pytest.mark.gpu
def test_ensemble_workload_sizes_fp64_against_oracle(require_gpu):
    """One member of each workload size in the FP64 form inside one ensemble;
    the 64x65x64 member against the oracle over initial() + 3 steps at 1e-9
    (the larger two are held to the lone-channel identity above)."""
    from oracle import cf_oracle as O
    from channel_flow_amd.ensemble import Ensemble
    from .helpers import state_errors
    members = []
    for size in ic.ENSEMBLE_SIZES:
        m = next(m for m in ic.ensemble_members(precision="fp64") if (m["nnx"], m["nny"], m["nnz"]) == size)
        members.append(m)
    with Ensemble(members) as ens:
        for i, m in enumerate(members):
            ens.member(i).initial(*seeded(m))
        ens.step(3)
        got = ens.member(0).get_state()
    m = members[0]
    assert (m["nnx"], m["nnz"]) == (64, 64)
    p = O.Params(m["nnx"], m["nny"], m["nnz"], m["aphi"], m["beta"], m["dt"], m["re"], m["re_p"])
    st = O.initial_presolve_state(*seeded(m), p)
    O.Stepper(p).step(st, 3)
    errs = state_errors(got, st)
    print("64x65x64 ensemble member vs oracle, 3 steps:", " ".join(f"{k}={v:.1e}" for k, v in errs.items()),
          flush=True)
    for k, e in errs.items():
        assert e < 1e-9, (k, e)
