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
