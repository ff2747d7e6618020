"""bench.py's algorithmic HBM bytes per launch (kernel_work), which the
roofline fractions divide by the launch time, against the HBM traffic that the
FETCH_SIZE and WRITE_SIZE counters measured for the same kernels at
256x193x256 (the committed counter summaries under profiles/, scaled by a
device copy of known size).

Only the kernel classes that run as one full-size launch per step are
compared: the transform kernels also run as two field groups on two streams,
so their per-dispatch averages mix two launch sizes.  A class whose counted
bytes fell outside [0.97, 1.15] x its measured traffic would mean a field
counted at the wrong precision, or counted twice or not at all.  The solve
sits highest (measured / algorithmic = 1.11): the inverses are counted once
per operator class, 1.55 GB, but a class with more than two groups runs as
two blocks that each stream them, 1.73 GB."""
import json
import os

import pytest

import bench
from channel_flow_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIG = "retau395_256x193x256"

# kernel class -> symbol prefix in the counter summary, per form
CLASSES = {
    "mixed": {"solve": "void k_solve<false>", "xpass": "void k_xpass_st<",
              "kinematics": "void k_kinematics<true>", "hassemble": "void k_hassemble<true>",
              "rhs_v": "void k_yrhs<16, 24, 2,", "rhs_o": "void k_yrhs<16, 24, 1,",
              "wsolve": "void k_wsolve<"},
    "fp64": {"solve": "void k_solve<false>", "xpass": "void k_xpass_st<",
             "kinematics": "void k_kinematics<false>", "hassemble": "void k_hassemble<false>",
             "rhs_v": "void k_yrhs<16, 24, 2,", "rhs_o": "void k_yrhs<16, 24, 1,",
             "wsolve": "void k_wsolve<"},
}


def measured(form):
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{CONFIG}_{form}.json")
    with open(path) as f:
        return json.load(f)["kernels"]


@pytest.mark.parametrize("form", ["mixed", "fp64"])
def test_algorithmic_bytes_match_measured_traffic(form):
    d = bench.dims(dict(ic.CONFIGS[CONFIG]))
    work = bench.kernel_work(d, form, yfft=True)
    data = measured(form)
    for cls, prefix in CLASSES[form].items():
        hits = [v for k, v in data.items() if k.startswith(prefix)]
        assert len(hits) == 1, (form, cls, prefix, list(data))
        traffic = hits[0]["fetch_bytes"] + hits[0]["write_bytes"]
        ratio = traffic / work[cls][0]
        assert 0.97 <= ratio <= 1.15, (form, cls, round(ratio, 4))


def test_mixed_form_counts_its_fp32_fields():
    """The mixed form's FP32 storage is where the two forms' counts differ:
    kinematics 6 against 8 spectral fields, h assembly 3.5 against 5 (plus
    the mean-mode block), the x-pass half the FP64 bytes; the solve and the
    right-hand sides are FP64 in both."""
    d = bench.dims(dict(ic.CONFIGS[CONFIG]))
    S = d["Ny"] * d["ncol"] * 16
    mx, fp = bench.kernel_work(d, "mixed", yfft=True), bench.kernel_work(d, "fp64", yfft=True)
    assert mx["kinematics"][0] == 6 * S and fp["kinematics"][0] == 8 * S
    assert mx["hassemble"][0] - fp["hassemble"][0] == -1.5 * S
    assert 2 * mx["xpass"][0] == fp["xpass"][0]
    for cls in ("solve", "rhs_v", "rhs_o", "wsolve"):
        assert mx[cls] == fp[cls], cls
