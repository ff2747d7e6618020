"""CPU checks of the C ABI: the library loads and exports every declared entry point."""
import ctypes
import os
import re

from channel_flow_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names.update(re.findall(r"\b(chf_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = runtime.load()
    declared = declared_functions()
    assert declared >= set(runtime.EXPORTS)
    for name in sorted(declared):
        assert hasattr(lib, name), name


def test_create_rejects_invalid_parameters_without_a_device():
    lib = runtime.load()
    bad = runtime.ChfParams(30, 32, 32, 1.0, 0.5, 1e-3, 100.0, 0.0, 0)   # nnx % 4 != 0
    ctx = ctypes.c_void_p()
    rc = lib.chf_create(ctypes.byref(bad), ctypes.byref(ctx))
    assert rc < 0
    assert b"nnx" in lib.chf_last_error(None)
    assert not ctx.value
