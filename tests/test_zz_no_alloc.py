"""chf_physical_noalloc: the unpadded physical transform into caller-owned
device memory, with its work plane in the library's fixed per-GPU array
instead of memory it allocates.

Its output must equal chf_physical's (the same kernels on the same inputs:
bit for bit) on two calls in a row, the second on a smaller channel (so the
array still holds the first call's larger contents), and a grid whose work
plane exceeds the array must be rejected with a message.  Device buffers are
allocated here through the HIP runtime (the test may allocate; the entry point
does not).
"""
import ctypes

import numpy as np
import pytest

from channel_flow_amd import ic

from .helpers import ic_of

pytestmark = pytest.mark.gpu


def hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


def make_channel(cfg):
    from channel_flow_amd.channel import Channel
    return Channel(cfg["nnx"], cfg["nny"], cfg["nnz"], cfg["aphi"], cfg["beta"], cfg["dt"], cfg["re"],
                   cfg["re_p"])


def noalloc_field(lib, ch, field):
    shape = (ch.Ny, ch.Mz, 2 * ch.Mx)
    nbytes = int(np.prod(shape)) * 8
    dev = ctypes.c_void_p()
    assert lib.hipMalloc(ctypes.byref(dev), nbytes) == 0
    try:
        ch.physical_noalloc(field, dev.value)
        ch.sync()
        out = np.empty(shape, dtype=np.float64)
        assert lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, nbytes, 2) == 0   # device to host
    finally:
        lib.hipFree(dev)
    return out


def test_noalloc_physical_equals_chf_physical(require_gpu):
    lib = hip()
    big = ic.CONFIGS["retau180_128x129x128"]
    small = dict(nnx=32, nny=32, nnz=24, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0, dt=1.0e-3,
                 u_c=1.0, rms=0.1, kmax=8.0, slope=-4.0 / 3.0, seed=51)
    for cfg, fields in ((big, ("u", "omega_z")), (small, ("v", "w"))):
        with make_channel(cfg) as ch:
            ch.initial(*ic_of(cfg))
            ch.step(2)
            for f in fields:
                got = noalloc_field(lib, ch, f)
                ref = ch.physical(f)
                print(f"{cfg['nnx']}x{cfg['nnz'] + 1}x{cfg['nny']} {f}: max |ref| {np.max(np.abs(ref)):.3e}",
                      flush=True)
                np.testing.assert_array_equal(got, ref, err_msg=f"{cfg['nnx']} {f}")


def test_noalloc_rejects_a_grid_beyond_its_scratch(require_gpu):
    """512 x 65 x 512: the work plane is 65 * 512 * 256 * 16 B = 136 MB, above
    the 128 MiB array."""
    from channel_flow_amd import runtime as rt
    cfg = dict(nnx=512, nny=512, nnz=64, aphi=1.0, beta=0.5, re=3000.0, re_p=2.0 / 3000.0, dt=1.0e-3)
    lib = hip()
    with make_channel(cfg) as ch:
        nbytes = ch.Ny * ch.Mz * 2 * ch.Mx * 8
        dev = ctypes.c_void_p()
        assert lib.hipMalloc(ctypes.byref(dev), nbytes) == 0
        try:
            with pytest.raises(rt.ChfError, match="scratch") as ei:
                ch.physical_noalloc("u", dev.value)
            assert ei.value.code < 0
        finally:
            lib.hipFree(dev)
