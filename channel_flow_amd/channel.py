"""Host-side interface of the GPU timestep, named after the reference call sites.

    ch = Channel(nnx=32, nny=32, nnz=32, aphi=1.0, beta=0.5, dt=1e-3,
                 re=3000.0, re_p=2/3000)
    ch.initial(v, omegay, au0, w0)       # initial(isign=-1), PF:1932-1946
    ch.step(100)                         # 100 x [solve ; getnonlinear], PF:220-532
    st = ch.get_state()                  # output(), PF:2460-2504

State dictionaries use the reference's array names (PC:28-38) and memory
layout: complex128[Ny][Mz][Mx] fields (== real*8 (0:nx,0:ny,0:nz) with
interleaved Re/Im) and complex128[Ny] mean-mode arrays (== real*8 (2,0:nz)).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import runtime as rt


class Channel:
    def __init__(self, nnx, nny, nnz, aphi, beta, dt, re, re_p, precision="fp64"):
        self._lib = rt.load()
        prec = {"fp64": 0, "mixed": 1}[precision]
        self.params = rt.ChfParams(nnx, nny, nnz, aphi, beta, dt, re, re_p, prec)
        self._ctx = ctypes.c_void_p()
        rc = self._lib.chf_create(ctypes.byref(self.params), ctypes.byref(self._ctx))
        if rc != 0:
            msg = self._lib.chf_last_error(None).decode()
            raise rt.error(f"chf_create failed ({rc}): {msg}", rc)
        self._owned = True
        self._init_dims(precision)

    def _init_dims(self, precision):
        mx, mz, ny = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._lib.chf_dims(self._ctx, ctypes.byref(mx), ctypes.byref(mz), ctypes.byref(ny))
        self.Mx, self.Mz, self.Ny = mx.value, mz.value, ny.value
        self.precision = precision

    @classmethod
    def _borrowed(cls, lib, ctx, params, precision):
        """A view of a context owned elsewhere (an ensemble member)."""
        self = cls.__new__(cls)
        self._lib = lib
        self._ctx = ctx
        self.params = params
        self._owned = False
        self._init_dims(precision)
        return self

    # -------------------------------------------------------------- plumbing
    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.chf_last_error(self._ctx).decode()
            raise rt.error(f"{what} failed ({rc}): {msg}", rc)

    def close(self):
        if self._ctx and self._owned:
            self._lib.chf_destroy(self._ctx)
        self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def field_shape(self):
        return (self.Ny, self.Mz, self.Mx)

    def _as_field(self, a, name):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != self.field_shape:
            raise ValueError(f"{name}: expected shape {self.field_shape}, got {a.shape}")
        return a

    def _as_mean(self, a, name):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != (self.Ny,):
            raise ValueError(f"{name}: expected shape {(self.Ny,)}, got {a.shape}")
        return a

    @staticmethod
    def _dp(a):
        return a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    # ------------------------------------------------------------ interface
    def set_state(self, st: dict):
        """Load a pre-solve state (the restart record, PF:1913-1917)."""
        arrs = {}
        for n in rt.STATE_FIELDS:
            arrs[n] = (self._as_mean if n in rt.MEAN_FIELDS else self._as_field)(st[n], n)
        cs = rt.ChfState(**{n: self._dp(arrs[n]) for n in rt.STATE_FIELDS})
        self._check(self._lib.chf_set_state(self._ctx, ctypes.byref(cs)), "chf_set_state")

    def get_state(self) -> dict:
        """Read the pre-solve state (what output writes, PF:2476-2479)."""
        out = {n: np.empty((self.Ny,) if n in rt.MEAN_FIELDS else self.field_shape, np.complex128)
               for n in rt.STATE_FIELDS}
        cs = rt.ChfState(**{n: self._dp(out[n]) for n in rt.STATE_FIELDS})
        self._check(self._lib.chf_get_state(self._ctx, ctypes.byref(cs)), "chf_get_state")
        return out

    def load_restart(self, path: str):
        """initial(isign=1) (PF:1861-1924): state from a restart file in the
        reference's two-record layout."""
        from . import restart
        p = self.params
        self.set_state(restart.read_restart(path, p.nnx, p.nny, p.nnz))

    def save_restart(self, path: str):
        """output (PF:2460-2504): the pre-solve state as a restart file."""
        from . import restart
        restart.write_restart(path, self.get_state())

    def initial(self, vnp1, omegaynp1, au0, w0):
        """Cold start from v, omega_y and the mean profiles (PF:1932-1946)."""
        v = self._as_field(vnp1, "vnp1")
        o = self._as_field(omegaynp1, "omegaynp1")
        a = self._as_mean(au0, "au0")
        w = self._as_mean(w0, "w0")
        self._check(self._lib.chf_init_fields(self._ctx, self._dp(v), self._dp(o), self._dp(a),
                                              self._dp(w)), "chf_init_fields")

    def getnonlinear(self):
        self._check(self._lib.chf_nonlinear(self._ctx), "chf_nonlinear")

    def solve(self):
        self._check(self._lib.chf_solve(self._ctx), "chf_solve")

    def step(self, n=1):
        self._check(self._lib.chf_step(self._ctx, int(n)), "chf_step")

    def sync(self):
        self._check(self._lib.chf_sync(self._ctx), "chf_sync")

    def fields(self) -> dict:
        """Post-solve u, w, omega_x, omega_z (PF:387-472)."""
        out = {n: np.empty(self.field_shape, np.complex128)
               for n in ("unp1", "wnp1", "omegaxnp1", "omegaznp1")}
        self._check(self._lib.chf_get_fields(self._ctx, *(self._dp(out[n]) for n in
                                                          ("unp1", "wnp1", "omegaxnp1", "omegaznp1"))),
                    "chf_get_fields")
        return out

    FIELDS = ("u", "v", "w", "omega_x", "omega_y", "omega_z")

    def physical(self, field) -> np.ndarray:
        """transform(1, .) (PF:948-961): the unpadded physical field after the
        last solve, double[Ny][nny][nnx] at y_k = cos(pi k/N)."""
        i = self.FIELDS.index(field) if isinstance(field, str) else int(field)
        out = np.empty((self.Ny, self.Mz, 2 * self.Mx), dtype=np.float64)
        self._check(self._lib.chf_physical(self._ctx, i, self._dp(out)), "chf_physical")
        return out

    def physical_noalloc(self, field, out_device_ptr: int):
        """chf_physical into caller-owned device memory (Ny * nny * nnx
        doubles at the device address `out_device_ptr`), allocating nothing;
        queued on the context's stream (sync() waits for it)."""
        i = self.FIELDS.index(field) if isinstance(field, str) else int(field)
        self._check(self._lib.chf_physical_noalloc(self._ctx, i, ctypes.c_void_p(out_device_ptr)),
                    "chf_physical_noalloc")

    def plane_statistics(self) -> np.ndarray:
        """The reductions of statistics (PF:2816-3020): double[12][Ny], rows
        ua va wa uu vv ww uv uw vw max|u| max|v| max|w|."""
        out = np.empty((12, self.Ny), dtype=np.float64)
        self._check(self._lib.chf_statistics(self._ctx, self._dp(out)), "chf_statistics")
        return out

    def statistics(self) -> dict:
        """The quantities statistics prints and writes (PF:2816-3190)."""
        from . import statistics
        p = self.params
        return statistics.diagnostics(self.plane_statistics(), p.nnx, p.nny, p.nnz, p.aphi, p.beta, p.dt,
                                      p.re, p.re_p)

    def tile_choices(self) -> dict:
        """Column blocks per wavefront the autotuner picked for each tuned
        contraction, with the measured ms per launch for 1, 2 and 4."""
        out = {}
        for i in range(self._lib.chf_kernel_count()):
            ms = (ctypes.c_double * 3)()
            cb = self._lib.chf_tile_choice(self._ctx, i, ms)
            if cb > 0:
                out[self._lib.chf_kernel_name(i).decode()] = dict(choice=cb, ms=[round(x, 4) for x in ms])
        return out

    def set_tile_choice(self, name: str, cb: int):
        """Overrides the autotuner's choice for one tuned contraction."""
        names = [self._lib.chf_kernel_name(i).decode() for i in range(self._lib.chf_kernel_count())]
        if name not in names:
            raise ValueError(f"unknown kernel class {name!r}")
        self._check(self._lib.chf_set_tile_choice(self._ctx, names.index(name), int(cb)), "chf_set_tile_choice")

    def device_bytes(self) -> int:
        return int(self._lib.chf_device_bytes(self._ctx))

    # --------------------------------------------------------- instrumentation
    def set_timing(self, on: bool):
        """Bracket every kernel launch with HIP events on the context stream."""
        self._check(self._lib.chf_set_timing(self._ctx, int(bool(on))), "chf_set_timing")

    def kernel_times(self) -> dict:
        """{kernel class: (total ms, launches)} since the last set_timing."""
        out = {}
        for k in range(self._lib.chf_kernel_count()):
            name = self._lib.chf_kernel_name(k).decode()
            ms, n = ctypes.c_double(), ctypes.c_int()
            self._check(self._lib.chf_timing(self._ctx, k, ctypes.byref(ms), ctypes.byref(n)),
                        "chf_timing")
            out[name] = (ms.value, n.value)
        return out
