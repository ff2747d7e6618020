"""Batched ensemble of independent channels on one GPU (C ABI chf_ensemble_*).

Members may differ in grid size, Reynolds number and precision.  Each member
is a full context (state exchange through `member(i)`, a Channel view); one
ensemble step advances every member by one time step (PF:220-532) in a single
hipGraph launch whose branches are the members.
"""
from __future__ import annotations

import ctypes

from . import runtime as rt
from .channel import Channel

PRECISION = {"fp64": 0, "mixed": 1}


class Ensemble:
    def __init__(self, members):
        """members: iterable of dicts with nnx, nny, nnz, aphi, beta, dt, re,
        re_p and optional precision ("fp64" or "mixed")."""
        self._lib = rt.load()
        members = list(members)
        arr = (rt.ChfParams * len(members))()
        self._prec = []
        for i, m in enumerate(members):
            prec = m.get("precision", "fp64")
            arr[i] = rt.ChfParams(m["nnx"], m["nny"], m["nnz"], m["aphi"], m["beta"], m["dt"], m["re"],
                                  m["re_p"], PRECISION[prec])
            self._prec.append(prec)
        self._params = arr
        self._ens = ctypes.c_void_p()
        rc = self._lib.chf_ensemble_create(arr, len(members), ctypes.byref(self._ens))
        if rc != 0:
            msg = self._lib.chf_ensemble_last_error(None).decode()
            raise rt.error(f"chf_ensemble_create failed ({rc}): {msg}", rc)
        self._members = []
        for i in range(len(members)):
            ctx = ctypes.c_void_p()
            self._check(self._lib.chf_ensemble_member(self._ens, i, ctypes.byref(ctx)), "chf_ensemble_member")
            self._members.append(Channel._borrowed(self._lib, ctx, arr[i], self._prec[i]))

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.chf_ensemble_last_error(self._ens).decode()
            raise rt.error(f"{what} failed ({rc}): {msg}", rc)

    def __len__(self):
        return len(self._members)

    def member(self, i) -> Channel:
        return self._members[i]

    def step(self, n=1):
        """n steps of every member; returns when they have completed."""
        self._check(self._lib.chf_ensemble_step(self._ens, int(n)), "chf_ensemble_step")

    def close(self):
        if self._ens:
            for m in self._members:
                m._ctx = ctypes.c_void_p()
            self._lib.chf_ensemble_destroy(self._ens)
            self._ens = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
