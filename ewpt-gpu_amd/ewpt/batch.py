"""Many independent lattices stepped together (ewpt_batch_* of libewpt).

The reference runs one lattice per process; a study over many seeded
realisations (several lattice sizes, both initial conditions) runs here as one
batch: each ``Lattice`` keeps its own state and diagnostics, and a batch step
advances all of them with one launch per RK4 stage (two steps captured in a
hipGraph and replayed).  ``tune`` times the candidate brick variants and
launch forms and keeps the fastest.
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import EwptError, EwptParams
from .lattice import Lattice, as_params


class Batch:
    def __init__(self, lattices: list[Lattice]):
        if not lattices:
            raise ValueError("a batch needs at least one lattice")
        self._L = _lib.load()
        self.lattices = list(lattices)
        arr = (ctypes.c_void_p * len(self.lattices))(*[lat._h.value for lat in self.lattices])
        h = ctypes.c_void_p()
        rc = self._L.ewpt_batch_create(ctypes.byref(h), arr, len(self.lattices))
        if rc != 0:
            raise EwptError(rc, "ewpt_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.ewpt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, where: str):
        if rc != 0:
            raise EwptError(rc, where, self._L.ewpt_batch_last_error(self._h).decode())

    def _params(self, p) -> ctypes.Array:
        ps = p if isinstance(p, (list, tuple)) else [p] * len(self.lattices)
        if len(ps) != len(self.lattices):
            raise ValueError("one parameter set per lattice")
        return (EwptParams * len(ps))(*[as_params(x) for x in ps])

    @property
    def variant(self) -> int:
        v = ctypes.c_int(0)
        self._check(self._L.ewpt_batch_get_variant(self._h, ctypes.byref(v)), "ewpt_batch_get_variant")
        return v.value

    @variant.setter
    def variant(self, v: int):
        self._check(self._L.ewpt_batch_set_variant(self._h, int(v)), "ewpt_batch_set_variant")

    def step(self, p, nsteps: int = 1):
        """nsteps steps of every lattice; p is one parameter set or one per lattice."""
        self._check(self._L.ewpt_batch_step(self._h, self._params(p), int(nsteps)), "ewpt_batch_step")

    def step_timed(self, p, nsteps: int) -> float:
        ms = ctypes.c_double(0.0)
        self._check(self._L.ewpt_batch_step_timed(self._h, self._params(p), int(nsteps), ctypes.byref(ms)),
                    "ewpt_batch_step_timed")
        return ms.value

    def tune(self, p, nsteps: int = 3) -> tuple[int, float]:
        """Pick the fastest variant (states are left unchanged); returns
        (variant, ms per step), ms = -1 when the choice was cached."""
        v, ms = ctypes.c_int(0), ctypes.c_double(0.0)
        self._check(self._L.ewpt_batch_tune(self._h, self._params(p), int(nsteps), ctypes.byref(v), ctypes.byref(ms)),
                    "ewpt_batch_tune")
        return v.value, ms.value
