"""Checkpoint and resume of an evolution.

The reference writes no restart files: its run() starts from an initial
condition every time (gpu-kernels-rk4/main_rk4.jl:5174-7081).  A checkpoint
here is one container file in the ``ewpt.io`` layout holding what an
evolution needs to go on as if it had not stopped:

* ``ckpt_version`` (1), ``Nx Ny Nz dx``, ``precision`` (0 FP64, 1 mixed),
  ``it`` (steps taken);
* FP64: ``state``, the 36 channels as one Nx x Ny x Nz x 36 float64 array
  (Julia order, x fastest: the library's channel-major buffer as it lies in
  memory); mixed: ``state_hi`` and ``state_lo``, the FP32 pair the state is
  evolved in (state = hi + lo), in the same layout;
* the caller's own records (``ewpt.run`` adds its parameters, snapshot
  arrays and counters).

Restoring the state sets exactly the numbers the lattice held, so a resumed
run continues bit for bit.  The file is written under a temporary name and
renamed, so an interrupted write leaves the previous checkpoint intact.
"""
from __future__ import annotations

import os

import numpy as np

from . import io
from ._lib import EWPT_FP64, EWPT_MIXED

VERSION = 1


def save(path: str, lat, it: int, records: dict | None = None) -> None:
    """Write lat's state after `it` steps, plus `records`, to path."""
    tmp = path + ".tmp"
    with io.Writer(tmp) as w:
        w.add("ckpt_version", VERSION)
        for name, v in (("Nx", lat.nx), ("Ny", lat.ny), ("Nz", lat.nz), ("dx", lat.dx),
                        ("precision", int(lat.precision)), ("it", int(it))):
            w.add(name, v)
        if lat.precision == EWPT_MIXED:
            hi, lo = lat.get_state_pair()
            w.add("state_hi", hi.T)     # (36, nz, ny, nx) C order = (nx, ny, nz, 36) Julia order
            del hi
            w.add("state_lo", lo.T)
        else:
            w.add("state", lat.get_state().T)
        for name, v in (records or {}).items():
            w.add(name, v)
    os.replace(tmp, path)


def load(path: str, mmap: bool = True) -> dict:
    """All records of a checkpoint (state arrays memory-mapped by default)."""
    ck = io.read(path, mmap=mmap)
    if int(ck.get("ckpt_version", -1)) != VERSION:
        raise ValueError(f"{path}: not a version-{VERSION} checkpoint")
    return ck


def restore(lat, ck: dict) -> int:
    """Set lat's state from a loaded checkpoint; returns the step count."""
    dims = tuple(int(ck[k]) for k in ("Nx", "Ny", "Nz"))
    if dims != (lat.nx, lat.ny, lat.nz) or float(ck["dx"]) != lat.dx:
        raise ValueError(f"checkpoint lattice {dims}, dx={float(ck['dx'])} does not match "
                         f"{(lat.nx, lat.ny, lat.nz)}, dx={lat.dx}")
    prec = int(ck["precision"])
    if prec != int(lat.precision):
        raise ValueError(f"checkpoint precision {prec} does not match the lattice's {int(lat.precision)}")
    if prec == EWPT_MIXED:
        lat.set_state_pair(np.asarray(ck["state_hi"]).T, np.asarray(ck["state_lo"]).T)
    else:
        assert prec == EWPT_FP64
        lat.set_state(np.asarray(ck["state"]).T)
    return int(ck["it"])
