"""ctypes binding of the literal C oracle (test infrastructure only).

Wraps ``libewpt_oracle.so`` built by ``oracle/Makefile`` from
``ewpt_oracle.c``, which restates gpu-kernels-rk4/main_rk4.jl:52-661 (RHS),
6146-6441 (RK4 loop), 1334-2301 (energy densities) and 2303-4250 (EM fields).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libewpt_oracle.so")

NCH = 36


class OracleParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("gw", "gy", "gp2", "beta_w", "beta_y", "vev", "lambda_", "dt", "jex", "gamma")]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(OracleParams)
        i = ctypes.c_int
        L.oracle_rhs.argtypes = [i, i, i, ctypes.c_double, pp, dp, dp]
        L.oracle_rk4.argtypes = [i, i, i, ctypes.c_double, pp, dp, i]
        L.oracle_energies.argtypes = [i, i, i, ctypes.c_double, pp, dp, dp]
        L.oracle_gauss.argtypes = [i, i, i, ctypes.c_double, pp, dp, dp]
        L.oracle_em_fields.argtypes = [i, i, i, ctypes.c_double, pp, ctypes.c_double, dp, dp]
        for f in (L.oracle_rhs, L.oracle_rk4, L.oracle_energies, L.oracle_gauss, L.oracle_em_fields):
            f.restype = None
        _lib = L
    return _lib


def to_params(p) -> OracleParams:
    """Accept a mapping or an object with the ten physics fields."""
    get = (lambda k: p[k]) if isinstance(p, dict) else (lambda k: getattr(p, k))
    return OracleParams(get("gw"), get("gy"), get("gp2"), get("beta_w"), get("beta_y"),
                        get("vev"), get("lambda_"), get("dt"), get("jex"), get("gamma"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(state: np.ndarray) -> tuple[int, int, int]:
    assert state.dtype == np.float64 and state.flags.c_contiguous
    assert state.ndim == 4 and state.shape[0] == NCH, state.shape
    nz, ny, nx = state.shape[1:]
    return nx, ny, nz


def rhs(state: np.ndarray, p, dx: float) -> np.ndarray:
    """One rk4_step! evaluation: (36,nz,ny,nx) -> (36,nz,ny,nx)."""
    nx, ny, nz = _check(state)
    out = np.empty_like(state)
    P = to_params(p)
    lib().oracle_rhs(nx, ny, nz, dx, ctypes.byref(P), _ptr(state), _ptr(out))
    return out


def rk4(state: np.ndarray, p, dx: float, nsteps: int) -> np.ndarray:
    """nsteps RK4 steps; returns a new array."""
    y = np.ascontiguousarray(state, dtype=np.float64).copy()
    nx, ny, nz = _check(y)
    P = to_params(p)
    lib().oracle_rk4(nx, ny, nz, dx, ctypes.byref(P), _ptr(y), int(nsteps))
    return y


def energies(state: np.ndarray, p, dx: float) -> np.ndarray:
    nx, ny, nz = _check(state)
    out = np.zeros(13)
    P = to_params(p)
    lib().oracle_energies(nx, ny, nz, dx, ctypes.byref(P), _ptr(state), _ptr(out))
    return out


def gauss(state: np.ndarray, p, dx: float) -> np.ndarray:
    nx, ny, nz = _check(state)
    out = np.zeros(8)
    P = to_params(p)
    lib().oracle_gauss(nx, ny, nz, dx, ctypes.byref(P), _ptr(state), _ptr(out))
    return out


def em_fields(state: np.ndarray, p, dx: float, theta_w: float) -> np.ndarray:
    """(9, nz, ny, nx): Bx, By, Bz, B2x, B2y, B2z, Ax, Ay, Az."""
    nx, ny, nz = _check(state)
    out = np.zeros((9, nz, ny, nx))
    P = to_params(p)
    lib().oracle_em_fields(nx, ny, nz, dx, ctypes.byref(P), float(theta_w), _ptr(state), _ptr(out))
    return out
