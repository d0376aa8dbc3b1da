"""Learned coarse-step corrector (ewpt_net_* of libewpt).

One RK4 step of ``factor * dt`` followed by a learned correction replaces
``factor`` RK4 steps of ``dt``; the network (periodic 3x3x3 convolutions
36 -> 64 -> 64 -> 36 with SiLU, BF16 on the matrix cores, FP32 accumulation)
is trained on the GPU on pairs the plain method produces: the state after
one coarse step and the state after the fine steps from the same start.
``step`` falls back to the fine steps whenever the Gauss-law residual grows
past the tolerance.  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import NCH, EwptError
from .lattice import Lattice, as_params


class Corrector:
    def __init__(self, seed: int = 0, device: int = 0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        rc = self._L.ewpt_net_create(ctypes.byref(h), device, ctypes.c_ulonglong(seed))
        if rc != 0:
            raise EwptError(rc, "ewpt_net_create")
        self._h = h
        self.nparams = int(self._L.ewpt_net_param_count())
        self.in_scale = np.ones(NCH, dtype=np.float32)
        self.out_scale = np.ones(NCH, dtype=np.float32)
        self._input_mode = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.ewpt_net_destroy(self._h)
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
            raise EwptError(rc, where, self._L.ewpt_net_last_error(self._h).decode())

    @staticmethod
    def _fp(a: np.ndarray):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        self._check(self._L.ewpt_net_get_params(self._h, self._fp(out)), "ewpt_net_get_params")
        return out

    @params.setter
    def params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.shape != (self.nparams,):
            raise ValueError(f"expected {self.nparams} parameters")
        self._check(self._L.ewpt_net_set_params(self._h, self._fp(a)), "ewpt_net_set_params")

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        self._check(self._L.ewpt_net_get_grads(self._h, self._fp(out)), "ewpt_net_get_grads")
        return out

    def set_norm(self, in_scale, out_scale):
        self.in_scale = np.ascontiguousarray(in_scale, dtype=np.float32)
        self.out_scale = np.ascontiguousarray(out_scale, dtype=np.float32)
        self._check(self._L.ewpt_net_set_norm(self._h, self._fp(self.in_scale), self._fp(self.out_scale)),
                    "ewpt_net_set_norm")

    @property
    def input_mode(self) -> int:
        return self._input_mode

    @input_mode.setter
    def input_mode(self, mode: int):
        self._check(self._L.ewpt_net_set_input(self._h, int(mode)), "ewpt_net_set_input")
        self._input_mode = int(mode)

    def features(self, state: np.ndarray) -> np.ndarray:
        """The network input before scaling: the state (mode 0) or its 7-point
        lattice Laplacian (mode 1)."""
        if self._input_mode == 0:
            return state
        lap = -6.0 * state
        for ax in (1, 2, 3):
            lap += np.roll(state, 1, axis=ax) + np.roll(state, -1, axis=ax)
        return lap

    def fit_norm(self, coarse: np.ndarray, fine: np.ndarray):
        """Channel scales from one training pair: the spread of the network
        input features (input) and of the fine-minus-coarse residual (output)."""
        sin = self.features(coarse).reshape(NCH, -1).std(axis=1)
        sout = (fine - coarse).reshape(NCH, -1).std(axis=1)
        sin = np.where(sin > 1e-30, sin, 1.0)
        sout = np.where(sout > 1e-30, sout, 1e-30)
        self.set_norm(sin, sout)

    def apply(self, lat: Lattice):
        self._check(self._L.ewpt_net_apply(self._h, lat._h), "ewpt_net_apply")

    def train_step(self, coarse: Lattice, fine: Lattice, lr: float) -> float:
        loss = ctypes.c_double(0.0)
        self._check(self._L.ewpt_net_train_step(self._h, coarse._h, fine._h, float(lr), ctypes.byref(loss)),
                    "ewpt_net_train_step")
        return loss.value

    def step(self, lat: Lattice, p, factor: int = 4, ncoarse: int = 1, gauss_tol: float = 1e-4) -> int:
        """ncoarse corrected coarse steps; returns how many fell back to fine steps."""
        P = as_params(p)
        nfb = ctypes.c_int(0)
        self._check(self._L.ewpt_step_corrected(lat._h, self._h, ctypes.byref(P), int(factor), int(ncoarse),
                                                float(gauss_tol), ctypes.byref(nfb)), "ewpt_step_corrected")
        return nfb.value

    def save(self, path: str):
        np.savez(path, params=self.params, in_scale=self.in_scale, out_scale=self.out_scale)

    def load(self, path: str):
        with np.load(path, allow_pickle=False) as f:
            self.params = f["params"]
            self.set_norm(f["in_scale"], f["out_scale"])


def make_pair(src: Lattice, coarse: Lattice, fine: Lattice, p, factor: int = 4):
    """coarse <- one step of factor*dt from src; fine <- factor steps of dt."""
    P = as_params(p)
    pc = type(P).from_buffer_copy(P)   # a copy: as_params returns an EwptParams unchanged
    pc.dt = P.dt * factor
    coarse.copy_from(src)
    coarse.step(pc, 1)
    fine.copy_from(src)
    fine.step(P, factor)


def train_thermal(net: "Corrector", n: int, iters: int, lr: float = 3e-4, factor: int = 4, horizon: int = 2000,
                  dx: float = 0.25, gy: float = 0.35, seed: int = 7, log=None) -> list:
    """Train net on n^3 thermal trajectories (T uniform in [0.3, 0.7]): each
    pair is one factor*dt RK4 step against factor dt steps from the same
    state.  The channel scales come from six states spread over the first
    `horizon` dt steps; a new trajectory (fresh seed and temperature, random
    start in the first horizon/2 steps) begins every 100 pairs, and the
    source advances 1-4 steps between pairs.  Returns the relative losses."""
    from .parameters import Parameters
    par = Parameters(N=n, dx=dx, gy=gy)
    p = par.step_params()
    rng = np.random.default_rng(seed)
    src, co, fi = (Lattice(n, dx) for _ in range(3))
    try:
        cs, rs = [], []
        for k in range(6):
            src.init_thermal(float(rng.uniform(0.3, 0.7)), par.meff_sq, 900 + k, "config")
            src.step(p, int(k * horizon / 12))
            make_pair(src, co, fi, p, factor)
            c, f = co.get_state(), fi.get_state()
            cs.append(net.features(c).reshape(NCH, -1).std(1))
            rs.append((f - c).reshape(NCH, -1).std(1))
        net.set_norm(np.maximum(np.mean(cs, 0), 1e-12), np.maximum(np.mean(rs, 0), 1e-30))
        losses = []
        for it in range(iters):
            if it % 100 == 0:
                src.init_thermal(float(rng.uniform(0.3, 0.7)), par.meff_sq, 1000 + it, "config")
                src.step(p, int(rng.integers(0, horizon // 2)))
            make_pair(src, co, fi, p, factor)
            losses.append(net.train_step(co, fi, lr))
            src.step(p, int(rng.integers(1, 5)))
            if log and (it + 1) % 250 == 0:
                log(it + 1, float(np.mean(losses[-100:])))
    finally:
        for x in (src, co, fi):
            x.close()
    return losses
