"""Host-side handle on one device lattice (one libewpt context).

``Lattice`` owns the 36-channel state on the GPU and exposes the operations
the reference performs on its CuArrays inside run()
(gpu-kernels-rk4/main_rk4.jl:5174-7081): one RHS evaluation (the rk4_step!
launch), full RK4 steps (the evolve loop body), and the energy / Gauss
reductions.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import EwptError, EwptLattice, EwptParams, NCH
from .ic import profile_id

# weak mixing angle of the reference, asin(sqrt(0.22)) (parameters.jl:17, 26)
THETA_W = float(np.arcsin(np.sqrt(0.22)))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def as_params(p) -> EwptParams:
    """EwptParams from an EwptParams, a dict, or an object with the ten fields."""
    if isinstance(p, EwptParams):
        return p
    names = ("gw", "gy", "gp2", "beta_w", "beta_y", "vev", "lambda_", "dt", "jex", "gamma")
    get = (lambda k: p[k]) if isinstance(p, dict) else (lambda k: getattr(p, k))
    return EwptParams(*[float(get(k)) for k in names])


class Lattice:
    """A periodic nx*ny*nz lattice of the electroweak fields on one GPU.

    ``slab=True``: one z-slab of a lattice split over several contexts (its
    share nx*ny*nz of the z axis, with ghost planes exchanged before each
    stage; see ``slab_step`` and the slab_* methods)."""

    def __init__(self, n, dx: float, precision: int = _lib.EWPT_FP64, device: int = 0, slab: bool = False):
        self.nx, self.ny, self.nz = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        self.dx = float(dx)
        self.precision = precision
        self.slab = bool(slab)
        self._L = _lib.load()
        self._lat = EwptLattice(self.nx, self.ny, self.nz, self.dx)
        h = ctypes.c_void_p()
        create = self._L.ewpt_create_slab if slab else self._L.ewpt_create
        rc = create(ctypes.byref(h), ctypes.byref(self._lat), precision, device)
        if rc != 0:
            raise EwptError(rc, "ewpt_create_slab" if slab else "ewpt_create")
        self._h = h

    # ---- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._L.ewpt_destroy(self._h)
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

    # ---- helpers
    @property
    def shape(self) -> tuple[int, int, int, int]:
        """numpy shape of a state: (36, nz, ny, nx), x fastest."""
        return (NCH, self.nz, self.ny, self.nx)

    def _check(self, rc: int, where: str):
        if rc != 0:
            raise EwptError(rc, where, self._L.ewpt_last_error(self._h).decode())

    def _host(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"state shape {a.shape} != {self.shape}")
        return a

    # ---- state
    def set_state(self, state: np.ndarray):
        a = self._host(state)
        self._check(self._L.ewpt_set_state(self._h, _dptr(a)), "ewpt_set_state")

    def init_modes(self, seed: int, nmodes: int = 4, kmax: int = 2, amp: float = 0.3, vev: float = 1.0,
                   mom_scale: float = 1.0) -> np.ndarray:
        """Seeded smooth state generated on the device; returns the mode table
        (36, nmodes, 5) = kx, ky, kz, phase, amplitude."""
        t = np.zeros((NCH, nmodes, 5))
        self._check(self._L.ewpt_init_modes(self._h, ctypes.c_ulonglong(seed), nmodes, kmax, amp, vev, mom_scale,
                                            _dptr(t)), "ewpt_init_modes")
        return t

    def init_bubbles(self, table: np.ndarray, profile="ref", mH: float = 2.0 * 0.129 ** 0.5):
        """Bubble-nucleation state (main_rk4.jl:5640-5698): zero fields plus the
        superposed profiles of the rows (ib, jb, kb, p1..p4) of ``table``
        (``ic.bubble_table``)."""
        t = np.ascontiguousarray(np.asarray(table, dtype=np.float64).reshape(-1, 7))
        self._check(self._L.ewpt_init_bubbles(self._h, _dptr(t), len(t), profile_id(profile), mH),
                    "ewpt_init_bubbles")

    def init_thermal(self, T: float, meff_sq: float, seed: int, profile="ref"):
        """Thermal state (main_rk4.jl:5349-5638) built on the device."""
        self._check(self._L.ewpt_init_thermal(self._h, T, meff_sq, ctypes.c_ulonglong(seed), profile_id(profile)),
                    "ewpt_init_thermal")

    def init_thermal_gauss(self, T: float, meff_sq: float, seed: int, p) -> dict:
        """Thermal fields (profile "config") with thermal E and pi projected
        onto the Gauss law for the couplings of p; returns the residual norms
        before and after the projection and the CG iteration count."""
        st = np.zeros(4)
        P = as_params(p)
        self._check(self._L.ewpt_init_thermal_gauss(self._h, T, meff_sq, ctypes.c_ulonglong(seed), ctypes.byref(P),
                                                    _dptr(st)), "ewpt_init_thermal_gauss")
        return {"residual_before": st[0], "residual_after": st[1], "iterations": int(st[2]), "su2_after": st[3]}

    def copy_from(self, other: "Lattice"):
        """Device-to-device copy of other's state (same lattice and precision)."""
        self._check(self._L.ewpt_copy_state(self._h, other._h), "ewpt_copy_state")

    def get_state(self) -> np.ndarray:
        out = np.empty(self.shape)
        self._check(self._L.ewpt_get_state(self._h, _dptr(out)), "ewpt_get_state")
        return out

    def get_state_pair(self) -> tuple[np.ndarray, np.ndarray]:
        """Mixed path: the FP32 pair (hi, lo) the state is evolved in."""
        hi, lo = np.empty(self.shape, np.float32), np.empty(self.shape, np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.ewpt_get_state_pair(self._h, hi.ctypes.data_as(fp), lo.ctypes.data_as(fp)),
                    "ewpt_get_state_pair")
        return hi, lo

    def set_state_pair(self, hi: np.ndarray, lo: np.ndarray):
        """Mixed path: set the state as its FP32 pair (state = hi + lo)."""
        a, b = (np.ascontiguousarray(x, dtype=np.float32) for x in (hi, lo))
        if a.shape != self.shape or b.shape != self.shape:
            raise ValueError(f"state pair shapes {a.shape}, {b.shape} != {self.shape}")
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self._L.ewpt_set_state_pair(self._h, a.ctypes.data_as(fp), b.ctypes.data_as(fp)),
                    "ewpt_set_state_pair")

    def set_state_device(self, ptr: int):
        """Copy the state in from a device buffer of 36*nx*ny*nz doubles
        (e.g. a torch tensor's data_ptr()) on the context's device."""
        self._check(self._L.ewpt_set_state_device(self._h, ctypes.c_void_p(ptr)), "ewpt_set_state_device")

    def get_state_device(self, ptr: int):
        """Copy the state (FP64; for the mixed path hi + lo of the FP32 pair)
        out to a device buffer of 36*nx*ny*nz doubles."""
        self._check(self._L.ewpt_get_state_device(self._h, ctypes.c_void_p(ptr)), "ewpt_get_state_device")

    # ---- the hot path
    def rhs(self, state: np.ndarray, p) -> np.ndarray:
        """One rk4_step! evaluation on host arrays (main_rk4.jl:52-661)."""
        a = self._host(state)
        out = np.empty(self.shape)
        P = as_params(p)
        self._check(self._L.ewpt_rhs(self._h, ctypes.byref(P), _dptr(a), _dptr(out)), "ewpt_rhs")
        return out

    def step(self, p, nsteps: int = 1):
        """nsteps RK4 steps with fixed scalars (main_rk4.jl:6146-6441)."""
        P = as_params(p)
        self._check(self._L.ewpt_step(self._h, ctypes.byref(P), int(nsteps)), "ewpt_step")

    def step_diag(self, p, nsteps: int = 1):
        """nsteps RK4 steps; returns (energy row, Gauss norms) of the state
        before them (columns 7, 8, 12 of the row are 0: em_energies)."""
        P = as_params(p)
        e, g = np.zeros(13), np.zeros(8)
        self._check(self._L.ewpt_step_diag(self._h, ctypes.byref(P), int(nsteps), _dptr(e), _dptr(g)),
                    "ewpt_step_diag")
        return e, g

    def step_diag_em(self, p, theta_w: float, nsteps: int = 1):
        """step_diag with the EM columns 7, 8, 12 of the energy row as well
        (em_energies of the state before the steps, Weinberg angle theta_w)."""
        P = as_params(p)
        e, g = np.zeros(13), np.zeros(8)
        self._check(self._L.ewpt_step_diag_em(self._h, ctypes.byref(P), float(theta_w), int(nsteps), _dptr(e),
                                              _dptr(g)), "ewpt_step_diag_em")
        return e, g

    def step_timed(self, p, nsteps: int):
        """Like step(); returns (mean ms per launch of stages 1..4, total ms)."""
        P = as_params(p)
        st = (ctypes.c_double * 4)()
        tot = ctypes.c_double()
        self._check(self._L.ewpt_step_timed(self._h, ctypes.byref(P), int(nsteps), st, ctypes.byref(tot)),
                    "ewpt_step_timed")
        return list(st), tot.value

    # ---- outputs
    def energies(self, p) -> np.ndarray:
        """13-column energy row (main_rk4.jl:6104-6116)."""
        out = np.zeros(13)
        P = as_params(p)
        self._check(self._L.ewpt_energies(self._h, ctypes.byref(P), _dptr(out)), "ewpt_energies")
        return out

    def em_energies(self, p, theta_w: float = THETA_W) -> np.ndarray:
        """Energy-table columns 8, 9, 13: 0.5 sum|B|^2 dx^3, 0.5 sum|B_2|^2 dx^3,
        sum B_2.A dx^3 (main_rk4.jl:6113-6114, 6901)."""
        out = np.zeros(3)
        P = as_params(p)
        self._check(self._L.ewpt_em_energies(self._h, ctypes.byref(P), float(theta_w), _dptr(out)),
                    "ewpt_em_energies")
        return out

    def energy_row(self, p, theta_w: float = THETA_W) -> np.ndarray:
        """The full 13-column row of total_energies (main_rk4.jl:6104-6118)."""
        row = self.energies(p)
        row[[7, 8, 12]] = self.em_energies(p, theta_w)
        return row

    def em_fields(self, p, theta_w: float = THETA_W) -> np.ndarray:
        """(9, nz, ny, nx): Bx, By, Bz, B2x, B2y, B2z, Ax, Ay, Az."""
        out = np.empty((9, self.nz, self.ny, self.nx))
        P = as_params(p)
        self._check(self._L.ewpt_em_fields(self._h, ctypes.byref(P), float(theta_w), _dptr(out)), "ewpt_em_fields")
        return out

    def y_slice(self, p, theta_w: float = THETA_W, j: int | None = None) -> np.ndarray:
        """(7, nz, nx) snapshot planes at y = j (default ny/2 - 1, the reference's
        1-based Ny/2): PE, KE, GE, gauge energy, sum B_i, sum B2_i, 0; times dx^3."""
        j = self.ny // 2 - 1 if j is None else int(j)
        out = np.empty((7, self.nz, self.nx))
        P = as_params(p)
        self._check(self._L.ewpt_y_slice(self._h, ctypes.byref(P), float(theta_w), j, _dptr(out)), "ewpt_y_slice")
        return out

    def spectrum(self, p, theta_w: float = THETA_W, field: str = "B", cut=None):
        """Shell spectrum (convolve_1d, spec_routines.jl:47-77) of B or B_2:
        returns (|K|, mean power) for the first Kc_bin_nums(cut) shells;
        cut defaults to spec_cut = (N/4, N/4, N/4) of parameters.jl:78."""
        if field not in ("B", "B2"):
            raise ValueError("field must be 'B' or 'B2'")
        c = (ctypes.c_int * 3)(*((self.nx // 4, self.ny // 4, self.nz // 4) if cut is None else cut))
        P = as_params(p)
        nb = ctypes.c_int(0)
        self._check(self._L.ewpt_spectrum(self._h, ctypes.byref(P), float(theta_w), int(field == "B2"), c, 0,
                                          None, None, ctypes.byref(nb)), "ewpt_spectrum")
        k = np.zeros(nb.value)
        pw = np.zeros(nb.value)
        self._check(self._L.ewpt_spectrum(self._h, ctypes.byref(P), float(theta_w), int(field == "B2"), c,
                                          nb.value, _dptr(k), _dptr(pw), ctypes.byref(nb)), "ewpt_spectrum")
        return k, pw

    def gauss(self, p) -> np.ndarray:
        """Gauss-law residual norms G^1..3, G^Y, C^1..3, C^Y."""
        out = np.zeros(8)
        P = as_params(p)
        self._check(self._L.ewpt_gauss(self._h, ctypes.byref(P), _dptr(out)), "ewpt_gauss")
        return out

    # ---- tuning
    @property
    def variant(self) -> int:
        v = ctypes.c_int()
        self._check(self._L.ewpt_get_variant(self._h, ctypes.byref(v)), "ewpt_get_variant")
        return v.value

    @variant.setter
    def variant(self, v: int):
        self._check(self._L.ewpt_set_variant(self._h, int(v)), "ewpt_set_variant")

    # ---- z-slabs and asynchronous stage control
    def halo_bytes(self, what: int = 0) -> int:
        """Bytes of one side's halo buffer: what = 0 state, 1..3 k_s."""
        b = ctypes.c_ulonglong()
        self._check(self._L.ewpt_slab_halo_bytes(self._h, int(what), ctypes.byref(b)), "ewpt_slab_halo_bytes")
        return b.value

    def slab_buffers(self) -> tuple[int, int]:
        """Device addresses of the context's own send buffers (low, high)."""
        lo, hi = ctypes.c_void_p(), ctypes.c_void_p()
        self._check(self._L.ewpt_slab_buffers(self._h, ctypes.byref(lo), ctypes.byref(hi)), "ewpt_slab_buffers")
        return lo.value, hi.value

    def slab_pack(self, what: int, send_lo: int, send_hi: int):
        """Pack the 3 lowest / highest interior planes of array `what` into
        the device buffers (asynchronous on the context stream)."""
        self._check(self._L.ewpt_slab_pack(self._h, int(what), ctypes.c_void_p(send_lo), ctypes.c_void_p(send_hi)),
                    "ewpt_slab_pack")

    def slab_unpack(self, what: int, recv_lo: int, recv_hi: int):
        """Fill the ghost planes of array `what` from the device buffers
        received from the lower / upper neighbour (asynchronous)."""
        self._check(self._L.ewpt_slab_unpack(self._h, int(what), ctypes.c_void_p(recv_lo), ctypes.c_void_p(recv_hi)),
                    "ewpt_slab_unpack")

    def stage(self, p, s: int):
        """Launch stage s (1..4) of a step, without exchange (asynchronous)."""
        P = as_params(p)
        self._check(self._L.ewpt_stage(self._h, ctypes.byref(P), int(s)), "ewpt_stage")

    @property
    def stream(self) -> int:
        """The context's HIP stream handle."""
        st = ctypes.c_void_p()
        self._check(self._L.ewpt_get_stream(self._h, ctypes.byref(st)), "ewpt_get_stream")
        return st.value or 0

    def sync(self):
        self._check(self._L.ewpt_sync(self._h), "ewpt_sync")

    @property
    def active_variant(self) -> int:
        """The stage variant the next step launches (what 0 resolves to)."""
        v = ctypes.c_int()
        self._check(self._L.ewpt_active_variant(self._h, ctypes.byref(v)), "ewpt_active_variant")
        return v.value

    def device_bytes(self) -> int:
        b = ctypes.c_ulonglong()
        self._check(self._L.ewpt_device_bytes(self._h, ctypes.byref(b)), "ewpt_device_bytes")
        return b.value


def _ring(slabs):
    if not slabs:
        raise ValueError("no slabs")
    arr = (ctypes.c_void_p * len(slabs))(*[x._h for x in slabs])
    return arr, slabs[0]


def slab_step(slabs, p, nsteps: int = 1):
    """nsteps RK4 steps of slab lattices stacked along z (slabs[0] lowest,
    periodic over the ring), exchanging halos with device copies."""
    arr, s0 = _ring(slabs)
    P = as_params(p)
    s0._check(s0._L.ewpt_slab_step(arr, len(slabs), ctypes.byref(P), int(nsteps)), "ewpt_slab_step")


def slab_step_timed(slabs, p, nsteps: int):
    """Like slab_step; returns (mean kernel ms per stage [4], mean exchange
    ms per step, total ms), measured on slabs[0]'s stream."""
    arr, s0 = _ring(slabs)
    P = as_params(p)
    st = np.zeros(4)
    ex, tot = ctypes.c_double(), ctypes.c_double()
    s0._check(s0._L.ewpt_slab_step_timed(arr, len(slabs), ctypes.byref(P), int(nsteps), _dptr(st),
                                         ctypes.byref(ex), ctypes.byref(tot)), "ewpt_slab_step_timed")
    return st.tolist(), ex.value, tot.value
