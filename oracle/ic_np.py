"""CPU restatement of the initial-condition modules (test infrastructure only).

Only ``tests/`` use this module, as the checker of ``ewpt_init_bubbles`` and
``ewpt_init_thermal``.

* Bubble profile, ``initializer!`` (gpu-kernels-rk4/main_rk4.jl:4930-4946):
  ``rmag = (1+a^2) exp(-mH r/sqrt2) / (1 + a^2 exp(-sqrt2 mH r))``,
  ``a = sqrt2 - 1``, with the periodicised distance
  ``r^2 = sum_d sin^2(rk_d ((i_d-1) - (ib_d-0.5)) dx) / rk_d^2``,
  ``rk_d = pi/((N_d-1) dx)`` (4935-4936).  Algebraically
  ``rmag = sqrt2 sech(mH r/sqrt2 + asinh 1)``, which the tests use as an
  independent form.  Profile 1 is a tanh wall
  ``(1 - tanh((r - 3/mH) mH/2))/2`` at the minimum-image distance.
* Orientation, ``gen_phi`` (convenients.jl:8-14): ``alpha = acos(s a)/2``,
  ``beta = 2 pi b``, ``gamma = 2 pi c``,
  ``phi = (cos alpha e^{i beta}, sin alpha e^{i gamma})``.
* Thermal modes, ``thermal_initializer_gauge!`` (5076-5152) and
  ``thermal_initializer_phi!`` (5029-5074), driver 5349-5638: lattice momentum
  ``k_d = 2 sin(pi m_d/N_d)/dx``,
  ``|k|^2 = 2 sum_d (1 - cos(2 pi m_d/N_d))/dx^2``, occupation
  ``n = 1/(exp(omega/T) - 1)``, amplitude ``sqrt(n/(omega vol))``,
  polarisations ``eps1 = r x k/|r x k|``, ``eps2 = eps1 x k/|eps1 x k|``,
  ``A = amp (eps1 x1 + eps2 x2) + i amp (eps1 x3 + eps2 x4)``; the field is
  ``real(bfft(A))`` (5415-5635), bfft being the unnormalised inverse DFT.
  The reference evaluates mode array index m (0-based) at m_d = m+1
  (its kernels use the 1-based thread index directly); profile 0 keeps that,
  profile 1 uses the signed index.  The normal draws N(0, sqrt2) come from a
  counter-based generator (splitmix64 finaliser + Box-Muller), keyed by
  (seed, stream, mode index), which the device reproduces; the Julia
  ``Random.seed!``/``rand(Normal)`` streams of the reference cannot be
  reproduced outside Julia.

Channel order of the returned (36, nz, ny, nx) states: 0-3 phi, 4-12 W^a_i
(a outer), 13-15 Y_i, 16-18 Gamma^a, 19 Sigma, 20-35 momenta (zero).
"""
from __future__ import annotations

import numpy as np

NCH = 36
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def gauss_sqrt2(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """N(0, sqrt2) draws (the reference's ``Normal(0.0, sqrt(2.0))``) for the
    given mode indices of one stream."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = _mix64(np.array([seed], dtype=np.uint64) + _GOLD * np.uint64(stream + 1))[0]
        a = _mix64(base + _GOLD * (np.uint64(2) * idx + np.uint64(1)))
        b = _mix64(base + _GOLD * (np.uint64(2) * idx + np.uint64(2)))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(2.0) * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------- bubbles

def rmag_ref(r, mH: float):
    """Reference bubble profile (main_rk4.jl:4937)."""
    a2 = (np.sqrt(2.0) - 1.0) ** 2
    return (1.0 + a2) * np.exp(-mH * r / np.sqrt(2.0)) / (1.0 + a2 * np.exp(-np.sqrt(2.0) * mH * r))


def rmag_tanh(r, mH: float):
    return 0.5 * (1.0 - np.tanh((r - 3.0 / mH) * mH * 0.5))


def _coords(nx, ny, nz):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i, j, k


def dist_ref(shape, dx: float, ib: float, jb: float, kb: float) -> np.ndarray:
    """Periodicised distance of initializer! (main_rk4.jl:4935); i 0-based
    here, i.e. the reference's (i-1)."""
    nx, ny, nz = shape
    i, j, k = _coords(nx, ny, nz)
    r2 = 0.0
    for idx, c, n in ((i, ib, nx), (j, jb, ny), (k, kb, nz)):
        rk = np.pi / ((n - 1) * dx)
        r2 = r2 + np.sin(rk * (idx - (c - 0.5)) * dx) ** 2 / rk ** 2
    return np.sqrt(r2)


def dist_min_image(shape, dx: float, cx: float, cy: float, cz: float) -> np.ndarray:
    nx, ny, nz = shape
    i, j, k = _coords(nx, ny, nz)
    r2 = 0.0
    for idx, c, n in ((i, cx, nx), (j, cy, ny), (k, cz, nz)):
        d = idx - c
        d = d - n * np.rint(d / n)
        r2 = r2 + (d * dx) ** 2
    return np.sqrt(r2)


def bubbles_state(shape, dx: float, table: np.ndarray, profile: int, mH: float) -> np.ndarray:
    """State with phi = sum_b rmag_b p_b, everything else zero."""
    nx, ny, nz = shape
    s = np.zeros((NCH, nz, ny, nx))
    for row in np.asarray(table, dtype=np.float64).reshape(-1, 7):
        if profile == 0:
            rm = rmag_ref(dist_ref(shape, dx, *row[:3]), mH)
        else:
            rm = rmag_tanh(dist_min_image(shape, dx, *row[:3]), mH)
        for c in range(4):
            s[c] += rm * row[3 + c]
    return s


def gen_phi(abc, sign: float) -> np.ndarray:
    """(p1, p2, p3, p4) = (re, im) of gen_phi's doublet (convenients.jl:8-14)."""
    a, b, c = abc
    alpha = 0.5 * np.arccos(sign * a)
    z1 = np.cos(alpha) * np.exp(1j * 2.0 * np.pi * b)
    z2 = np.sin(alpha) * np.exp(1j * 2.0 * np.pi * c)
    return np.array([z1.real, z1.imag, z2.real, z2.imag])


# ---------------------------------------------------------------- thermal

def mode_numbers(shape, profile: int):
    """Integer mode numbers (m_x, m_y, m_z) of every (x fastest) array slot."""
    nx, ny, nz = shape
    i, j, k = _coords(nx, ny, nz)
    if profile == 0:
        return i + 1, j + 1, k + 1
    return tuple(np.where(a <= n // 2, a, a - n) for a, n in ((i, nx), (j, ny), (k, nz)))


def lattice_k(shape, dx: float, profile: int):
    """(kx, ky, kz, |k|) per slot (main_rk4.jl:5040-5046, 5089-5095)."""
    nx, ny, nz = shape
    mi, mj, mk = mode_numbers(shape, profile)
    kx = 2.0 * np.sin(0.5 * mi * 2.0 * np.pi / nx) / dx
    ky = 2.0 * np.sin(0.5 * mj * 2.0 * np.pi / ny) / dx
    kz = 2.0 * np.sin(0.5 * mk * 2.0 * np.pi / nz) / dx
    modk = np.sqrt(2.0 * ((1.0 - np.cos(mi * 2.0 * np.pi / nx)) / dx ** 2 +
                          (1.0 - np.cos(mj * 2.0 * np.pi / ny)) / dx ** 2 +
                          (1.0 - np.cos(mk * 2.0 * np.pi / nz)) / dx ** 2))
    return kx, ky, kz, modk


def be_amplitude(omega, T: float, vol: float):
    """sqrt(n/(omega vol)), n = 1/(exp(omega/T)-1); 0 where omega == 0."""
    w = np.where(omega > 0, omega, 1.0)
    amp = np.sqrt((1.0 / np.expm1(w / T)) / (w * vol))
    return np.where(omega > 0, amp, 0.0)


def gauge_modes(shape, dx: float, T: float, seed: int, grp: int, profile: int) -> np.ndarray:
    """Complex mode arrays (3, nz, ny, nx) of gauge group grp (0..2 W^a, 3 Y)."""
    nx, ny, nz = shape
    n3 = nx * ny * nz
    q = np.arange(n3, dtype=np.uint64)
    kx, ky, kz, modk = (a.reshape(-1) for a in lattice_k(shape, dx, profile))
    s0 = 10 * grp
    r1, r2, r3 = (gauss_sqrt2(seed, s0 + t, q) for t in range(3))
    x1, x2, x3, x4 = (gauss_sqrt2(seed, s0 + 6 + t, q) for t in range(4))
    k = np.stack([kx, ky, kz])
    e1 = np.cross(np.stack([r1, r2, r3]), k, axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        e1 = e1 / np.linalg.norm(e1, axis=0)
        e2 = np.cross(e1, k, axis=0)
        e2 = e2 / np.linalg.norm(e2, axis=0)
    amp = be_amplitude(modk, T, n3 * dx ** 3)
    out = np.zeros((3, n3), dtype=np.complex128)
    nz_ = modk != 0
    out[:, nz_] = amp[nz_] * (e1[:, nz_] * x1[nz_] + e2[:, nz_] * x2[nz_]) + \
        1j * amp[nz_] * (e1[:, nz_] * x3[nz_] + e2[:, nz_] * x4[nz_])
    return out.reshape(3, nz, ny, nx)


def higgs_modes(shape, dx: float, T: float, meff_sq: float, seed: int, b: int, profile: int) -> np.ndarray:
    nx, ny, nz = shape
    n3 = nx * ny * nz
    q = np.arange(n3, dtype=np.uint64)
    modk = lattice_k(shape, dx, profile)[3].reshape(-1)
    amp = be_amplitude(np.sqrt(modk ** 2 + meff_sq), T, n3 * dx ** 3)
    return (amp * gauss_sqrt2(seed, 40 + b, q)).reshape(nz, ny, nx).astype(np.complex128)


def bfft(a: np.ndarray) -> np.ndarray:
    """Unnormalised inverse DFT over the three lattice axes (FFTW's bfft)."""
    return np.fft.ifftn(a, axes=(-3, -2, -1)) * (a.shape[-1] * a.shape[-2] * a.shape[-3])


def _d1(A, axis, dx):
    c = lambda s: np.roll(A, -s, axis=axis)
    return (45.0 * (c(1) - c(-1)) - 9.0 * (c(2) - c(-2)) + (c(3) - c(-3))) / (60.0 * dx)


def thermal_state(shape, dx: float, T: float, meff_sq: float, seed: int, profile: int) -> np.ndarray:
    nx, ny, nz = shape
    s = np.zeros((NCH, nz, ny, nx))
    for grp in range(4):
        f = bfft(gauge_modes(shape, dx, T, seed, grp, profile)).real
        ch = 4 + 3 * grp if grp < 3 else 13
        s[ch:ch + 3] = f
    for b in range(4):
        s[b] = bfft(higgs_modes(shape, dx, T, meff_sq, seed, b, profile)).real
    if profile == 1:
        for grp in range(4):
            ch = 4 + 3 * grp if grp < 3 else 13
            div = sum(_d1(s[ch + ax], -1 - ax, dx) for ax in range(3))
            s[16 + grp] = div
    return s
