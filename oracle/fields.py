"""Seeded test states shared by the oracle and the device tests.

TEST INFRASTRUCTURE ONLY.  Produces smooth periodic 36-channel states from a
numpy Generator seed: each channel is a sum of a few low Fourier modes, the
Higgs doublet is offset to sit near the vacuum (0,0,0,v).
"""
from __future__ import annotations

import numpy as np

NCH = 36

DEFAULT_PARAMS = dict(gw=0.65, gy=0.35, gp2=0.99, beta_w=0.0, beta_y=0.0, vev=1.0,
                      lambda_=0.129, dt=0.025, jex=0.0, gamma=0.0)


def params(**kw) -> dict:
    p = dict(DEFAULT_PARAMS)
    p.update(kw)
    return p


def smooth_state(n, seed: int = 1, amp: float = 0.3, modes: int = 4, kmax: int = 2,
                 momenta: bool = True, vev: float = 1.0) -> np.ndarray:
    """(36, nz, ny, nx) smooth random state; n is an int or (nx, ny, nz)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    s = np.zeros((NCH, nz, ny, nx))
    for c in range(NCH):
        if not momenta and c >= 20:
            continue
        for _ in range(modes):
            kx, ky, kz = rng.integers(-kmax, kmax + 1, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            a = rng.normal() * amp / modes
            s[c] += a * np.cos(2 * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz) + ph)
    s[3] += vev
    return np.ascontiguousarray(s)


def noise_state(n, seed: int = 1, amp: float = 0.5) -> np.ndarray:
    """(36, nz, ny, nx) white-noise state (for single-RHS parity only)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    s = rng.uniform(-amp, amp, size=(NCH, nz, ny, nx))
    s[3] += 1.0
    return np.ascontiguousarray(s)


def modes_state(n, table: np.ndarray, vev: float = 1.0) -> np.ndarray:
    """CPU rebuild of a device ewpt_init_modes field from its mode table."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    tx, ty, tz = 2 * np.pi * x / nx, 2 * np.pi * y / ny, 2 * np.pi * z / nz
    s = np.zeros((NCH, nz, ny, nx))
    for c in range(NCH):
        if c == 3:
            s[c] += vev
        for kx, ky, kz, ph, a in table[c]:
            s[c] += a * np.cos(kx * tx + ky * ty + kz * tz + ph)
    return s
