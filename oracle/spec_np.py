"""CPU restatement of the magnetic-field diagnostics (test infrastructure only).

Only ``tests/`` use this module, as the checker of ``ewpt_em_energies``,
``ewpt_em_fields`` and ``ewpt_spectrum``.

* EM fields in vector form, read off compute_Bx!/By!/Bz!
  (gpu-kernels-rk4/main_rk4.jl:2303-2822), compute_B_2_*! (3718-4250) and
  compute_Ax!/Ay!/Az! (3012-3528):
  ``n^a = -phi.T_a-like bilinears / |phi|^2`` (2459-2467),
  ``B_i = -(sin th n^a F^a_jk + cos th FY_jk + 4 sin th/(gw v^2) X_i)``,
  ``X_i = sum over the pairs (1,2), (3,4) of D_j phi_p D_k phi_q - D_j phi_q D_k phi_p``,
  (i, j, k) cyclic; B_2 has ``|phi|^4`` in place of ``v^2`` and is 0 where
  ``phi = 0``; ``A_i = sin th n^a W^a_i + cos th Y_i``.  The covariant
  derivatives and field strengths come from ``rhs_np.derived`` (generator
  form), so this is independent of the C transcription in ewpt_oracle.c.
* Shell spectrum, ``convolve_1d`` (spec_routines.jl:47-77) with
  ``K_c_mag`` (5-29) and ``Kc_bin_nums`` (31-43): the power
  ``sum_c |fft(B_c)|^2`` averaged over the lattice points of each distinct
  ``|K|`` (signed integer indices), for the first ``Kc_bin_nums(cut)``
  distinct values, as applied at main_rk4.jl:6066-6069 and 7069-7072.
"""
from __future__ import annotations

import numpy as np

from . import rhs_np


def em_fields(s: np.ndarray, p, dx: float, theta_w: float) -> np.ndarray:
    f, g = rhs_np.derived(s, p, dx)
    phi, W, Y = f["phi"], f["W"], f["Y"]
    D, Bw, BY = g["D"], g["B"], g["BY"]          # D[b,i]; Bw[a,i] = F^a_jk; BY[i]
    sw, cw = np.sin(theta_w), np.cos(theta_w)
    mag = np.sum(phi * phi, axis=0)
    safe = np.where(mag == 0.0, 1.0, mag)
    n = np.stack([-2.0 * (phi[0] * phi[2] + phi[1] * phi[3]),
                  2.0 * (phi[1] * phi[2] - phi[0] * phi[3]),
                  -phi[0] ** 2 - phi[1] ** 2 + phi[2] ** 2 + phi[3] ** 2]) / safe
    n = np.where(mag == 0.0, 0.0, n)
    out = np.zeros((9,) + s.shape[1:])
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        X = (D[0, j] * D[1, k] - D[1, j] * D[0, k]) + (D[2, j] * D[3, k] - D[3, j] * D[2, k])
        F = sw * np.sum(n * Bw[:, i], axis=0) + cw * BY[i]
        out[i] = -(F + 4.0 * sw / (p["gw"] * p["vev"] ** 2) * X)
        out[3 + i] = np.where(mag == 0.0, 0.0, -(F + 4.0 * sw / (p["gw"] * safe ** 2) * X))
        out[6 + i] = sw * np.sum(n * W[:, i], axis=0) + cw * Y[i]
    return out


def em_energies(em: np.ndarray, dx: float) -> np.ndarray:
    """Energy-table columns 8, 9, 13 (main_rk4.jl:6113-6114, 6901)."""
    vol = dx ** 3
    return np.array([0.5 * np.sum(em[0:3] ** 2) * vol, 0.5 * np.sum(em[3:6] ** 2) * vol,
                     np.sum(em[3:6] * em[6:9]) * vol])


def k2_signed(shape) -> np.ndarray:
    """Integer K^2 of every slot of an (nz, ny, nx) array, K_c_mag's signed
    index convention (x-1 <= N/2 -> x-1, else x-1-N for 1-based x)."""
    nz, ny, nx = shape
    ks = [np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n) for n in (nz, ny, nx)]
    return (ks[0][:, None, None] ** 2 + ks[1][None, :, None] ** 2 + ks[2][None, None, :] ** 2)


def kc_bin_nums(cutx: int, cuty: int, cutz: int) -> int:
    """Number of distinct |K| on a cutx*cuty*cutz lattice (Kc_bin_nums)."""
    return int(np.unique(k2_signed((cutz, cuty, cutx))).size)


def convolve_1d(power: np.ndarray, cut) -> tuple[np.ndarray, np.ndarray]:
    """(|K|, mean power) over the first kc_bin_nums(cut) distinct |K| shells."""
    nbins = kc_bin_nums(*cut)
    k2 = k2_signed(power.shape).ravel()
    vals, inv = np.unique(k2, return_inverse=True)
    sums = np.bincount(inv, weights=power.ravel())
    cnts = np.bincount(inv)
    vals, sums, cnts = vals[:nbins], sums[:nbins], cnts[:nbins]
    return np.sqrt(vals.astype(np.float64)), sums / cnts


def spectrum(B3: np.ndarray, cut) -> tuple[np.ndarray, np.ndarray]:
    """Shell spectrum of a 3-component real field (3, nz, ny, nx)."""
    F = np.fft.fftn(B3, axes=(-3, -2, -1))
    power = np.sum((np.conj(F) * F).real, axis=0)
    return convolve_1d(power, cut)
