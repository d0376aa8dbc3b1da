"""Initial conditions: the host half of the two IC branches of run().

The reference driver draws bubble centres and orientations on the host and
launches one ``initializer!`` per bubble (gpu-kernels-rk4/main_rk4.jl:5640-5698);
``bubble_table`` does the host part and ``Lattice.init_bubbles`` evaluates all
profiles in one device pass.  The thermal branch (5349-5638) is entirely on
the device here (``Lattice.init_thermal``: counter-based normal draws,
Bose-Einstein weights, hipFFT inverse transform).

Two profiles of each IC:

* ``"ref"`` (0) reproduces the reference: bubble centres on the ``bub_diam``
  sub-lattice sampled without replacement (5643-5659), the special two-bubble
  placement (5662-5667), Hopf-parametrised orientations (``gen_phi``,
  convenients.jl:8-14), the sech-like profile with periodicised distance; the
  thermal modes with the reference's index offset and Gamma = Sigma = 0.
* ``"config"`` (1) is the benchmark description: ``no_bubbles`` centres
  uniform on the lattice, uniform random SU(2) orientations (unit
  quaternions), a ``(1 - tanh)/2`` wall of radius 3/mH; thermal modes with
  signed indices and Gamma = div W, Sigma = div Y.

The Julia RNG streams (``Random.seed!``, ``sample``, ``rand``) cannot be
reproduced outside Julia: the host draws use numpy's PCG64 seeded with the
same seed, so a given seed yields a different (equally distributed)
realisation than the reference's.
"""
from __future__ import annotations

import math

import numpy as np

PROFILES = {"ref": 0, "config": 1}


def profile_id(profile) -> int:
    if isinstance(profile, str):
        if profile not in PROFILES:
            raise ValueError(f"profile must be one of {sorted(PROFILES)}, got {profile!r}")
        return PROFILES[profile]
    if profile not in (0, 1):
        raise ValueError(f"profile must be 0 or 1, got {profile!r}")
    return int(profile)


def gen_phi(abc, sign: float) -> np.ndarray:
    """Unit doublet (p1, p2, p3, p4) from three uniforms (convenients.jl:8-14):
    alpha = acos(sign*a)/2, beta = 2 pi b, gamma = 2 pi c,
    phi = (cos alpha e^{i beta}, sin alpha e^{i gamma})."""
    a, b, c = abc
    alpha = 0.5 * math.acos(sign * a)
    beta, gamma = 2.0 * math.pi * b, 2.0 * math.pi * c
    return np.array([math.cos(alpha) * math.cos(beta), math.cos(alpha) * math.sin(beta),
                     math.sin(alpha) * math.cos(gamma), math.sin(alpha) * math.sin(gamma)])


def bubble_centres_ref(N: tuple[int, int, int], bub_diam: int, no_bubbles: int, seed: int) -> np.ndarray:
    """(no_bubbles, 3) 1-based centres: a sample without replacement of the
    sub-lattice range(bub_diam, N, step=bub_diam)^3 (main_rk4.jl:5643-5659);
    two bubbles sit at (N/2 -+ 2 bub_diam, N/2, N/2) (5662-5667)."""
    nx, ny, nz = N
    if bub_diam < 1:
        raise ValueError("bub_diam must be >= 1")
    xs, ys, zs = (np.arange(bub_diam, n + 1, bub_diam) for n in (nx, ny, nz))
    cand = np.array([(a, b, c) for a in xs for b in ys for c in zs], dtype=np.float64).reshape(-1, 3)
    if no_bubbles > len(cand):
        raise ValueError(f"no_bubbles={no_bubbles} exceeds the {len(cand)} sub-lattice sites")
    rng = np.random.default_rng(seed)
    locs = cand[rng.choice(len(cand), size=no_bubbles, replace=False)] if no_bubbles else cand[:0]
    if no_bubbles == 2:
        locs = np.array([[nx // 2 - 2 * bub_diam, ny // 2, nz // 2],
                         [nx // 2 + 2 * bub_diam, ny // 2, nz // 2]], dtype=np.float64)
    return locs


def bubble_table(N, no_bubbles: int, seed: int, profile="ref", bub_diam: int | None = None,
                 with_angles: bool = False):
    """(no_bubbles, 7) rows (ib, jb, kb, p1, p2, p3, p4) for init_bubbles.

    N is an int or (nx, ny, nz).  ``"ref"`` needs ``bub_diam``
    (``Parameters.bub_diam``).  with_angles=True returns (table, angles):
    for ``"ref"`` the (no_bubbles, 3) uniform draws that gen_phi turns into
    the orientations (the reference's Hoft_arr, main_rk4.jl:5677, written as
    Hoft_angles at 6946); for ``"config"``, whose orientations are uniform
    unit quaternions with no Hopf draws, None."""
    N = (N, N, N) if np.isscalar(N) else tuple(N)
    prof = profile_id(profile)
    if no_bubbles < 0:
        raise ValueError("no_bubbles must be >= 0")
    if prof == 0:
        if bub_diam is None:
            raise ValueError("the reference profile needs bub_diam")
        locs = bubble_centres_ref(N, bub_diam, no_bubbles, seed)
        rng = np.random.default_rng(seed)                    # Random.seed!(seed_value), 5676
        hoft = rng.uniform(0.0, 1.0, size=(no_bubbles, 3))   # Hoft_arr, 5677
        signs = rng.choice([-1.0, 1.0], size=no_bubbles)     # gen_phi's sample([-1,1])
        orient = np.array([gen_phi(h, s) for h, s in zip(hoft, signs)]).reshape(-1, 4)
    else:
        rng = np.random.default_rng(seed)
        locs = rng.uniform(0.0, 1.0, size=(no_bubbles, 3)) * np.array(N, dtype=np.float64)
        q = rng.standard_normal(size=(no_bubbles, 4))
        orient = q / np.linalg.norm(q, axis=1, keepdims=True)
        hoft = None
    table = np.ascontiguousarray(np.hstack([locs, orient]), dtype=np.float64)
    return (table, hoft) if with_angles else table
