"""The run() driver of the reference (gpu-kernels-rk4/main_rk4.jl:5174-7081)
over libewpt: initial condition, evolve loop with the per-step J_ext and damp
scalars, snapshots of the energy table, y-slices and spectra, and the output
files (``ewpt.io``).  Plots are not produced.

Command line, the reference's 12 positional ARGS (parameters.jl:14-55)::

    python -m ewpt.run gamma no_bubbles nte dx dt_div N gp2 nsnaps T beta_Y beta_W seed_mult \\
        [--bubbles] [--profile ref|config] [--precision fp64|mixed] [--out DIR] [--fft] \\
        [--checkpoint-every K] [--resume]

``--checkpoint-every K`` writes ``DIR/checkpoint.ewpt`` (``ewpt.checkpoint``)
every K steps; ``--resume`` continues from it, giving bitwise the outputs of
the uninterrupted run.  The reference has no restart files.

The reference's entry point runs the thermal IC (run(thermal_init=true),
main_rk4.jl:7082); ``--bubbles`` selects the bubble branch.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import checkpoint, io
from ._lib import EWPT_FP64, EWPT_MIXED
from .ic import bubble_table, profile_id
from .lattice import Lattice
from .parameters import J_ext, Parameters, damp


def _fft_record(lat: Lattice, p, theta_w: float, path: str):
    """3d-fft-data file: full forward DFTs of B and B_2 (main_rk4.jl:6861-6868)."""
    em = lat.em_fields(p, theta_w)
    with io.Writer(path) as w:
        for c, name in enumerate(("3D_fft_x", "3D_fft_y", "3D_fft_z", "3D_fft_x2", "3D_fft_y2", "3D_fft_z2")):
            # em[c] is (nz, ny, nx) x fastest; its transpose is the Julia (Nx, Ny, Nz) array
            w.add(name, np.fft.fftn(em[c].T))


# what must agree between a checkpoint and the run that resumes from it
_RUN_KEYS = ("N", "dx", "dt", "nte", "nsnaps", "gamma_inp", "T", "gp2", "gy", "beta_W", "beta_Y", "no_bubbles",
             "seed_mult")


def run(P: Parameters, thermal_init: bool = True, profile="ref", precision: int = EWPT_FP64,
        out_dir: str | None = ".", write_fft: bool = False, nsnaps_fft: int = 50, verbose: bool = True,
        checkpoint_path: str | None = None, checkpoint_every: int = 0, resume: bool = False) -> dict:
    """Evolve one lattice as run(thermal_init) does; returns the outputs and,
    when out_dir is given, writes data.ewpt (and 3d-fft-data-<it>.ewpt files
    every dsnaps_fft steps with write_fft).  With checkpoint_every > 0 the
    run writes checkpoint_path every that many steps; resume=True continues
    from checkpoint_path instead of drawing the initial condition."""
    N, dx, dt, mH = P.N, P.dx, P.dt, P.mH
    prof = profile_id(profile)
    theta_w = P.theta_w
    dsnaps = P.dsnaps
    dsnaps_fft = max(1, P.nte // nsnaps_fft)
    if dsnaps < 1:
        raise ValueError("nte must be >= nsnaps")
    if (checkpoint_every or resume) and not checkpoint_path:
        raise ValueError("checkpointing needs checkpoint_path")
    ident = np.array([float(getattr(P, k)) for k in _RUN_KEYS] + [float(thermal_init), float(prof)])
    lat = Lattice(N, dx, precision=precision)
    res: dict = {}
    try:
        p0 = P.step_params()
        cols = [0, 1, 2, 3, 4, 5, 6, 9, 10, 11]
        cols_all = cols + [7, 8, 12]
        if resume:
            ck = checkpoint.load(checkpoint_path)
            if not np.array_equal(np.asarray(ck["run_ident"]), ident):
                raise ValueError(f"{checkpoint_path} was written by a run with other parameters")
            it0 = checkpoint.restore(lat, ck)
            energies, slices = np.array(ck["energies"]), np.array(ck["slices"])
            k0, s0 = np.array(ck["k0"]), np.array(ck["s0"])
            snp = int(ck["snp"])
            pend = np.asarray(ck["pending"])
            pending = None if pend[0] < 0 else (int(pend[0]), int(pend[1]))
            if not thermal_init:
                res["bubble_table"] = np.array(ck["bubble_table"])
                res["hoft_angles"] = np.array(ck["hoft_angles"]) if "hoft_angles" in ck else None
            del ck
        else:
            it0 = 0
            seed = P.base_seed
            if thermal_init:
                lat.init_thermal(P.T, P.meff_sq, seed, prof)
            else:
                table, hoft = bubble_table(N, P.no_bubbles, seed, prof, bub_diam=P.bub_diam, with_angles=True)
                lat.init_bubbles(table, prof, mH)
                res["bubble_table"] = table
                res["hoft_angles"] = hoft

            energies = np.zeros((P.nsnaps + 1, 13))
            slices = np.zeros((P.nsnaps + 1, 7, N, N))     # Julia (nsnaps+1, 7, Nx, Nz)
            energies[0] = lat.energy_row(p0, theta_w)
            slices[0] = np.transpose(lat.y_slice(p0, theta_w), (0, 2, 1))
            k0, s0 = lat.spectrum(p0, theta_w)
            if out_dir and write_fft:
                _fft_record(lat, p0, theta_w, os.path.join(out_dir, "3d-fft-data-0.ewpt"))
            snp = 0
            pending = None   # snapshot whose energy columns the next step's stage 1 reduces
        t0 = time.perf_counter()

        def save_checkpoint(it_k):
            rec = {"run_ident": ident, "energies": energies, "slices": slices, "k0": k0, "s0": s0, "snp": snp,
                   "pending": np.array(pending if pending is not None else (-1, -1))}
            if not thermal_init:
                rec["bubble_table"] = res["bubble_table"]
                if res["hoft_angles"] is not None:
                    rec["hoft_angles"] = res["hoft_angles"]
            checkpoint.save(checkpoint_path, lat, it_k, rec)

        def report(k, it_k):
            if verbose:
                print(f"it={it_k} t*mH={it_k * dt * mH:.3f} E_tot={energies[k, :7].sum():.9e} "
                      f"B={energies[k, 7]:.6e} ({time.perf_counter() - t0:.1f} s)", flush=True)

        for it in range(it0 + 1, P.nte + 1):
            jex = J_ext(it, dt, mH, P.t_sat, thermal_init=thermal_init)
            gam = damp(it, dt, mH, P.t_sat, P.gamma_inp, thermal_init=thermal_init)
            p = P.step_params(jex=jex, gamma=gam)
            if pending is not None:
                # the energy row of the snapshot state (EM columns included),
                # taken by this step's stage 1 from the values it stages
                # anyway (ewpt_step_diag_em)
                e, _ = lat.step_diag_em(p, theta_w, 1)
                energies[pending[0], cols_all] = e[cols_all]
                report(*pending)
                pending = None
            else:
                lat.step(p, 1)
            if it % dsnaps == 0 and snp < P.nsnaps:
                snp += 1
                slices[snp] = np.transpose(lat.y_slice(p, theta_w), (0, 2, 1))
                pending = (snp, it)
            if out_dir and write_fft and it % dsnaps_fft == 0:
                _fft_record(lat, p, theta_w, os.path.join(out_dir, f"3d-fft-data-{it}.ewpt"))
            if checkpoint_every and it % checkpoint_every == 0 and it < P.nte:
                save_checkpoint(it)
        if pending is not None:
            energies[pending[0], cols] = lat.energies(P.step_params())[cols]
            energies[pending[0], [7, 8, 12]] = lat.em_energies(P.step_params(), theta_w)
            report(*pending)
        k1, s1 = lat.spectrum(P.step_params(), theta_w)
        res.update(energies=energies, slices=slices, spectrum=np.stack([np.stack([k0, s0], 1), np.stack([k1, s1], 1)]),
                   gauss=lat.gauss(P.step_params()))
        res["final_state"] = lat.get_state() if N <= 64 else None
    finally:
        lat.close()

    if out_dir:
        with io.Writer(os.path.join(out_dir, "data.ewpt")) as w:
            for name, v in (("Nx", N), ("gp2", P.gp2), ("dx", dx), ("dt", dt), ("nte", P.nte),
                            ("gamma", P.gamma_inp), ("T", P.T), ("meff_sq", P.meff_sq), ("t_sat_jex", P.t_sat),
                            ("nsnaps", P.nsnaps), ("dsnaps", dsnaps)):
                w.add(name, v)
            if not thermal_init:
                t = res["bubble_table"]
                w.add("no_bubbles", P.no_bubbles)
                w.add("bubble_locations", t[:, :3])
                if res["hoft_angles"] is not None:
                    w.add("Hoft_angles", res["hoft_angles"])        # nb x 3 uniform draws (5677, 6946)
                else:
                    w.add("bubble_orientations", t[:, 3:])          # config profile: unit quaternions
            w.add("energies", res["energies"])
            w.add("y_slices_snaps", res["slices"])
            w.add("B_linear_fft", res["spectrum"])
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="run() of the electroweak lattice evolution")
    ap.add_argument("args", nargs=12, help="the 12 positional ARGS of parameters.jl")
    ap.add_argument("--bubbles", action="store_true", help="bubble-nucleation IC instead of thermal")
    ap.add_argument("--profile", choices=("ref", "config"), default="ref")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--gy", type=float, default=None, help="override the derived g_y")
    ap.add_argument("--out", default=".")
    ap.add_argument("--fft", action="store_true", help="write 3d-fft-data files")
    ap.add_argument("--checkpoint-every", type=int, default=0, help="write DIR/checkpoint.ewpt every K steps")
    ap.add_argument("--resume", action="store_true", help="continue from DIR/checkpoint.ewpt")
    a = ap.parse_args(argv)
    over = {} if a.gy is None else {"gy": a.gy}
    P = Parameters.from_args(a.args, **over)
    os.makedirs(a.out, exist_ok=True)
    ck = os.path.join(a.out, "checkpoint.ewpt") if (a.checkpoint_every or a.resume) else None
    run(P, thermal_init=not a.bubbles, profile=a.profile,
        precision=EWPT_FP64 if a.precision == "fp64" else EWPT_MIXED, out_dir=a.out, write_fft=a.fft,
        checkpoint_path=ck, checkpoint_every=a.checkpoint_every, resume=a.resume)
    return 0


if __name__ == "__main__":
    sys.exit(main())
