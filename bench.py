#!/usr/bin/env python3
"""Benchmark of the RK4 time step of the electroweak lattice evolution.

One "step" is one full RK4 step (4 stage launches) of the 36-channel
Higgs-gauge state on an N^3 periodic lattice (default 512^3, FP64), with the
state resident in HBM before the timed region starts.  The state is the
seeded thermal initial condition (Bose-Einstein modes at T = 0.5, built on
the device with hipFFT); ``--ic modes`` uses a smooth synthetic field instead.

Prints ONE JSON line (rank 0):
  value      = site-updates/s (lattice sites x RK4 steps / second)
  roofline   = achieved algorithmic HBM bandwidth of the dominant stage kernel
               against MI355X's 8 TB/s, from HIP-event launch timings
  cpu_baseline = the CPU oracle (literal C restatement, OpenMP) on a bounded
               sample of the same workload, on this host's cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "ewpt-gpu_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "site-updates/s per RK4 step at 512^3, FP64 vs mixed; % of HBM roofline"
HBM_PEAK = 8.0e12
# Algorithmic HBM bytes per site per stage launch (stage-separated RK4 with 20
# stored derivative channels per stage), as channel accesses times the word
# size.  FP64: S1 reads y, writes k1; S2 reads y, k1, writes k2; S3 reads y,
# the 16 momentum k1's, k2, writes k3; S4 reads y, k1..k3, writes y'.  Mixed:
# the same in FP32 plus the FP32 compensation word of y, read and written at
# stage 4 only.
STAGE_BYTES = {
    "fp64": {1: (36 + 20) * 8, 2: (36 + 20 + 20) * 8, 3: (36 + 16 + 20 + 20) * 8, 4: (36 + 60 + 36) * 8},
    "mixed": {1: (36 + 20) * 4, 2: (36 + 20 + 20) * 4, 3: (36 + 16 + 20 + 20) * 4,
              4: (36 + 36 + 60 + 36 + 36) * 4},
}


def cpu_baseline(n: int = 128, steps: int = 16):
    """Time the CPU oracle's RK4 (literal C restatement) on n^3 x steps."""
    from oracle import cref, fields
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    s = fields.smooth_state(n, seed=11)
    p = fields.params()
    cref.rk4(s, p, 0.25, 1)  # warm (page-in, thread pool)
    t0 = time.perf_counter()
    cref.rk4(s, p, 0.25, steps)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle/ewpt_oracle.c FP64 RK4, {n}^3 lattice x {steps} steps, {dt:.1f} s"}


def stream_copy(nbytes: int = 4 << 30, reps: int = 10) -> int:
    """Time a plain device-to-device copy of nbytes (read + write) with torch."""
    import torch
    n = nbytes // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"stream_copy_bytes_read": nbytes, "stream_copy_bytes_written": nbytes,
                      "ms": ms, "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9}))
    return 0


def batch_bench(a) -> int:
    """48 independent realisations stepped as one batch (mixed precision)."""
    import numpy as np

    import ewpt
    from ewpt import _lib, ic

    rng = np.random.default_rng(2024)
    dx = 0.25
    par = ewpt.Parameters(N=64, dx=dx, gy=0.35)
    shapes = [64] * 16 + [96] * 16 + [128] * 16
    lats, ps = [], []
    print(f"bench: batch of {len(shapes)} lattices, warmup={a.warmup} steps={a.steps}", file=sys.stderr, flush=True)
    try:
        for k, n in enumerate(shapes):
            lat = ewpt.Lattice(n, dx, precision=_lib.EWPT_MIXED)
            lats.append(lat)
            if k % 2 == 0:
                lat.init_thermal(float(rng.uniform(0.3, 0.7)), par.meff_sq, 5000 + k, "config")
                ps.append(par.step_params(jex=par.mH ** 2))   # thermal runs start with J_ext = mH^2
            else:
                nb = int(rng.integers(6, 21))
                lat.init_bubbles(ic.bubble_table(n, nb, 7000 + k, "config"), "config", par.mH)
                ps.append(par.step_params())
        sites = sum(n ** 3 for n in shapes)
        with ewpt.Batch(lats) as bt:
            variant, tune_ms = bt.tune(ps, 3)
            if a.warmup:
                bt.step(ps, a.warmup)
            ms = bt.step_timed(ps, a.steps)
            e = [lat.energies(p) for lat, p in zip(lats, ps)]
    finally:
        for lat in lats:
            lat.close()
    value = sites * a.steps / (ms * 1e-3)
    out = {
        "metric": "site-updates/s, 48 batched realisations (16 x 64^3, 16 x 96^3, 16 x 128^3), mixed precision",
        "value": value, "unit": "site-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic: seeded thermal (T in [0.3, 0.7]) and bubble (6-20) ICs on the device",
        "config": {"workload": "batched RK4 steps, one launch per stage over all 48 lattices, hipGraph-captured",
                   "lattices": "16x64^3+16x96^3+16x128^3", "variant": variant, "tuned_ms_per_step": tune_ms},
        "all_finite": bool(all(np.all(np.isfinite(x[:7])) for x in e)),
    }
    print(json.dumps(out))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512, help="lattice edge")
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both",
                    help="both = FP64 headline value plus the mixed path in the same run")
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--ic", choices=("thermal", "modes"), default="thermal",
                    help="initial state: thermal IC (T=0.5) or a smooth seeded mode sum")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch", action="store_true",
                    help="the batched configuration instead: 48 realisations (16 each at 64^3, 96^3, 128^3), "
                         "half thermal, half bubbles, mixed precision, autotuned, graph-captured steps")
    ap.add_argument("--stream-copy", action="store_true",
                    help="only time a plain device copy (HBM ceiling / counter calibration)")
    a = ap.parse_args(argv)

    if a.stream_copy:
        return stream_copy()

    if a.batch:
        if a.gpus != 1:
            print("bench.py: the batched configuration runs on one GPU", file=sys.stderr)
            return 2
        return batch_bench(a)

    if a.gpus != 1:
        print(f"bench.py: this build runs the hot path on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2

    import ewpt
    from ewpt import _lib

    n = a.n
    dx = 0.25
    T = 0.5
    params = ewpt.Parameters(N=n, dx=dx, gy=0.35, T=T)
    P = params.step_params()
    sites = n ** 3

    def run(precision: str, variant: int) -> dict:
        prec = _lib.EWPT_FP64 if precision == "fp64" else _lib.EWPT_MIXED
        print(f"bench: {n}^3 {precision} variant={variant} warmup={a.warmup} steps={a.steps}",
              file=sys.stderr, flush=True)
        lat = ewpt.Lattice(n, dx, precision=prec)
        try:
            lat.variant = variant
            if a.ic == "thermal":
                lat.init_thermal(T, params.meff_sq, params.base_seed, "config")
            else:
                lat.init_modes(seed=2024, nmodes=4, kmax=3, amp=0.3)
            if a.warmup:
                lat.step(P, a.warmup)
            stage_ms, total_ms = lat.step_timed(P, a.steps)   # synchronises before returning
            e = lat.energies(P)
        finally:
            lat.close()
        value = sites * a.steps / (total_ms * 1e-3)
        stage_bytes = {s: STAGE_BYTES[precision][s] * sites for s in (1, 2, 3, 4)}
        dom = max(range(4), key=lambda s: stage_ms[s]) + 1
        achieved = stage_bytes[dom] / (stage_ms[dom - 1] * 1e-3)
        return {
            "value": value,
            "ms_per_step": total_ms / a.steps,
            "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK, "traffic": None,
                         "kernel": f"stage {dom} ({precision})", "bytes_per_site": stage_bytes[dom] // sites},
            "step_roofline_frac": value * sum(stage_bytes.values()) / sites / HBM_PEAK,
            "stage_ms": stage_ms,
            "energy_check": float(sum(e[:7])),
        }

    precs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    res = {pr: run(pr, a.variant) for pr in precs}
    head = res[precs[0]]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precs[0] == "fp64" else "f32+f64",
        "data": ("synthetic: seeded thermal IC (Bose-Einstein modes, T=0.5) built on the device"
                 if a.ic == "thermal" else "synthetic: seeded smooth Fourier-mode fields generated on the device"),
        "config": {"workload": f"RK4 step of the 36-channel Higgs-gauge lattice, {n}^3 periodic, "
                               f"{precs[0]}, dx=0.25, dt=dx/10, {a.ic} initial state",
                   "lattice": n, "precision": precs[0], "variant": a.variant, "ic": a.ic},
        "roofline": head["roofline"],
        "step_roofline_frac": head["step_roofline_frac"],
        "stage_ms": head["stage_ms"],
        "energy_check": head["energy_check"],
    }
    if "mixed" in res and precs[0] != "mixed":
        m = res["mixed"]
        out["mixed"] = {"value": m["value"], "ms_per_step": m["ms_per_step"],
                        "speedup_vs_fp64": m["value"] / head["value"], "dtype": "f32 storage/stencils, f64 reductions",
                        "roofline": m["roofline"], "step_roofline_frac": m["step_roofline_frac"],
                        "stage_ms": m["stage_ms"], "energy_check": m["energy_check"],
                        "energy_rel_diff_vs_fp64": abs(m["energy_check"] - head["energy_check"]) / abs(head["energy_check"])}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
