#!/usr/bin/env python3
"""Benchmark of the RK4 time step of the electroweak lattice evolution.

One "step" is one full RK4 step (4 stage launches) of the 36-channel
Higgs-gauge state on an N^3 periodic lattice (default 512^3, FP64), with the
state resident in HBM before the timed region starts.  The state is the
seeded thermal initial condition (Bose-Einstein modes at T = 0.5, built on
the device with hipFFT); ``--ic modes`` uses a smooth synthetic field instead.

Prints ONE JSON line (rank 0):
  value      = site-updates/s (lattice sites x RK4 steps / second), FP64
  roofline   = achieved algorithmic HBM bandwidth of the dominant stage kernel
               against MI355X's 8 TB/s, from HIP-event launch timings
  mixed      = the same for the mixed-precision path (FP32 storage/stencils,
               compensated update, FP64 reductions)
  precision  = mixed against FP64 after --precision-steps steps (default 1000)
               from the same initial state: per-channel relative L2, energy
               drift of each path, Gauss-norm ratio
  batch      = 48 batched realisations (16 each at 64^3, 96^3, 128^3), mixed,
               --batch-steps steps (default 300), autotuned, graph-captured
  cpu_baseline = the CPU oracle (literal C restatement, OpenMP) on a bounded
               sample of the same workload, on this host's cores
"""
from __future__ import annotations

import argparse
import ctypes
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
# Stored-stage-input mask of a stage variant: bit s-2 set = stage s (2..4)
# reads its input as one array written by stage s-1.
STORED_MASK = {15: 7, 24: 7, 31: 7, 32: 7, 33: 7, 34: 7}
# Variants that rebuild y from the stored inputs and the derivatives instead
# of reading it: stage 2 reads k1 (20 channels) for the next input's y, and
# stage 4 reads no y.
RECON_Y = {34}


def stage_bytes_per_site(precision: str, variant: int) -> dict:
    """Algorithmic HBM bytes per site of each stage launch: channel accesses
    times the word size (8 FP64, 4 mixed).  Formed stage inputs: S1 reads y,
    writes k1; S2 reads y, k1, writes k2; S3 reads y, the 16 momentum k1's,
    k2, writes k3; S4 reads y, k1..k3, writes y'.  A stored input u_s
    replaces the formed reads of stage s by one 36-channel array, and stage
    s-1 then also writes u_s (reading y for it).  Mixed adds the FP32
    compensation word of y, read and written at stage 4.  A variant that
    rebuilds y replaces the y read of stage 2 by the 20 k1 channels and drops
    stage 4's y read.  Variant 0 or any non-stored variant gives the minimal
    (formed) figures."""
    m = STORED_MASK.get(variant % 100, 0) if variant > 0 else 0
    recon = variant > 0 and variant % 100 in RECON_Y
    st = {s: s >= 2 and bool((m >> (s - 2)) & 1) for s in (1, 2, 3, 4, 5)}
    formed = {1: 36, 2: 36 + 20, 3: 36 + 16 + 20}
    ch = {}
    for s in (1, 2, 3):
        reads = 36 if st[s] else formed[s]
        if st[s + 1] and st[s]:
            reads += 20 if recon and s == 2 else 36    # y (or k1) for the next stage input
        ch[s] = reads + 20 + (36 if st[s + 1] else 0)
    ch[4] = 36 + 60 + (36 if st[4] and not recon else 0) + 36
    word = 8 if precision == "fp64" else 4
    if precision != "fp64":
        ch[4] += 36 + 36
    return {s: c * word for s, c in ch.items()}


STAGE_BYTES = {pr: stage_bytes_per_site(pr, 0) for pr in ("fp64", "mixed")}

# Bytes per site that one stage launch of the automatic variants moves across
# the L2/fabric boundary at 512^3, measured with rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE on this bench: FETCH_SIZE x 2 (what that counter reads for a
# plain device copy of known size) + WRITE_SIZE x 1.  Each entry names the
# profile summary that holds the counters and the commit of the library that
# was measured; a kernel changed since then needs a new counter run.
TRAFFIC_SOURCE = {
    "r1": "profiles/stage_traffic_512.md (library at commit be83b44 / 1ea9c7a)",
    "r6": "profiles/stage_traffic_512.md (library at commit c01baa4)",
}
MEASURED_TRAFFIC = {
    ("fp64", 334, 1): (1113, "r6"), ("fp64", 334, 2): (1275, "r6"), ("fp64", 334, 3): (1410, "r6"),
    ("fp64", 334, 4): (1466, "r6"),
    ("fp64", 332, 1): (1110, "r1"), ("fp64", 332, 2): (1504, "r1"), ("fp64", 332, 3): (1500, "r1"),
    ("fp64", 332, 4): (1882, "r1"),
    ("fp64", 324, 1): (1154, "r1"), ("fp64", 324, 2): (1387, "r1"), ("fp64", 324, 3): (1386, "r1"),
    ("fp64", 324, 4): (2407, "r1"),
    ("fp64", 311, 1): (849, "r1"), ("fp64", 311, 2): (1630, "r1"), ("fp64", 311, 3): (2212, "r1"),
    ("fp64", 311, 4): (3725, "r1"),
    ("mixed", 333, 1): (602, "r6"), ("mixed", 333, 2): (674, "r6"), ("mixed", 333, 3): (676, "r6"),
    ("mixed", 333, 4): (1214, "r6"),
    ("mixed", 331, 1): (632, "r1"), ("mixed", 331, 2): (850, "r1"), ("mixed", 331, 3): (845, "r1"),
    ("mixed", 331, 4): (1495, "r1"),
    ("mixed", 315, 1): (628, "r1"), ("mixed", 315, 2): (696, "r1"), ("mixed", 315, 3): (698, "r1"),
    ("mixed", 315, 4): (1358, "r1"),
}


def measured_traffic(precision: str, variant: int, stage: int, n: int):
    """(bytes per launch, source) of the dominant kernel at 512^3, or (None, None)."""
    e = MEASURED_TRAFFIC.get((precision, variant, stage)) if n == 512 else None
    if e is None:
        return None, None
    return e[0] * n ** 3, TRAFFIC_SOURCE[e[1]]


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
    """Time plain device-to-device copies of nbytes (read + write): torch's
    copy_ and libewpt's one-8-byte-access-per-lane copy kernel (the stage
    kernels' access width), for the HBM ceiling and counter calibration."""
    import ctypes

    import torch

    import ewpt
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
    ms2 = ctypes.c_double(0.0)
    rc = ewpt.load().ewpt_copy_probe(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), n, reps,
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), ctypes.byref(ms2))
    if rc != 0:
        raise RuntimeError(f"ewpt_copy_probe failed ({rc})")
    print(json.dumps({"stream_copy_bytes_read": nbytes, "stream_copy_bytes_written": nbytes,
                      "torch_copy_ms": ms, "torch_copy_GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9,
                      "fp64_lane_copy_ms": ms2.value,
                      "fp64_lane_copy_GB_per_s": 2 * nbytes / (ms2.value * 1e-3) / 1e9}))
    return 0


BATCH_SHAPES = [64] * 16 + [96] * 16 + [128] * 16


def batch_leg(steps: int, warmup: int, check_solo: bool = True) -> dict:
    """48 independent realisations (16 each at 64^3, 96^3, 128^3; even
    members thermal with T uniform in [0.3, 0.7] and J_ext = mH^2, odd
    members 6-20 bubbles; distinct seeds) stepped as one batch in mixed
    precision: one launch per stage over every lattice, graph-captured step
    pairs, tile shape from the autotuner.  With check_solo, one member of
    each size is stepped alone with the same variant from the same initial
    state and must come out bitwise equal."""
    import numpy as np

    import ewpt
    from ewpt import _lib, ic

    rng = np.random.default_rng(2024)
    dx = 0.25
    par = ewpt.Parameters(N=64, dx=dx, gy=0.35)
    lats, ps, inits = [], [], []
    print(f"bench: batch of {len(BATCH_SHAPES)} lattices, warmup={warmup} steps={steps}", file=sys.stderr, flush=True)

    def init(lat, k):
        kind, arg, seed = inits[k]
        if kind == "thermal":
            lat.init_thermal(arg, par.meff_sq, seed, "config")
        else:
            lat.init_bubbles(ic.bubble_table(lat.nx, arg, seed, "config"), "config", par.mH)

    try:
        for k, n in enumerate(BATCH_SHAPES):
            lat = ewpt.Lattice(n, dx, precision=_lib.EWPT_MIXED)
            lats.append(lat)
            if k % 2 == 0:
                inits.append(("thermal", float(rng.uniform(0.3, 0.7)), 5000 + k))
                ps.append(par.step_params(jex=par.mH ** 2))   # thermal runs start with J_ext = mH^2
            else:
                inits.append(("bubbles", int(rng.integers(6, 21)), 7000 + k))
                ps.append(par.step_params())
            init(lat, k)
        sites = sum(n ** 3 for n in BATCH_SHAPES)
        with ewpt.Batch(lats) as bt:
            variant, tune_ms = bt.tune(ps, 3)
            for k, lat in enumerate(lats):   # restart from the initial states (tuning steps use dt = 0)
                init(lat, k)
            if warmup:
                bt.step(ps, warmup)
            ms = bt.step_timed(ps, steps)
        e = [lat.energies(p) for lat, p in zip(lats, ps)]
        solo_ok = None
        if check_solo:
            solo_ok = True
            for k in (0, 17, 33):            # one member per size, both IC kinds
                with ewpt.Lattice(BATCH_SHAPES[k], dx, precision=_lib.EWPT_MIXED) as one:
                    one.variant = variant % 1000
                    init(one, k)
                    one.step(ps[k], warmup + steps)
                    solo_ok &= bool(np.array_equal(one.get_state(), lats[k].get_state(), equal_nan=True))
    finally:
        for lat in lats:
            lat.close()
    return {
        "value": sites * steps / (ms * 1e-3), "unit": "site-updates/s", "ms_per_step": ms / steps,
        "steps": steps, "warmup": warmup, "lattices": "16x64^3+16x96^3+16x128^3", "precision": "mixed",
        "ic": "even members thermal (T in [0.3, 0.7], J_ext = mH^2), odd members 6-20 bubbles",
        "variant": variant, "tuned_ms_per_step": tune_ms,
        "all_finite": bool(all(np.all(np.isfinite(x[:7])) for x in e)),
        "members_bitwise_equal_solo": solo_ok,
    }


def batch_bench(a) -> int:
    """The batched configuration as its own line."""
    r = batch_leg(a.steps, a.warmup)
    out = {
        "metric": "site-updates/s, 48 batched realisations (16 x 64^3, 16 x 96^3, 16 x 128^3), mixed precision",
        "value": r["value"], "unit": "site-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic: seeded thermal (T in [0.3, 0.7]) and bubble (6-20) ICs on the device",
        "config": {"workload": "batched RK4 steps, one launch per stage over all 48 lattices, hipGraph-captured",
                   "lattices": r["lattices"], "variant": r["variant"], "tuned_ms_per_step": r["tuned_ms_per_step"]},
        "all_finite": r["all_finite"], "members_bitwise_equal_solo": r["members_bitwise_equal_solo"],
    }
    print(json.dumps(out))
    return 0


def corrector_bench(a) -> int:
    """Learned coarse-step corrector: train on 128^3 thermal trajectories,
    evaluate at 256^3 over 2000 fine-step equivalents against plain RK4 at
    dt, 2 dt and 4 dt (errors against the dt run)."""
    import numpy as np

    import ewpt

    from ewpt.corrector import train_thermal

    dx, factor = 0.25, 4
    n_train, n_eval, fine_steps = a.corr_train_n, a.corr_eval_n, a.corr_fine_steps
    t0 = time.perf_counter()
    net = ewpt.Corrector(seed=3)
    net.input_mode = 1            # lattice Laplacian of the coarse state

    def log(it, loss):
        print(f"bench: corrector training {it}/{a.corr_iters}, relative loss {loss:.3f} "
              f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)

    losses = train_thermal(net, n_train, a.corr_iters, a.corr_lr, factor, fine_steps, dx, log=log)
    train_s = time.perf_counter() - t0
    print(f"bench: corrector trained {a.corr_iters} iterations on {n_train}^3 in {train_s:.1f} s, "
          f"relative loss {np.mean(losses[:50]):.3f} -> {np.mean(losses[-200:]):.3f}", file=sys.stderr, flush=True)

    pe = ewpt.Parameters(N=n_eval, dx=dx, gy=0.35)
    p = pe.step_params()

    def with_dt(m):
        q = pe.step_params()
        q.dt = m * p.dt
        return q

    base = ewpt.Lattice(n_eval, dx)
    base.init_thermal(0.5, pe.meff_sq, 424242, "config")
    e0 = base.energies(p)
    runs = {}
    plan = [("rk4_dt", 1), ("rk4_2dt", 2), ("rk4_4dt", 4), ("corrected_4dt", 4)]
    for name, m in plan:
        lat = ewpt.Lattice(n_eval, dx)
        lat.copy_from(base)
        nfb = 0
        w0 = time.perf_counter()
        if name.startswith("corrected"):
            nfb = net.step(lat, p, factor=m, ncoarse=fine_steps // m, gauss_tol=a.corr_tol)
        else:
            lat.step(with_dt(m), fine_steps // m)
        wall = time.perf_counter() - w0
        runs[name] = (lat, wall, nfb)
        print(f"bench: corrector eval {name}: {wall:.1f} s", file=sys.stderr, flush=True)
    ref = runs["rk4_dt"][0].get_state()[:16]
    refn = np.linalg.norm(ref)
    tphys = fine_steps * p.dt
    res = {}
    for name, (lat, wall, nfb) in runs.items():
        e = lat.energies(p)
        g = lat.gauss(p)
        err = 0.0 if name == "rk4_dt" else float(np.linalg.norm(lat.get_state()[:16] - ref) / refn)
        res[name] = {"wall_s": wall, "physical_time_per_s": tphys / wall, "field_rel_l2_vs_dt": err,
                     "energy_drift": float((sum(e[:7]) - sum(e0[:7])) / sum(e0[:7])),
                     "gauss_l2": float(np.sqrt(np.sum(np.asarray(g[:4]) ** 2))), "fallbacks": nfb}
        lat.close()
    base.close()
    net.close()
    # plain RK4 at the corrected run's error: RK4 error scales as dt^4 between the 2 dt and 4 dt runs
    e2, e4, ec = res["rk4_2dt"]["field_rel_l2_vs_dt"], res["rk4_4dt"]["field_rel_l2_vs_dt"], \
        res["corrected_4dt"]["field_rel_l2_vs_dt"]
    order = float(np.log(e4 / e2) / np.log(2.0)) if e2 > 0 and e4 > 0 else None
    m_eq = 4.0 * (ec / e4) ** (1.0 / order) if order and ec > 0 else None
    rk4_eq = res["rk4_dt"]["physical_time_per_s"] * m_eq if m_eq else None
    out = {
        "metric": f"physical time per second at {n_eval}^3 FP64: learned 4*dt corrector vs plain RK4",
        "value": res["corrected_4dt"]["physical_time_per_s"], "unit": "physical time units / s", "n_gpus": 1,
        "steps": fine_steps, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 state, bf16 network (fp32 accumulate)",
        "data": "synthetic: seeded thermal ICs (training T in [0.3, 0.7], evaluation T = 0.5)",
        "config": {"workload": f"{fine_steps} fine-step equivalents ({fine_steps // factor} corrected 4*dt steps) at "
                               f"{n_eval}^3, network trained on {n_train}^3",
                   "train_iters": a.corr_iters, "train_s": train_s, "lr": a.corr_lr, "gauss_tol": a.corr_tol,
                   "input": "7-point lattice Laplacian of the coarse state"},
        "error_reduction_vs_rk4_4dt": ec / e4 if e4 > 0 else None,
        "rk4_observed_order": order,
        "rk4_step_multiple_at_same_error": m_eq,
        "rk4_physical_time_per_s_at_same_error": rk4_eq,
        "runs": res,
    }
    print(json.dumps(out))
    return 0


XGMI_LINK = 153e9   # bytes/s of one MI355X xGMI link (one direction)


def slab_bench(a) -> int:
    """One GPU's share of the 512^3 lattice split into --shards z-slabs: the
    stage kernels on the 512 x 512 x (512/shards) slab reading their z halo
    from ghost planes, plus the halo pack/unpack of every stage (here the
    slab is its own neighbour, so the exchange is two device copies)."""
    import ewpt
    from ewpt import _lib

    n, dx, T = a.n, 0.25, 0.5
    if n % a.shards:
        print(f"bench.py: --shards {a.shards} does not divide {n}", file=sys.stderr)
        return 2
    local = (n, n, n // a.shards)
    params = ewpt.Parameters(N=n, dx=dx, gy=0.35, T=T)
    P = params.step_params()
    sites = n * n * (n // a.shards)
    out = {}
    for precision in ("fp64", "mixed"):
        prec = _lib.EWPT_FP64 if precision == "fp64" else _lib.EWPT_MIXED
        print(f"bench: slab {local} {precision} warmup={a.warmup} steps={a.steps}", file=sys.stderr, flush=True)
        with ewpt.Lattice(local, dx, precision=prec) as src:
            src.init_thermal(T, params.meff_sq, params.base_seed, "config")
            with ewpt.Lattice(local, dx, precision=prec, slab=True) as slab:
                slab.copy_from(src)
                ewpt.slab_step([slab], P, a.warmup)
                stage_ms, exch_ms, total_ms = ewpt.slab_step_timed([slab], P, a.steps)
                variant = slab.active_variant
                halo = sum(slab.halo_bytes(w) for w in range(4))    # per side per step
            # the same lattice as a periodic context (z wrapped, no exchange)
            src.variant = variant
            if a.warmup:
                src.step(P, a.warmup)
            p_stage, p_total = src.step_timed(P, a.steps)
        out[precision] = {
            "value": sites * a.steps / (total_ms * 1e-3), "ms_per_step": total_ms / a.steps, "variant": variant,
            "stage_ms": stage_ms, "exchange_ms_per_step": exch_ms,
            "periodic_ms_per_step": p_total / a.steps, "periodic_stage_ms": p_stage,
            "halo_bytes_per_side_per_step": halo,
            "xgmi_ms_per_step_estimate": halo / XGMI_LINK * 1e3,
        }
    f = out["fp64"]
    print(json.dumps({
        "metric": f"site-updates/s of one GPU's share of {n}^3 over {a.shards} GPUs (z-slab with halo exchange)",
        "value": f["value"], "unit": "site-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": f["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic: seeded thermal IC (T=0.5) of the slab's own periodic lattice",
        "config": {"workload": f"RK4 steps of a {local[0]}x{local[1]}x{local[2]} z-slab with 3 ghost planes, "
                               "halo pack/unpack before every stage (the slab is its own z-neighbour; the "
                               "transfer between GPUs is not measured: xgmi_ms_per_step_estimate is bytes / "
                               "one link's 153 GB/s)", "slab": list(local), "shards": a.shards},
        "fp64": f, "mixed": out["mixed"],
    }))
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
    ap.add_argument("--precision-steps", type=int, default=1000,
                    help="steps of the FP64-vs-mixed comparison from the same initial state (0 = off); the "
                         "warmup and timed steps are the first ones")
    ap.add_argument("--batch-steps", type=int, default=300,
                    help="steps of the batched 48-realisation leg in the default line (0 = off)")
    ap.add_argument("--batch", action="store_true",
                    help="the batched configuration instead: 48 realisations (16 each at 64^3, 96^3, 128^3), "
                         "half thermal, half bubbles, mixed precision, autotuned, graph-captured steps")
    ap.add_argument("--corrector", action="store_true",
                    help="the learned-corrector configuration instead: train on 128^3, evaluate at 256^3 "
                         "over 2000 fine-step equivalents against plain RK4")
    ap.add_argument("--corr-train-n", type=int, default=128)
    ap.add_argument("--corr-eval-n", type=int, default=256)
    ap.add_argument("--corr-fine-steps", type=int, default=2000)
    ap.add_argument("--corr-iters", type=int, default=3000)
    ap.add_argument("--corr-lr", type=float, default=3e-4)
    ap.add_argument("--corr-tol", type=float, default=1e-4)
    ap.add_argument("--slab", action="store_true",
                    help="one GPU's share of the --n lattice split into --shards z-slabs, with halo exchange")
    ap.add_argument("--shards", type=int, default=8)
    ap.add_argument("--stream-copy", action="store_true",
                    help="only time a plain device copy (HBM ceiling / counter calibration)")
    a = ap.parse_args(argv)

    if a.stream_copy:
        return stream_copy()

    if a.slab:
        if a.gpus != 1:
            print("bench.py: the slab configuration measures one GPU's share on one GPU", file=sys.stderr)
            return 2
        return slab_bench(a)

    if a.corrector:
        if a.gpus != 1:
            print("bench.py: the corrector configuration runs on one GPU", file=sys.stderr)
            return 2
        return corrector_bench(a)

    if a.batch:
        if a.gpus != 1:
            print("bench.py: the batched configuration runs on one GPU", file=sys.stderr)
            return 2
        return batch_bench(a)

    if a.gpus != 1:
        print(f"bench.py: the benchmark runs the RK4 step on one GPU; --gpus {a.gpus} is not supported "
              "(z-slab sharding runs its one-GPU share with --slab; transfers between GPUs are not run)",
              file=sys.stderr)
        return 2

    import numpy as np

    import ewpt
    from ewpt import _lib

    n = a.n
    dx = 0.25
    T = 0.5
    params = ewpt.Parameters(N=n, dx=dx, gy=0.35, T=T)
    P = params.step_params()
    sites = n ** 3
    total_steps = max(a.precision_steps, a.warmup + a.steps + 8)   # + the 8 steps of the snapshot timing
    try:
        import torch
        dev_ok = torch.cuda.is_available()
    except ImportError:
        dev_ok = False
    finals = {}

    def run(precision: str, variant: int) -> dict:
        """IC, warmup, the timed steps, then (precision leg) the remaining
        steps up to total_steps from the same IC with the same scalars; the
        final state is kept on the device for the comparison."""
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
            e0, g0 = lat.energies(P), lat.gauss(P)
            if a.warmup:
                lat.step(P, a.warmup)
            stage_ms, total_ms = lat.step_timed(P, a.steps)   # synchronises before returning
            active = lat.active_variant
            e_timed = lat.energies(P)
            # snapshot cost: a step that also reduces the whole energy row (EM
            # columns included) and the Gauss norms of its input (fused into
            # stage 1 where the variant has it) against a plain step and the
            # standalone reductions (wall clock, each call synchronous; mean of
            # 3 after one warm call)
            def wall(f, n=3):
                f()
                t0 = time.perf_counter()
                for _ in range(n):
                    f()
                return (time.perf_counter() - t0) / n * 1e3
            th = params.theta_w
            snap = {"step_ms": wall(lambda: lat.step(P, 1)),
                    "step_diag_ms": wall(lambda: lat.step_diag_em(P, th, 1)),
                    "standalone_energies_gauss_em_ms": wall(lambda: (lat.energies(P), lat.gauss(P),
                                                                     lat.em_energies(P, th)))}
            snap["fused_overhead_frac"] = snap["step_diag_ms"] / snap["step_ms"] - 1.0
            # variants with the diagnostic stage 1 (FP64 32/34, mixed 31/33)
            snap["fused"] = active % 100 in ((32, 34) if precision == "fp64" else (31, 33))
            done_snap = 8             # wall() runs each stepping call 1 + 3 times
            done = a.warmup + a.steps + done_snap
            if a.precision_steps:
                t0 = time.perf_counter()
                while done < total_steps:
                    k = min(100, total_steps - done)
                    lat.step(P, k)
                    done += k
                    print(f"bench: precision leg {precision} {done}/{total_steps} steps "
                          f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
                if dev_ok:
                    buf = torch.empty(lat.shape, dtype=torch.float64, device="cuda")
                    lat.get_state_device(buf.data_ptr())
                else:
                    buf = lat.get_state()
                finals[precision] = buf
            e1, g1 = lat.energies(P), lat.gauss(P)
        finally:
            lat.close()
        value = sites * a.steps / (total_ms * 1e-3)
        # achieved = the canonical (minimal, formed-input) bytes of the stage,
        # whatever the variant moves; the variant's own schedule beside it
        canon = STAGE_BYTES[precision]
        own = stage_bytes_per_site(precision, active)
        dom = max(range(4), key=lambda s: stage_ms[s]) + 1
        achieved = canon[dom] * sites / (stage_ms[dom - 1] * 1e-3)
        traffic, source = measured_traffic(precision, active, dom, n)
        if traffic is None and n == 512:
            if a.variant == 0:
                # the automatic variant must carry its counter figures: a kernel
                # that changed under it needs a new counter run first
                raise SystemExit(f"bench: no measured traffic for the automatic {precision} variant {active} "
                                 f"(stage {dom}) at 512^3; add its rocprofv3 counters to MEASURED_TRAFFIC")
            print(f"bench: WARNING no measured traffic for {precision} variant {active} stage {dom}: "
                  "roofline.traffic is null", file=sys.stderr, flush=True)
        return {
            "value": value,
            "ms_per_step": total_ms / a.steps,
            "variant": active,
            "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK, "traffic": traffic,
                         "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE)",
                         "traffic_source": source,
                         "kernel": f"stage {dom} ({precision}, variant {active})", "bytes_per_site": canon[dom],
                         "variant_bytes_per_site": own[dom],
                         "variant_frac": own[dom] * sites / (stage_ms[dom - 1] * 1e-3) / HBM_PEAK},
            # the step against the minimal (formed-input) traffic of an RK4 step
            "step_roofline_frac": value * sum(STAGE_BYTES[precision].values()) / HBM_PEAK,
            # every stage kernel the same way (the dominant one above is the
            # longest; stages of nearly equal time trade that place between runs)
            "stages": [{"stage": s, "ms": stage_ms[s - 1], "bytes_per_site": canon[s],
                        "frac": canon[s] * sites / (stage_ms[s - 1] * 1e-3) / HBM_PEAK,
                        "variant_bytes_per_site": own[s],
                        "variant_frac": own[s] * sites / (stage_ms[s - 1] * 1e-3) / HBM_PEAK,
                        "traffic_bytes_per_site": (MEASURED_TRAFFIC.get((precision, active, s)) or (None,))[0]
                        if n == 512 else None}
                       for s in (1, 2, 3, 4)],
            "stage_ms": stage_ms,
            "snapshot": snap,
            "energy_check": float(sum(e_timed[:7])),
            "steps_run": done,
            "energy_drift": float((sum(e1[:7]) - sum(e0[:7])) / sum(e0[:7])),
            "gauss_l2_start": float(np.sqrt(np.sum(np.asarray(g0[:4]) ** 2))),
            "gauss_l2_end": float(np.sqrt(np.sum(np.asarray(g1[:4]) ** 2))),
        }

    # same-run copy ceiling: libewpt's one-double-per-lane copy of 4 GiB
    ms_copy = ctypes.c_double(0.0)
    rc = ewpt.load().ewpt_copy_probe(None, None, (4 << 30) // 8, 10, None, ctypes.byref(ms_copy))
    copy_gbs = 2 * (4 << 30) / (ms_copy.value * 1e-3) / 1e9 if rc == 0 and ms_copy.value > 0 else None

    precs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    res = {pr: run(pr, a.variant) for pr in precs}
    for r in res.values():
        r["roofline"]["copy_ceiling_GB_per_s"] = copy_gbs
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
                   "lattice": n, "precision": precs[0], "variant": head["variant"], "ic": a.ic},
        "roofline": head["roofline"],
        "step_roofline_frac": head["step_roofline_frac"],
        "stage_ms": head["stage_ms"],
        "stages": head["stages"],
        "snapshot": head["snapshot"],
        "energy_check": head["energy_check"],
    }
    if "mixed" in res and precs[0] != "mixed":
        m = res["mixed"]
        out["mixed"] = {"value": m["value"], "ms_per_step": m["ms_per_step"], "variant": m["variant"],
                        "speedup_vs_fp64": m["value"] / head["value"], "dtype": "f32 storage/stencils, f64 reductions",
                        "roofline": m["roofline"], "step_roofline_frac": m["step_roofline_frac"],
                        "stage_ms": m["stage_ms"], "stages": m["stages"], "snapshot": m["snapshot"],
                        "energy_check": m["energy_check"],
                        "energy_rel_diff_vs_fp64": abs(m["energy_check"] - head["energy_check"]) / abs(head["energy_check"])}
    if len(finals) == 2:
        ref, mix = finals["fp64"], finals["mixed"]
        rel = []
        for c in range(36):
            if dev_ok:
                den = float(torch.linalg.vector_norm(ref[c]))
                num = float(torch.linalg.vector_norm(mix[c] - ref[c]))
            else:
                den = float(np.linalg.norm(ref[c]))
                num = float(np.linalg.norm(mix[c] - ref[c]))
            rel.append(num / den if den > 0 else num)
        del finals["fp64"], finals["mixed"], ref, mix
        f, m = res["fp64"], res["mixed"]
        out["precision"] = {
            "workload": f"mixed vs FP64 from the same {a.ic} initial state, {total_steps} RK4 steps at {n}^3, "
                        "same scalars (the timed steps included)",
            "steps": total_steps,
            "max_field_rel_l2": max(rel[:16]),
            "field_rel_l2": rel[:16], "aux_rel_l2": rel[16:20], "momentum_rel_l2": rel[20:],
            "energy_drift_fp64": f["energy_drift"], "energy_drift_mixed": m["energy_drift"],
            "energy_drift_diff": m["energy_drift"] - f["energy_drift"],
            "gauss_l2_fp64": f["gauss_l2_end"], "gauss_l2_mixed": m["gauss_l2_end"],
            "gauss_ratio_mixed_over_fp64": (m["gauss_l2_end"] / f["gauss_l2_end"]) if f["gauss_l2_end"] > 0 else None,
        }
    if a.batch_steps:
        out["batch"] = batch_leg(a.batch_steps, 2)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
