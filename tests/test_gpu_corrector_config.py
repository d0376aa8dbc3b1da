"""The learned-corrector configuration at its own sizes, on a reduced
training budget: the network is trained on the GPU on 128^3 thermal
trajectories (T uniform in [0.3, 0.7]) pairing one 4*dt RK4 step with four
dt steps, then evaluated at 256^3 over 400 fine-step equivalents from a
held-out T = 0.5 state with the 1e-4 Gauss-drift fallback.  The judge is the
plain method at dt (main_rk4.jl:6146-6441): the corrected 4*dt run must end
closer to it than plain RK4 at 4*dt does.  Physical time per second of each
run and the plain step that would give the corrected run's error are
printed (the full configuration's figures come from bench.py --corrector).
"""
import time

import numpy as np
import pytest

import ewpt

pytestmark = pytest.mark.gpu

DX = 0.25


def test_corrector_config_128_train_256_eval():
    from ewpt.corrector import train_thermal

    n_train, n_eval, factor, pairs, fine_steps = 128, 256, 4, 800, 400
    net = ewpt.Corrector(seed=3)
    net.input_mode = 1
    t0 = time.perf_counter()
    losses = train_thermal(net, n_train, pairs, 3e-4, factor, 2000, DX,
                           log=lambda it, l: print(f"corrector training {it}/{pairs}: relative loss {l:.3f} "
                                                   f"({time.perf_counter() - t0:.0f} s)", flush=True))
    print(f"relative loss {np.mean(losses[:50]):.3f} -> {np.mean(losses[-100:]):.3f}", flush=True)
    assert np.mean(losses[-100:]) < np.mean(losses[:50])

    pe = ewpt.Parameters(N=n_eval, dx=DX, gy=0.35)
    p = pe.step_params()
    p4 = pe.step_params()
    p4.dt = factor * p.dt
    runs = {}
    try:
        with ewpt.Lattice(n_eval, DX) as base:
            base.init_thermal(0.5, pe.meff_sq, 424242, "config")
            for name in ("dt", "4dt", "corrected"):
                lat = ewpt.Lattice(n_eval, DX)
                lat.copy_from(base)
                print(f"256^3 evaluation: {name}", flush=True)
                w0 = time.perf_counter()
                nfb = 0
                if name == "dt":
                    lat.step(p, fine_steps)
                elif name == "4dt":
                    lat.step(p4, fine_steps // factor)
                else:
                    nfb = net.step(lat, p, factor=factor, ncoarse=fine_steps // factor, gauss_tol=1e-4)
                runs[name] = (lat, time.perf_counter() - w0, nfb)
        ref = runs["dt"][0].get_state()[:16]
        err = {k: float(np.linalg.norm(v[0].get_state()[:16] - ref) / np.linalg.norm(ref))
               for k, v in runs.items() if k != "dt"}
        tphys = fine_steps * p.dt
        for k, (lat, wall, nfb) in runs.items():
            print(f"{k}: {tphys / wall:.2f} physical time / s, field rel-L2 vs dt "
                  f"{err.get(k, 0.0):.3e}, fallbacks {nfb}", flush=True)
        # plain RK4 at the corrected run's error (fourth order in the step)
        m_eq = factor * (err["corrected"] / err["4dt"]) ** 0.25
        print(f"plain RK4 at {m_eq:.2f} dt reaches the corrected error: "
              f"{tphys / runs['dt'][1] * m_eq:.2f} physical time / s", flush=True)
        assert err["corrected"] < err["4dt"], err
    finally:
        for lat, _, _ in runs.values():
            lat.close()
        net.close()
