"""The run() driver (ewpt.run) and the y-slice snapshot against the oracle.

The driver's snapshots are compared with the C oracle evolving the same
initial state step by step with the same per-step J_ext / damp scalars
(tolerance 1e-10 relative on energies after tens of steps; the state itself is
held to 1e-11 by the parity tests).
"""
import math
import os

import numpy as np
import pytest

import ewpt
from ewpt import io, run as runmod
from oracle import cref, fields, ic_np, rhs_np, spec_np

pytestmark = pytest.mark.gpu

TH = math.asin(math.sqrt(0.22))
DX = 0.25


def densities(s, p, dx):
    f, g = rhs_np.derived(s, p, dx)
    phi, pi, E, EY = f["phi"], f["pi"], f["E"], f["EY"]
    r2 = (phi ** 2).sum(0)
    return dict(PE=p["lambda_"] * (r2 - p["vev"] ** 2) ** 2, KE=(pi ** 2).sum(0), GE=(g["D"] ** 2).sum((0, 1)),
                EEW=0.5 * (E ** 2).sum((0, 1)), MEW=0.5 * (g["B"] ** 2).sum((0, 1)),
                EEY=0.5 * (EY ** 2).sum(0), MEY=0.5 * (g["BY"] ** 2).sum(0))


def test_y_slice_matches_oracle():
    shape = (24, 16, 20)
    s = fields.noise_state(shape, seed=31)
    s[3] += 1.0
    p = fields.params()
    with ewpt.Lattice(shape, DX) as lat:
        lat.set_state(s)
        got = lat.y_slice(p, TH)
        got5 = lat.y_slice(p, TH, j=5)
    d = densities(s, p, DX)
    em = cref.em_fields(s, p, DX, TH)
    vol = DX ** 3
    for j, g in ((shape[1] // 2 - 1, got), (5, got5)):
        want = [d["PE"], d["KE"], d["GE"], d["MEW"] + d["EEW"] + d["MEY"] + d["EEY"],
                em[0] + em[1] + em[2], em[3] + em[4] + em[5]]
        for t, w in enumerate(want):
            w = w[:, j, :] * vol
            assert np.max(np.abs(g[t] - w)) < 1e-12 * np.max(np.abs(w)), (j, t)
        assert np.all(g[6] == 0.0)


@pytest.mark.parametrize("thermal", [False, True])
def test_driver_matches_oracle_evolution(thermal, tmp_path):
    N, nte, nsnaps = 32, 24, 4
    P = ewpt.Parameters(N=N, dx=DX, nte=nte, nsnaps=nsnaps, no_bubbles=3, gamma_inp=0.2, T=0.5, gy=0.35)
    print(f"driver thermal={thermal}", flush=True)
    res = runmod.run(P, thermal_init=thermal, profile="ref", out_dir=str(tmp_path), write_fft=True,
                     nsnaps_fft=2, verbose=False)
    # the same evolution in the oracle
    if thermal:
        y = ic_np.thermal_state((N, N, N), DX, P.T, P.meff_sq, P.base_seed, 0)
    else:
        y = ic_np.bubbles_state((N, N, N), DX, res["bubble_table"], 0, P.mH)
    base = dict(gw=P.gw, gy=P.gy, gp2=P.gp2, beta_w=P.beta_W, beta_y=P.beta_Y, vev=P.vev, lambda_=P.lambda_, dt=P.dt)
    snaps = [cref.energies(y, dict(base, jex=0.0, gamma=0.0), DX)]
    for it in range(1, nte + 1):
        jex = ewpt.J_ext(it, P.dt, P.mH, P.t_sat, thermal_init=thermal)
        gam = ewpt.damp(it, P.dt, P.mH, P.t_sat, P.gamma_inp, thermal_init=thermal)
        y = cref.rk4(y, dict(base, jex=jex, gamma=gam), DX, 1)
        if it % P.dsnaps == 0:
            snaps.append(cref.energies(y, dict(base, jex=jex, gamma=gam), DX))
            if it == nte // 2:
                y_mid = y.copy()
    e = res["energies"]
    for r, want in enumerate(snaps):
        np.testing.assert_allclose(e[r, :7], want[:7], rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(e[r, 9], want[9], rtol=1e-10)
        np.testing.assert_allclose(e[r, 11], want[11], rtol=1e-10)
    em_end = spec_np.em_energies(cref.em_fields(y, base | dict(jex=0.0, gamma=0.0), DX, P.theta_w), DX)
    np.testing.assert_allclose(e[-1, [7, 8, 12]], em_end, rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(res["final_state"], y, atol=1e-10 * np.max(np.abs(y)))
    # files
    d = io.read(str(tmp_path / "data.ewpt"))
    assert int(d["Nx"]) == N and int(d["nte"]) == nte and int(d["dsnaps"]) == P.dsnaps
    np.testing.assert_array_equal(d["energies"], e)
    assert d["y_slices_snaps"].shape == (nsnaps + 1, 7, N, N)
    assert ("bubble_locations" in d) == (not thermal)
    if not thermal:
        # Hoft_angles holds the nb x 3 uniform draws of gen_phi (main_rk4.jl:5677, 6946)
        assert d["Hoft_angles"].shape == (3, 3)
        np.testing.assert_array_equal(d["Hoft_angles"], res["hoft_angles"])
        assert np.all((d["Hoft_angles"] >= 0) & (d["Hoft_angles"] < 1))
    assert d["B_linear_fft"].shape[0] == 2 and d["B_linear_fft"].shape[2] == 2
    f = io.read(str(tmp_path / f"3d-fft-data-{nte // 2}.ewpt"))
    B = cref.em_fields(y_mid, base | dict(jex=0.0, gamma=0.0), DX, P.theta_w)
    for c, name in enumerate(("3D_fft_x", "3D_fft_y", "3D_fft_z")):
        want = np.fft.fftn(B[c].T)
        assert np.max(np.abs(f[name] - want)) < 1e-9 * np.max(np.abs(want)), name
    assert sorted(os.listdir(tmp_path)).count("data.ewpt") == 1


def test_cli(tmp_path):
    args = ["0.0", "2", "8", "0.25", "10", "32", "0.99", "2", "0.5", "0", "0", "1"]
    assert runmod.main(args + ["--bubbles", "--out", str(tmp_path)]) == 0
    d = io.read(str(tmp_path / "data.ewpt"))
    assert d["energies"].shape == (3, 13) and int(d["no_bubbles"]) == 2
    assert d["Hoft_angles"].shape == (2, 3)


@pytest.mark.parametrize("thermal,precision", [(True, ewpt.EWPT_FP64), (False, ewpt.EWPT_FP64),
                                               (True, ewpt.EWPT_MIXED)])
def test_resume_from_checkpoint_is_bitwise(thermal, precision, tmp_path):
    """A run that writes checkpoints every 6 steps, and a second run resumed
    from its last checkpoint (step 18, taken right after a snapshot, so the
    resumed run starts with the step that reduces that snapshot's energies),
    give bitwise the state, energies and slices of the uninterrupted run."""
    N, nte, nsnaps = 32, 20, 3           # snapshots every 6 steps
    P = ewpt.Parameters(N=N, dx=DX, nte=nte, nsnaps=nsnaps, no_bubbles=3, gamma_inp=0.2, T=0.5, gy=0.35)
    print(f"resume thermal={thermal} precision={precision}", flush=True)
    ck = str(tmp_path / "checkpoint.ewpt")
    os.makedirs(tmp_path / "a")
    full = runmod.run(P, thermal_init=thermal, precision=precision, out_dir=str(tmp_path / "a"),
                      verbose=False, checkpoint_path=ck, checkpoint_every=6)
    c = ewpt.checkpoint.load(ck)
    assert int(c["it"]) == 18 and int(c["snp"]) == 3 and int(np.asarray(c["pending"])[0]) == 3
    os.makedirs(tmp_path / "b")
    resumed = runmod.run(P, thermal_init=thermal, precision=precision, out_dir=str(tmp_path / "b"),
                         verbose=False, checkpoint_path=ck, resume=True)
    for k in ("energies", "slices", "gauss", "final_state"):
        assert np.array_equal(full[k], resumed[k]), k
    # the spectrum's shell sums are FP64 atomics (order-dependent rounding):
    # the same state gives the same spectrum to rounding, not bit for bit
    np.testing.assert_allclose(resumed["spectrum"], full["spectrum"], rtol=1e-12, atol=0)
    a, b = io.read(str(tmp_path / "a" / "data.ewpt")), io.read(str(tmp_path / "b" / "data.ewpt"))
    assert a.keys() == b.keys()
    for k in a:
        if k == "B_linear_fft":
            np.testing.assert_allclose(b[k], a[k], rtol=1e-12, atol=0)
        else:
            assert np.array_equal(a[k], b[k]), k
    with pytest.raises(ValueError):
        runmod.run(ewpt.Parameters(N=N, dx=DX, nte=nte + 2, nsnaps=nsnaps, gy=0.35), thermal_init=thermal,
                   precision=precision, out_dir=None, verbose=False, checkpoint_path=ck, resume=True)
