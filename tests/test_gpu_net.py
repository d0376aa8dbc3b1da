"""Corrector network on the GPU against its CPU restatement (oracle/net_np.py).

Layer kernels are compared with identical bf16 inputs, so the difference is
FP32 accumulation on the matrix cores against FP64 (bound: 1e-5 of the
magnitude sum |X||W| per output); bf16-rounded outputs may differ by one
bf16 unit where that accumulation difference crosses a rounding boundary.
End-to-end gradients pass through three bf16 roundings of intermediate
values that depend on the accumulated ones, so they are compared by relative
L2 error (2e-2) and direction (cosine > 0.999).
"""
import math

import numpy as np
import pytest

import ewpt
from oracle import fields, net_np

pytestmark = pytest.mark.gpu

SHAPE = (16, 8, 8)   # nx, ny, nz
MEFF = 8119.43 / 174.0 ** 2


def _rand_bf16(rng, shape, scale=1.0, zero_from=None):
    a = net_np.bf16(scale * rng.standard_normal(shape))
    if zero_from is not None:
        a[..., zero_from:] = 0.0
    return a


def _lib():
    return ewpt.load()


def _call_conv(mode, X, W, bias=None, zprev=None):
    import ctypes
    nz, ny, nx = X.shape[:3]
    Xb = np.ascontiguousarray(net_np.to_bits(X))
    Wb = np.ascontiguousarray(net_np.to_bits(W))
    Z = np.zeros(X.shape[:3] + (64,), dtype=np.float32)
    A = np.zeros(X.shape[:3] + (64,), dtype=np.uint16)
    b = np.ascontiguousarray(bias if bias is not None else np.zeros(64), dtype=np.float32)
    zp = np.ascontiguousarray(zprev if zprev is not None else np.zeros(X.shape[:3] + (64,)), dtype=np.float32)
    hp = ctypes.POINTER(ctypes.c_ushort)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = _lib().ewpt_net_conv_test(nx, ny, nz, mode, Xb.ctypes.data_as(hp), Wb.ctypes.data_as(hp),
                                   b.ctypes.data_as(fp), zp.ctypes.data_as(fp), Z.ctypes.data_as(fp),
                                   A.ctypes.data_as(hp))
    assert rc == 0
    return Z.astype(np.float64), net_np.from_bits(A)


def _close_bf16(got, want):
    return np.all(np.abs(got - want) <= 2.0 ** -7 * np.abs(want) + 1e-30)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_conv_layer_matches_oracle(mode):
    rng = np.random.default_rng(10 + mode)
    X = _rand_bf16(rng, SHAPE[::-1] + (64,))
    W = _rand_bf16(rng, (27, 64, 64), 0.05)
    b = rng.standard_normal(64).astype(np.float32).astype(np.float64)
    bound = 1e-5 * np.max(net_np.conv(np.abs(X), np.abs(W)))
    if mode < 2:
        Z, A = _call_conv(mode, X, W, bias=b)
        want = net_np.conv(X, W, b)
        assert np.max(np.abs(Z - want)) < bound
        if mode == 0:
            assert _close_bf16(A, net_np.bf16(net_np.silu(Z)))
            assert np.mean(A != net_np.bf16(net_np.silu(want))) < 0.01
    else:
        zprev = rng.standard_normal(SHAPE[::-1] + (64,)).astype(np.float32).astype(np.float64)
        _, A = _call_conv(2, X, net_np.mirrored(W), zprev=zprev)
        g = net_np.conv_dgrad(X, W)
        want = net_np.bf16(g * net_np.dsilu(zprev))
        assert np.max(np.abs(A - want)) <= 2.0 ** -7 * np.max(np.abs(want)) + bound


def test_wgrad_matches_oracle():
    import ctypes
    rng = np.random.default_rng(21)
    dZ = _rand_bf16(rng, SHAPE[::-1] + (64,), 0.01)
    X = _rand_bf16(rng, SHAPE[::-1] + (64,))
    dW = np.zeros((27, 64, 64), dtype=np.float32)
    db = np.zeros(64, dtype=np.float32)
    hp, fp = ctypes.POINTER(ctypes.c_ushort), ctypes.POINTER(ctypes.c_float)
    a, x = np.ascontiguousarray(net_np.to_bits(dZ)), np.ascontiguousarray(net_np.to_bits(X))
    rc = _lib().ewpt_net_wgrad_test(SHAPE[0], SHAPE[1], SHAPE[2], a.ctypes.data_as(hp), x.ctypes.data_as(hp),
                                    dW.ctypes.data_as(fp), db.ctypes.data_as(fp))
    assert rc == 0
    wW, wb = net_np.conv_wgrad(dZ, X)
    scale, _ = net_np.conv_wgrad(np.abs(dZ), np.abs(X))
    assert np.max(np.abs(dW - wW)) < 1e-5 * np.max(scale)
    assert np.max(np.abs(db - wb)) < 1e-5 * np.sum(np.abs(dZ), axis=(0, 1, 2)).max()


def _pair(shape, seed):
    s = fields.smooth_state(shape, seed=seed, amp=0.3)
    f = s + 1e-3 * fields.smooth_state(shape, seed=seed + 1, amp=1.0, vev=0.0)
    return s, f


def test_training_gradients_match_oracle():
    c, f = _pair(SHAPE, 5)
    P = net_np.init_params(7).astype(np.float32)
    in_scale = np.maximum(c.reshape(36, -1).std(1), 1e-3).astype(np.float32)
    out_scale = np.maximum((f - c).reshape(36, -1).std(1), 1e-6).astype(np.float32)
    with ewpt.Lattice(SHAPE, 0.25) as lc, ewpt.Lattice(SHAPE, 0.25) as lf, ewpt.Corrector(seed=1) as net:
        lc.set_state(c)
        lf.set_state(f)
        net.params = P
        net.set_norm(in_scale, out_scale)
        loss = net.train_step(lc, lf, 0.0)
        g = net.grads().astype(np.float64)
        assert np.array_equal(net.params, P)            # lr = 0 leaves the parameters
    inv = (np.float32(1.0) / in_scale).astype(np.float64)
    X0 = np.zeros(c.shape[1:] + (64,))
    X0[..., :36] = np.moveaxis(c * inv[:, None, None, None], 0, -1)
    X0 = net_np.bf16(X0)
    T = np.moveaxis((f - c) / out_scale.astype(np.float64)[:, None, None, None], 0, -1)
    lref, gref = net_np.loss_and_grads(P.astype(np.float64), X0, T, emulate=True)
    assert loss == pytest.approx(lref, rel=1e-4)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    cos = np.dot(g, gref) / (np.linalg.norm(g) * np.linalg.norm(gref))
    assert rel < 2e-2 and cos > 0.999, (rel, cos)
    # per layer too (a wrong layer would hide in the total)
    for l in range(3):
        o = l * (net_np.NW + net_np.NC)
        sl = slice(o, o + net_np.NW + net_np.NC)
        assert np.linalg.norm(g[sl] - gref[sl]) < 3e-2 * np.linalg.norm(gref[sl]), l


@pytest.mark.parametrize("shape", [SHAPE, (16, 12, 8)])   # 8x8x4 brick kernel; 8x4x4 kernel + apply_corr
def test_apply_matches_oracle(shape):
    """y + sout * net(y / in_scale) on the device (the last layer adds the
    correction to the FP64 state in its epilogue where the lattice tiles
    8x8x4) against the oracle's forward pass with the same bf16 roundings."""
    c, _ = _pair(shape, 9)
    P = net_np.init_params(3).astype(np.float32)
    rng = np.random.default_rng(4)
    o = 2 * (net_np.NW + net_np.NC)
    P[o:] = (0.05 * rng.standard_normal(net_np.NW + net_np.NC)).astype(np.float32)   # a nonzero last layer
    in_scale = (0.5 + rng.random(36)).astype(np.float32)
    out_scale = (1e-3 * (0.5 + rng.random(36))).astype(np.float32)
    with ewpt.Lattice(shape, 0.25) as lat, ewpt.Corrector(seed=1) as net:
        net.params = P
        net.set_norm(in_scale, out_scale)
        net.input_mode = 0
        lat.set_state(c)
        net.apply(lat)
        got = lat.get_state() - c
    inv = (np.float32(1.0) / in_scale).astype(np.float64)
    X0 = np.zeros(c.shape[1:] + (64,))
    X0[..., :36] = np.moveaxis(c * inv[:, None, None, None], 0, -1)
    Z3 = net_np.forward(P.astype(np.float64), net_np.bf16(X0), emulate=True)["Z3"]
    want = np.moveaxis(Z3[..., :36], -1, 0) * out_scale.astype(np.float64)[:, None, None, None]
    assert np.abs(want).max() > 0
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel < 1e-2, rel
    assert np.abs(got - want).max() <= 2.0 ** -6 * np.abs(want).max(), np.abs(got - want).max()


def test_step_corrected_fallback_paths():
    shape = (32, 16, 16)
    p = fields.params(dt=0.025)
    p4 = dict(p, dt=0.1)
    s = fields.smooth_state(shape, seed=3, amp=0.2)
    with ewpt.Corrector(seed=2) as net:
        rng = np.random.default_rng(0)
        P = net.params
        P[2 * (net_np.NW + net_np.NC):] = 1e-3 * rng.standard_normal(net_np.NW + net_np.NC)  # nonzero last layer
        net.params = P
        net.set_norm(np.ones(36), 1e-4 * np.ones(36))
        with ewpt.Lattice(shape, 0.25) as a, ewpt.Lattice(shape, 0.25) as b:
            # tolerance below any drift: every coarse step falls back to 4 fine steps
            a.set_state(s)
            assert net.step(a, p, factor=4, ncoarse=2, gauss_tol=-1.0) == 2
            b.set_state(s)
            b.step(p, 8)
            assert np.array_equal(a.get_state(), b.get_state())
            # no tolerance: coarse step + correction, never a fallback
            a.set_state(s)
            assert net.step(a, p, factor=4, ncoarse=1, gauss_tol=math.inf) == 0
            b.set_state(s)
            b.step(p4, 1)
            net.apply(b)
            got, want = a.get_state(), b.get_state()
            assert np.array_equal(got, want)
            b.set_state(s)
            b.step(p4, 1)
            assert not np.array_equal(got, b.get_state())   # the correction did something


def _checked_reference(lat, net, p, p4, factor, ncoarse, tol):
    """The corrected stepping with the Gauss check evaluated right after each
    correction by the standalone reductions: the decisions ewpt_step_corrected
    must reproduce.  Returns (fallbacks, drifts)."""
    def norms():
        g, e = lat.gauss(p), lat.energies(p)
        return math.sqrt(float(np.sum(g[:4] ** 2))), math.sqrt(2.0 * (e[3] + e[5]))
    g0, _ = norms()
    fb, drifts = 0, []
    for _ in range(ncoarse):
        save = lat.get_state()
        lat.step(p4, 1)
        net.apply(lat)
        g1, e1 = norms()
        drifts.append((g1 - g0) / max(e1, 1e-12))
        if not drifts[-1] <= tol:
            lat.set_state(save)
            lat.step(p, factor)
            g1, _ = norms()
            fb += 1
        g0 = g1
    return fb, drifts


def test_step_corrected_checks_match_reference():
    """Checks taken from the fused stage-1 reductions of the next coarse step
    give the fallbacks and states of checking each corrected state at once:
    all accepted, and a tolerance between the drifts that sends some steps
    back to fine steps."""
    shape = (32, 16, 16)
    p = fields.params(dt=0.025)
    p4 = dict(p, dt=0.1)
    s = fields.smooth_state(shape, seed=5, amp=0.3)
    ncoarse = 6
    with ewpt.Corrector(seed=2) as net:
        rng = np.random.default_rng(1)
        P = net.params
        P[2 * (net_np.NW + net_np.NC):] = 3e-3 * rng.standard_normal(net_np.NW + net_np.NC)
        net.params = P
        net.set_norm(np.ones(36), 1e-3 * np.ones(36))
        with ewpt.Lattice(shape, 0.25) as a, ewpt.Lattice(shape, 0.25) as b:
            b.set_state(s)
            fb, drifts = _checked_reference(b, net, p, p4, 4, ncoarse, 1e30)
            assert fb == 0
            a.set_state(s)
            assert net.step(a, p, factor=4, ncoarse=ncoarse, gauss_tol=1e30) == 0
            assert np.array_equal(a.get_state(), b.get_state())
            # a tolerance well inside the spread of the drifts
            d = np.sort(np.asarray(drifts))
            gaps = np.diff(d)
            k = int(np.argmax(gaps))
            assert gaps[k] > 1e-6 * np.abs(d).max()
            tol = 0.5 * (d[k] + d[k + 1])
            b.set_state(s)
            fb, drifts2 = _checked_reference(b, net, p, p4, 4, ncoarse, tol)
            assert fb >= 1
            assert min(abs(x - tol) for x in drifts2) > 1e-9 * np.abs(d).max()
            a.set_state(s)
            assert net.step(a, p, factor=4, ncoarse=ncoarse, gauss_tol=tol) == fb
            assert np.array_equal(a.get_state(), b.get_state())


def test_trained_corrector_reduces_coarse_error():
    """Trained on the GPU for 1500 pairs at 64^3, the corrector shrinks the
    error of eight 4*dt steps against 32 dt steps on held-out thermal states
    (early and late in the evolution); the untrained corrector (zero last
    layer) leaves the coarse result unchanged."""
    n, dx = 64, 0.25
    P = ewpt.Parameters(N=n, dx=dx, gy=0.35)
    p = P.step_params()
    p4 = P.step_params()
    p4.dt = 4 * p.dt
    rng = np.random.default_rng(0)
    lats = [ewpt.Lattice(n, dx) for _ in range(5)]
    src, co, fi, ref, c2 = lats
    print("corrector training 64^3 x 1500", flush=True)
    try:
        with ewpt.Corrector(seed=0) as net:
            net.input_mode = 1
            cs, rs = [], []
            for k in range(4):
                src.init_thermal(float(rng.uniform(0.3, 0.7)), MEFF, 500 + k, "config")
                src.step(p, 100 * k)
                ewpt.make_pair(src, co, fi, p, 4)
                c, f = co.get_state(), fi.get_state()
                cs.append(net.features(c).reshape(36, -1).std(1))
                rs.append((f - c).reshape(36, -1).std(1))
            net.set_norm(np.maximum(np.mean(cs, 0), 1e-12), np.maximum(np.mean(rs, 0), 1e-30))
            # untrained: zero output layer, no change
            ewpt.make_pair(src, co, fi, p, 4)
            before = co.get_state()
            net.apply(co)
            assert np.array_equal(co.get_state(), before)
            losses = []
            for it in range(1500):
                if it % 100 == 0:
                    src.init_thermal(float(rng.uniform(0.3, 0.7)), MEFF, 1000 + it, "config")
                    src.step(p, int(rng.integers(0, 400)))
                ewpt.make_pair(src, co, fi, p, 4)
                losses.append(net.train_step(co, fi, 3e-4))
                src.step(p, int(rng.integers(1, 5)))
            assert np.mean(losses[-200:]) < 0.8 * np.mean(losses[:50])
            ratios = []
            for seed, burn in ((99991, 10), (99993, 400)):
                src.init_thermal(0.5, MEFF, seed, "config")
                src.step(p, burn)
                ref.copy_from(src)
                ref.step(p, 32)
                co.copy_from(src)
                co.step(p4, 8)
                c2.copy_from(src)
                for _ in range(8):
                    c2.step(p4, 1)
                    net.apply(c2)
                R = ref.get_state()
                ratios.append(np.linalg.norm(c2.get_state() - R) / np.linalg.norm(co.get_state() - R))
            print("corrected / coarse error:", ratios, flush=True)
            assert max(ratios) < 0.85
    finally:
        for x in lats:
            x.close()
