"""Pins of the corrector network's CPU restatement (oracle/net_np.py).

The convolution is checked against scipy.ndimage.correlate (periodic), the
mirrored-weight input gradient against the adjoint identity
<dgrad(dZ), X> = <dZ, conv(X)>, and the whole backward pass against central
finite differences of the FP64 loss; a planted transposed weight index is
shown to fail the adjoint check.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import net_np


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((4, 6, 8, 3))
    W = rng.standard_normal((27, 2, 3))
    out = net_np.conv(X, W)
    for o in range(2):
        want = np.zeros(X.shape[:3])
        for i in range(3):
            k = np.zeros((3, 3, 3))
            for t, (dx, dy, dz) in enumerate(net_np.TAPS):
                k[dz + 1, dy + 1, dx + 1] = W[t, o, i]
            want += ndimage.correlate(X[..., i], k, mode="wrap")
        np.testing.assert_allclose(out[..., o], want, atol=1e-12)


def test_dgrad_is_the_adjoint():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((4, 4, 8, 5))
    dZ = rng.standard_normal((4, 4, 8, 6))
    W = rng.standard_normal((27, 6, 5))
    lhs = np.sum(net_np.conv_dgrad(dZ, W) * X)
    rhs = np.sum(dZ * net_np.conv(X, W))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # the device's mirrored operand used as forward weights gives the same
    np.testing.assert_allclose(net_np.conv(dZ, net_np.mirrored(W)), net_np.conv_dgrad(dZ, W), atol=1e-12)
    # planted mistake: transposing the tap's matrix without mirroring the tap
    bad = np.stack([W[t].T for t in range(27)])
    assert abs(np.sum(net_np.conv(dZ, bad) * X) - rhs) > 1e-3 * abs(rhs)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    P = net_np.init_params(3)
    X0 = np.zeros((4, 4, 8, net_np.NC))
    X0[..., :net_np.NCH] = rng.standard_normal((4, 4, 8, net_np.NCH))
    T = 0.1 * rng.standard_normal((4, 4, 8, net_np.NCH))
    loss, g = net_np.loss_and_grads(P, X0, T, emulate=False)
    idx = list(rng.choice(net_np.NPAR, 24, replace=False))
    # make sure every layer's weights and biases are probed
    for l in range(3):
        o = l * (net_np.NW + net_np.NC)
        idx += [o + 5 * 64 + 7, o + net_np.NW + 3]
    for q in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += h
        Pm[q] -= h
        lp, _ = net_np.loss_and_grads(Pp, X0, T, emulate=False)
        lm, _ = net_np.loss_and_grads(Pm, X0, T, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert g[q] == pytest.approx(fd, rel=1e-5, abs=1e-11), q


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e-5])
    r = net_np.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0                       # tie to even
    assert r[2] == 1.0 + 2.0 ** -6                           # tie rounds up to even mantissa
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-5) <= 2.0 ** -9 * 3.0e-5
    np.testing.assert_array_equal(net_np.from_bits(net_np.to_bits(r)), r)
