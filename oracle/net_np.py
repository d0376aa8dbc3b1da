"""CPU restatement of the corrector network (test infrastructure only).

Only ``tests/`` use this module, as the checker of the device convolution,
weight-gradient and training kernels (ewpt_net.hip).  No reference
counterpart: the reference has no learned component; the network here is
36 -> 64 -> 64 -> 36 periodic 3x3x3 convolutions with SiLU.

Layout as on the device: activations (nz, ny, nx, 64) channel-last, weights
[27 taps][64 out][64 in] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1) and
out[x] = sum_tap W[tap] . in[x + (dx, dy, dz)] + b (periodic).  FP64
arithmetic; ``bf16`` rounds to bfloat16 (round to nearest even) where the
device stores bfloat16, so a comparison isolates the FP32 accumulation.
"""
from __future__ import annotations

import numpy as np

NC = 64
NCH = 36
TAPS = [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]   # (dx, dy, dz)
NW = 27 * NC * NC
NPAR = 3 * (NW + NC)


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def to_bits(a) -> np.ndarray:
    """bfloat16 bit patterns (uint16) of values already on the bf16 grid."""
    return (np.asarray(a, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(h) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def shift(X, d):
    """X[x + d] for a (nz, ny, nx, C) array, periodic."""
    dx, dy, dz = d
    return np.roll(X, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv(X, W, b=None):
    out = np.zeros(X.shape[:3] + (W.shape[1],))
    for t, d in enumerate(TAPS):
        out += shift(X, d) @ W[t].T
    if b is not None:
        out += b
    return out


def conv_wgrad(dZ, X):
    """dW[tap][out][in] = sum_x dZ[x][out] X[x + tap][in]; db = sum_x dZ."""
    dW = np.stack([dZ.reshape(-1, dZ.shape[-1]).T @ shift(X, d).reshape(-1, X.shape[-1]) for d in TAPS])
    return dW, dZ.reshape(-1, dZ.shape[-1]).sum(0)


def conv_dgrad(dZ, W):
    """dX[y][in] = sum_tap W[tap]^T . dZ[y - tap]."""
    out = np.zeros(dZ.shape[:3] + (W.shape[2],))
    for t, (dx, dy, dz) in enumerate(TAPS):
        out += shift(dZ, (-dx, -dy, -dz)) @ W[t]
    return out


def mirrored(W):
    """The device's input-gradient operand Wt[tap][in][out] = W[26 - tap][out][in]."""
    return np.stack([W[26 - t].T for t in range(27)])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack_params(P):
    P = np.asarray(P, dtype=np.float64)
    out = []
    for l in range(3):
        o = l * (NW + NC)
        out.append((P[o:o + NW].reshape(27, NC, NC), P[o + NW:o + NW + NC]))
    return out


def pack_input(state, in_scale, emulate=True):
    """(36, nz, ny, nx) FP64 state -> (nz, ny, nx, 64) network input."""
    X = np.zeros(state.shape[1:] + (NC,))
    X[..., :NCH] = np.moveaxis(state / np.asarray(in_scale, dtype=np.float64)[:, None, None, None], 0, -1)
    return bf16(X) if emulate else X


def forward(P, X0, emulate=True):
    (W1, b1), (W2, b2), (W3, b3) = unpack_params(P)
    r = bf16 if emulate else (lambda a: a)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    Z1 = conv(X0, W1, b1)
    A1 = r(silu(Z1))
    Z2 = conv(A1, W2, b2)
    A2 = r(silu(Z2))
    Z3 = conv(A2, W3, b3)
    return dict(Z1=Z1, A1=A1, Z2=Z2, A2=A2, Z3=Z3, W=(W1, W2, W3))


def loss_and_grads(P, X0, T, emulate=True):
    """T: (nz, ny, nx, 36) target.  Returns the relative loss
    sum (Z3 - T)^2 / sum T^2 and its gradient vector."""
    f = forward(P, X0, emulate)
    r = bf16 if emulate else (lambda a: a)
    d = f["Z3"][..., :NCH] - T
    tsq = float(np.sum(T * T))
    loss = float(np.sum(d * d)) / tsq
    dZ3 = np.zeros_like(f["Z3"])
    dZ3[..., :NCH] = r(2.0 * d / tsq)
    W1, W2, W3 = f["W"]
    dW3, db3 = conv_wgrad(dZ3, f["A2"])
    dZ2 = r(conv_dgrad(dZ3, W3) * dsilu(f["Z2"]))
    dW2, db2 = conv_wgrad(dZ2, f["A1"])
    dZ1 = r(conv_dgrad(dZ2, W2) * dsilu(f["Z1"]))
    dW1, db1 = conv_wgrad(dZ1, X0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return loss, g


def init_params(seed: int) -> np.ndarray:
    """A random parameter vector with the device's padding (input channels
    >= 36 of layer 1 and output channels >= 36 of layer 3 zero)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPAR)
    for l in range(3):
        o = l * (NW + NC)
        W = rng.uniform(-1, 1, (27, NC, NC)) * np.sqrt(3.0 / (27 * NC))
        if l == 0:
            W[:, :, NCH:] = 0.0
        if l == 2:
            W[:, NCH:, :] = 0.0
        P[o:o + NW] = W.ravel()
        b = rng.uniform(-0.1, 0.1, NC)
        if l == 2:
            b[NCH:] = 0.0
        P[o + NW:o + NW + NC] = b
    return P
