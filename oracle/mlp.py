"""ORACLE (test infrastructure only): numpy restatement of the surrogate network.

MLP 11 -> 64 -> 64 -> 6 with SiLU, in the parameter layout of include/mphip.h
(W1 [64][16], b1, W2 [64][64], b2, W3 [16][64], b3; row-major, out x in).  Two forms:

  * the kernels' arithmetic: operands (normalised features, weights, activations and
    backward signals) rounded to BF16 with round-to-nearest-even before each product,
    products accumulated in FP64 and rounded to FP32 (the hardware accumulates in FP32;
    the difference is at the FP32 rounding level), elementwise work in FP32;
  * exact FP64 math with no rounding, for finite-difference and autograd checks.

The network has no counterpart in the reference; it stands in for the substep loop of
multiphysics_kernel! (src/fdtd/Kernels.jl:362-412).  Features and targets are defined in
include/mphip.h (mp_mlp) and restated in orc_surrogate.h.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 11, 64, 6
OW1, OB1, OW2, OB2, OW3, OB3, NPARAM = 0, 1024, 1088, 5184, 5248, 6272, 6288
N_XM, N_XIS, N_YM, N_YS, N_YIS, NNORM = 0, 16, 32, 48, 64, 80


def unpack(p):
    p = np.asarray(p)
    return (p[OW1:OB1].reshape(64, 16), p[OB1:OW2], p[OW2:OB2].reshape(64, 64), p[OB2:OW3],
            p[OW3:OB3].reshape(16, 64), p[OB3:NPARAM])


def pack(W1, b1, W2, b2, W3, b3, dtype=np.float32):
    out = np.zeros(NPARAM, dtype)
    out[OW1:OB1] = np.asarray(W1).reshape(-1)
    out[OB1:OW2] = b1
    out[OW2:OB2] = np.asarray(W2).reshape(-1)
    out[OB2:OW3] = b2
    out[OW3:OB3] = np.asarray(W3).reshape(-1)
    out[OB3:NPARAM] = b3
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights on the real rows/columns, zero biases and padding."""
    rng = np.random.default_rng(seed)

    def glorot(fo, fi):
        a = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, (fo, fi))

    W1 = np.zeros((64, 16))
    W1[:, :N_IN] = glorot(64, N_IN)
    W2 = glorot(64, 64)
    W3 = np.zeros((16, 64))
    W3[:N_OUT] = glorot(N_OUT, 64)
    return pack(W1, np.zeros(64), W2, np.zeros(64), W3, np.zeros(16))


def make_norm(X, T) -> np.ndarray:
    """Per-feature and per-target mean / scale from sample rows (FP64 statistics)."""
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    xm, xs = X.mean(0), X.std(0)
    ym, ys = T.mean(0), T.std(0)
    norm = np.zeros(NNORM, np.float32)
    norm[N_XM:N_XM + N_IN] = xm
    norm[N_XIS:N_XIS + N_IN] = np.where(xs > 1e-12 * np.maximum(np.abs(xm), 1.0), 1.0 / np.where(xs > 0, xs, 1.0), 0.0)
    ys = np.where(ys > 0, ys, 1.0)
    norm[N_YM:N_YM + N_OUT] = ym
    norm[N_YS:N_YS + N_OUT] = ys
    norm[N_YIS:N_YIS + N_OUT] = 1.0 / ys
    return norm


def bf16(x) -> np.ndarray:
    """Round FP32 values to BF16 (round to nearest, ties to even), returned as FP32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _mm(A, B):
    """Product of BF16-valued FP32 operands, accumulated in FP64, rounded to FP32."""
    return (np.asarray(A, np.float64) @ np.asarray(B, np.float64)).astype(np.float32)


def _sig(z):
    z = np.asarray(z, np.float32)
    with np.errstate(over="ignore"):
        return (np.float32(1.0) / (np.float32(1.0) + np.exp(-z))).astype(np.float32)


def _normalise(params, norm, X):
    X = np.asarray(X, np.float32)
    xm = norm[N_XM:N_XM + N_IN].astype(np.float32)
    xis = norm[N_XIS:N_XIS + N_IN].astype(np.float32)
    xn = np.zeros((X.shape[0], 16), np.float32)
    xn[:, :N_IN] = (X - xm) * xis
    return bf16(xn).T  # (16, B)


def forward(params, norm, X, return_cache=False):
    """Raw rates (B, 6) from raw features (B, 11), in the kernels' arithmetic."""
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float32) for a in unpack(params))
    xb = _normalise(params, norm, X)
    z1 = _mm(bf16(W1), xb) + b1[:, None]
    h1 = bf16(z1 * _sig(z1))
    z2 = _mm(bf16(W2), h1) + b2[:, None]
    h2 = bf16(z2 * _sig(z2))
    y = _mm(bf16(W3), h2) + b3[:, None]
    ys = norm[N_YS:N_YS + N_OUT].astype(np.float32)
    ym = norm[N_YM:N_YM + N_OUT].astype(np.float32)
    rates = np.ascontiguousarray((y[:N_OUT].T * ys + ym).astype(np.float32))
    if return_cache:
        return rates, dict(xb=xb, z1=z1, h1=h1, z2=z2, h2=h2, y=y)
    return rates


def loss_and_grad(params, norm, X, T):
    """Mean squared error over normalised targets and its gradient, in the kernels'
    arithmetic (mp_mlp_backward)."""
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float32) for a in unpack(params))
    B = X.shape[0]
    _, c = forward(params, norm, X, return_cache=True)
    ym = norm[N_YM:N_YM + N_OUT].astype(np.float32)
    yis = norm[N_YIS:N_YIS + N_OUT].astype(np.float32)
    tn = ((np.asarray(T, np.float32) - ym) * yis).T  # (6, B)
    e = (c["y"][:N_OUT] - tn).astype(np.float32)
    loss = float(np.sum(e.astype(np.float64) ** 2) / (B * N_OUT))
    gscale = np.float32(2.0 / (B * N_OUT))
    dy = np.zeros((16, B), np.float32)
    dy[:N_OUT] = e * gscale
    dyb = bf16(dy)
    dh2 = _mm(bf16(W3).T, dyb)
    s2 = _sig(c["z2"])
    dz2 = (dh2 * (s2 * (np.float32(1) + c["z2"] * (np.float32(1) - s2)))).astype(np.float32)
    dz2b = bf16(dz2)
    dh1 = _mm(bf16(W2).T, dz2b)
    s1 = _sig(c["z1"])
    dz1 = (dh1 * (s1 * (np.float32(1) + c["z1"] * (np.float32(1) - s1)))).astype(np.float32)
    dz1b = bf16(dz1)
    g = pack(_mm(dz1b, c["xb"].T), dz1.astype(np.float64).sum(1), _mm(dz2b, c["h1"].T),
             dz2.astype(np.float64).sum(1), _mm(dyb, c["h2"].T), dy.astype(np.float64).sum(1))
    return loss, g


def loss64(params, norm, X, T) -> float:
    """The same loss in exact FP64 math (no BF16 rounding)."""
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float64) for a in unpack(params))
    n = np.asarray(norm, np.float64)
    xn = np.zeros((X.shape[0], 16))
    xn[:, :N_IN] = (np.asarray(X, np.float64) - n[N_XM:N_XM + N_IN]) * n[N_XIS:N_XIS + N_IN]
    silu = lambda z: z / (1.0 + np.exp(-z))
    h1 = silu(W1 @ xn.T + b1[:, None])
    h2 = silu(W2 @ h1 + b2[:, None])
    y = (W3 @ h2 + b3[:, None])[:N_OUT]
    tn = ((np.asarray(T, np.float64) - n[N_YM:N_YM + N_OUT]) * n[N_YIS:N_YIS + N_OUT]).T
    return float(np.mean((y - tn) ** 2))


def grad64(params, norm, X, T) -> np.ndarray:
    """Analytic gradient of loss64 (exact FP64 backpropagation)."""
    W1, b1, W2, b2, W3, b3 = (np.asarray(a, np.float64) for a in unpack(params))
    n = np.asarray(norm, np.float64)
    B = X.shape[0]
    xn = np.zeros((B, 16))
    xn[:, :N_IN] = (np.asarray(X, np.float64) - n[N_XM:N_XM + N_IN]) * n[N_XIS:N_XIS + N_IN]
    x = xn.T
    z1 = W1 @ x + b1[:, None]
    s1 = 1.0 / (1.0 + np.exp(-z1))
    h1 = z1 * s1
    z2 = W2 @ h1 + b2[:, None]
    s2 = 1.0 / (1.0 + np.exp(-z2))
    h2 = z2 * s2
    y = W3 @ h2 + b3[:, None]
    tn = np.zeros((16, B))
    tn[:N_OUT] = ((np.asarray(T, np.float64) - n[N_YM:N_YM + N_OUT]) * n[N_YIS:N_YIS + N_OUT]).T
    dy = np.zeros((16, B))
    dy[:N_OUT] = 2.0 * (y[:N_OUT] - tn[:N_OUT]) / (B * N_OUT)
    dz2 = (W3.T @ dy) * (s2 * (1.0 + z2 * (1.0 - s2)))
    dz1 = (W2.T @ dz2) * (s1 * (1.0 + z1 * (1.0 - s1)))
    return pack(dz1 @ x.T, dz1.sum(1), dz2 @ h1.T, dz2.sum(1), dy @ h2.T, dy.sum(1), np.float64)


def adam(params, grad, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step in FP32, as mp_mlp_adam computes it; returns new (params, m1, m2)."""
    f = np.float32
    g = np.asarray(grad, np.float32)
    a = (f(b1) * m1 + (f(1) - f(b1)) * g).astype(np.float32)
    b = (f(b2) * m2 + (f(1) - f(b2)) * (g * g)).astype(np.float32)
    c1 = f(1.0 - float(f(b1)) ** step)
    c2 = f(1.0 - float(f(b2)) ** step)
    mh, vh = a / c1, b / c2
    p = (params - f(lr) * (mh / (np.sqrt(vh) + f(eps)) + f(wd) * params)).astype(np.float32)
    return p, a, b
