"""NumPy restatement of the radiance-cache network -- TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference.  It is a 96 -> 64 -> 64 ->
64 -> 64 -> 3 MLP with ReLU hidden layers; parameters are ordered W1 (64 x
96, row-major), b1, W2, b2, W3, b3, W4, b4, W5 (3 x 64), b5 (18,883 values).

`forward` / `loss_and_grad` can round the operands to BF16 at the points the
GPU kernels do (weights, inputs, hidden activations and output gradients),
so the GPU's BF16-MFMA results can be compared with a tolerance that only
has to cover FP32 accumulation order; with `bf16=False` everything stays in
float64, which the finite-difference tests use.
"""
from __future__ import annotations

import numpy as np

SIZES = [(64, 96), (64, 64), (64, 64), (64, 64), (3, 64)]
NPARAM = sum(o * i + o for o, i in SIZES)
assert NPARAM == 18883


def to_bf16_bits(x):
    """float32 -> BF16 bit patterns, round to nearest even (NaN kept quiet)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def from_bf16_bits(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(x):
    return from_bf16_bits(to_bf16_bits(np.asarray(x, np.float32)))


def unpack(params):
    out, off = [], 0
    for o, i in SIZES:
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    return out


def forward(params, feats, bf16_ops=True):
    """feats: (n, 96) float (already BF16 values when bf16_ops).  Returns
    the output (n, 3) and the list of layer inputs [x, h1, .., h4]."""
    layers = unpack(np.asarray(params, np.float64))
    x = np.asarray(feats, np.float64)
    acts = [x]
    for L, (W, b) in enumerate(layers):
        Wq = bf16(W).astype(np.float64) if bf16_ops else W
        z = x @ Wq.T + b
        if L < 4:
            h = np.maximum(z, 0.0)
            if bf16_ops:
                h = bf16(h).astype(np.float64)
            acts.append(h)
            x = h
        else:
            return z, acts
    raise AssertionError


Y_MAX = 20.0        # outputs stand for radiance expm1(min(y, Y_MAX))
REL_EPS = 0.01      # relative-loss denominator floor


def radiance(y):
    """Radiance an output stands for when it answers a query: expm1 of the
    output clamped at Y_MAX, negative values and NaN as 0."""
    r = np.expm1(np.minimum(np.asarray(y, np.float64), Y_MAX))
    return np.where(r > 0, r, 0.0)


def rel_l2_terms(y, t, den=None):
    """Relative L2 loss (Mueller et al. 2021, neural radiance caching) of
    outputs y against radiance targets t, elementwise: with
    yh = expm1(min(y, Y_MAX)), loss = (yh - t)^2 / den, den = yh^2 + REL_EPS
    held constant, so d loss / d y = 2 (yh - t) (yh + 1) / den (taken at the
    clamp above Y_MAX).  `den` may be given (a denominator frozen elsewhere).
    Returns (loss, dloss/dy, den)."""
    y = np.asarray(y, np.float64)
    yh = np.expm1(np.where(y > Y_MAX, Y_MAX, y))
    if den is None:
        den = yh * yh + REL_EPS
    d = yh - np.asarray(t, np.float64)
    return d * d / den, 2.0 * d * (yh + 1.0) / den, den


def loss_and_grad(params, feats, targets, bf16_ops=True, den=None):
    """Mean over samples of the relative L2 loss summed over the outputs,
    and its gradient with respect to every parameter, the denominators held
    constant (the GPU's rounding points when bf16_ops).  With `den` given the
    denominators are those values (finite differences of the loss then see
    the same constant the gradient does)."""
    params = np.asarray(params, np.float64)
    layers = unpack(params)
    n = len(feats)
    y, acts = forward(params, feats, bf16_ops)
    lterm, dterm, _ = rel_l2_terms(y, targets, den)
    loss = float(np.sum(lterm) / n)
    dz = dterm / n
    grads = []
    for L in range(4, -1, -1):
        if bf16_ops:
            dz = bf16(dz).astype(np.float64)
        W, _ = layers[L]
        grads.append((dz.T @ acts[L], dz.sum(axis=0)))
        if L == 0:
            break
        Wq = bf16(W).astype(np.float64) if bf16_ops else W
        dh = dz @ Wq
        dz = dh * (acts[L] > 0)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam(params, m, v, grad, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, ICLR 2015) in float32, as the GPU does it."""
    f = np.float32
    g = np.asarray(grad, f)
    m = f(beta1) * m + (f(1) - f(beta1)) * g
    v = f(beta2) * v + (f(1) - f(beta2)) * (g * g)
    bc1 = f(1) - f(beta1) ** f(step)
    bc2 = f(1) - f(beta2) ** f(step)
    p = params - f(lr) * (m / bc1) / (np.sqrt(v / bc2) + f(eps))
    return p.astype(f), m.astype(f), v.astype(f)
