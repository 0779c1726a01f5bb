"""GPU checks of the radiance-cache network kernels (BF16 MFMA forward,
backward through the activations, split-K weight-gradient GEMM, Adam)
against the NumPy restatement in oracle/ptref_mlp.py, through the C ABI.

Tolerances: the oracle rounds operands to BF16 at the same points as the
kernels and accumulates in float64; the kernels accumulate in FP32 in MFMA
order, which moves a BF16 rounding of a hidden activation by one BF16 ulp
(2^-8 relative) now and then.  Outputs: max error <= 2e-2 of the output
scale, mean <= 1e-3; gradients: per-layer relative norm error <= 2e-2 and
cosine >= 0.999; Adam: the closed-form float32 update to 1e-6 relative.
"""
import sys

import numpy as np
import pytest

import ptamd
import ptref_mlp as M

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ptamd.Context(0)
    yield c
    c.close()


def features(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, (n, 96)).astype(np.float32)
    f[:, 83:] = 0.0  # padding of the renderer's 83 features
    return M.to_bf16_bits(f)


def layer_slices():
    out, off = [], 0
    for o, i in M.SIZES:
        out.append(slice(off, off + o * i + o))
        off += o * i + o
    return out


def test_forward_matches_oracle(ctx):
    say("[gpu] cache forward")
    ctx.cache_init(1)
    p = ctx.cache_params()
    f = features(1000, 1)  # not a multiple of 16: tail lanes are masked
    y = ctx.cache_infer(f)
    yr, _ = M.forward(p, M.from_bf16_bits(f), bf16_ops=True)
    err = np.abs(y - yr)
    scale = np.abs(yr).max()
    assert err.max() <= 2e-2 * scale and err.mean() <= 1e-3 * scale, (err.max(), err.mean(), scale)


def test_gradients_match_oracle(ctx):
    say("[gpu] cache gradients")
    ctx.cache_init(2)
    p = ctx.cache_params()
    n = 2500  # padded to 2528: three split-K chunks, a partial last one
    f = features(n, 2)
    t = np.random.default_rng(3).uniform(-1, 2, (n, 3)).astype(np.float32)
    loss, g = ctx.cache_train(f, t, lr=0.0)
    lref, gref = M.loss_and_grad(p, M.from_bf16_bits(f), t, bf16_ops=True)
    assert abs(loss - lref) <= 1e-3 * lref
    for sl in layer_slices():
        a, b = g[sl].astype(np.float64), gref[sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert rel <= 2e-2 and cos >= 0.999, (sl, rel, cos)
    # lr = 0 leaves the parameters unchanged
    np.testing.assert_array_equal(ctx.cache_params(), p)


def test_adam_matches_closed_form(ctx):
    say("[gpu] cache Adam")
    ctx.cache_init(3)
    p0 = ctx.cache_params()
    f = features(512, 4)
    t = np.random.default_rng(5).uniform(0, 1, (512, 3)).astype(np.float32)
    m = np.zeros(M.NPARAM, np.float32)
    v = np.zeros(M.NPARAM, np.float32)
    p = p0
    for step in (1, 2, 3):
        _, g = ctx.cache_train(f, t, lr=1e-3)
        p, m, v = M.adam(p, m, v, g, step, lr=1e-3)
        got = ctx.cache_params()
        np.testing.assert_allclose(got, p, rtol=1e-6, atol=1e-9)
        p = got


def test_training_fits_a_smooth_function(ctx):
    say("[gpu] cache training run")
    ctx.cache_init(4)
    rng = np.random.default_rng(6)
    n = 8192
    x = rng.uniform(-1, 1, (n, 96)).astype(np.float32)
    x[:, 83:] = 0
    t = np.stack([np.sin(3 * x[:, 0]) + x[:, 1], x[:, 2] * x[:, 3], x[:, :8].mean(1)], 1).astype(np.float32)
    f = M.to_bf16_bits(x)
    first, _ = ctx.cache_train(f, t, lr=1e-3)
    for _ in range(300):
        last, _ = ctx.cache_train(f, t, lr=1e-3)
    assert last < 0.1 * first, (first, last)
    y = ctx.cache_infer(f)
    assert np.mean((y - t) ** 2) * 3 < 0.1 * first
