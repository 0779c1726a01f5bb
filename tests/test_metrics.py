"""The image-error measures of ptamd.metrics (CPU): relMSE and its trimmed
form, the stratified pixel subset and pixel lookup."""
import numpy as np
import pytest

from ptamd.metrics import accum_at, relmse, relmse_trimmed, stratified_pixels


def test_relmse_closed_forms():
    ref = np.full((10, 3), 2.0)
    img = ref + 1.0
    assert relmse(img, ref) == pytest.approx(1.0 / (4.0 + 0.01))
    assert relmse(ref, ref) == 0.0
    with pytest.raises(ValueError):
        relmse(img[:5], ref)


def test_trimmed_relmse_drops_only_the_largest_pixels():
    rng = np.random.default_rng(1)
    ref = rng.uniform(0.2, 1.0, (5000, 3))
    img = ref * (1 + rng.normal(0, 0.05, ref.shape))
    base = relmse(img, ref)
    # no outliers: trimming 0.1 % (5 pixels) changes the mean by little
    assert relmse_trimmed(img, ref) == pytest.approx(base, rel=0.05)
    # five planted fireflies decide the full relMSE, not the trimmed one
    hot = img.copy()
    hot[rng.choice(len(hot), 5, replace=False)] = 1e4
    assert relmse(hot, ref) > 1e4 * base
    assert relmse_trimmed(hot, ref) == pytest.approx(relmse_trimmed(img, ref), rel=0.05)
    # trimming keeps the smallest errors: exact on a known ladder
    e = np.arange(1000, dtype=np.float64)  # per-pixel errors 0..999 at R = 0, eps = 1
    lad = np.sqrt(e)[:, None].repeat(3, 1)
    assert relmse_trimmed(lad, np.zeros_like(lad), trim=0.1, eps=1.0) == pytest.approx(np.arange(900).mean())
    with pytest.raises(ValueError):
        relmse_trimmed(img[:1], ref[:1])


def test_stratified_pixels_one_per_block():
    W, H = 100, 37
    pix = stratified_pixels(W, H)
    x, y = pix % W, pix // W
    blocks = set(zip((x // 8).tolist(), (y // 8).tolist()))
    assert len(blocks) == len(pix) and np.all(np.diff(pix.astype(np.int64)) > 0)
    assert len(pix) <= ((W + 7) // 8) * ((H + 7) // 8)
    acc = np.arange(W * H * 3, dtype=np.float64).reshape(H, W, 3)
    got = accum_at(acc, pix, W, H)
    assert np.array_equal(got, acc[H - 1 - y, x])
