"""Pins the oracle's per-path RNG (programs/random.cuh:33-65) and its point
samplers (programs/sampling.cuh:22-71) against known answers.

rng_vectors.json is written by tests/golden/make_rng_vectors.py, an
independent numpy restatement checked against the published TEA test vector
and the Numerical Recipes LCG sequence.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import ptref

GOLDEN = Path(__file__).resolve().parent / "golden" / "rng_vectors.json"


@pytest.fixture(scope="module")
def vectors():
    return json.loads(GOLDEN.read_text())


def test_tea_seed_known_answers(vectors):
    for v in vectors["vectors"]:
        assert ptref.tea(v["index"] & 0xFFFFFFFF, v["frame"] & 0xFFFFFFFF, 64) == v["seed"], v


def test_rnd_sequence_known_answers(vectors):
    for v in vectors["vectors"]:
        vals, state = ptref.rnd_sequence(v["seed"], len(v["rnd"]))
        assert state == v["lcg_states"][-1]
        np.testing.assert_array_equal(vals, np.array(v["rnd"], np.float32))


def test_rnd_range_and_resolution():
    vals, _ = ptref.rnd_sequence(12345, 100000)
    assert vals.min() >= 0.0 and vals.max() < 1.0
    # every value is a multiple of 2^-24
    assert np.all(vals * np.float32(2 ** 24) == np.floor(vals * np.float32(2 ** 24)))
    assert abs(vals.mean() - 0.5) < 0.005


def test_point_samplers_domains():
    n = 20000
    disk, _ = ptref.random_points("disk", 7, n)
    assert np.all(disk[:, 2] == 0.0)
    assert np.all(np.sum(disk[:, :2].astype(np.float64) ** 2, axis=1) < 1.0)
    ball, _ = ptref.random_points("ball", 8, n)
    r2 = np.sum(ball.astype(np.float64) ** 2, axis=1)
    assert np.all(r2 <= 1.0 + 1e-6)
    # uniform in the ball: P(r < 1/2) = 1/8
    assert abs(np.mean(r2 < 0.25) - 0.125) < 0.01
    sph, _ = ptref.random_points("sphere", 9, n)
    np.testing.assert_allclose(np.linalg.norm(sph.astype(np.float64), axis=1), 1.0, atol=2e-6)
    # uniform on the sphere: each coordinate has mean 0 and variance 1/3
    np.testing.assert_allclose(sph.mean(axis=0), 0.0, atol=0.02)
    np.testing.assert_allclose((sph.astype(np.float64) ** 2).mean(axis=0), 1 / 3, atol=0.01)
