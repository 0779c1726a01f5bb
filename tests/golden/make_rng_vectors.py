"""Writes tests/golden/rng_vectors.json: known answers of the per-path RNG.

An independent restatement in numpy uint32 arithmetic, written from the
published algorithms rather than from the oracle:
  * TEA (Wheeler & Needham, "TEA, a Tiny Encryption Algorithm", FSE 1994):
    N cycles of  sum += 0x9E3779B9;
                 v0 += ((v1 << 4) + k0) ^ (v1 + sum) ^ ((v1 >> 5) + k1);
                 v1 += ((v0 << 4) + k2) ^ (v0 + sum) ^ ((v0 >> 5) + k3).
    The path tracer's seed (programs/random.cuh:33-47) is v0 after 64 cycles
    with the fixed key (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E).
  * The Numerical Recipes LCG (a = 1664525, c = 1013904223, mod 2^32);
    random.cuh:50-55 keeps the low 24 bits and random.cuh:63-65 divides by
    2^24.

The restatement is itself pinned before any vector is written:
  * standard 32-cycle TEA with the all-zero key and block encrypts
    (0, 0) to (0x41EA3A0A, 0x94BAA940), the published test vector;
  * the LCG from state 0 yields 1013904223, 1196435762, 3519870697,
    2868466484, the sequence printed in Numerical Recipes.

Run:  python tests/golden/make_rng_vectors.py
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

M32 = np.uint64(0xFFFFFFFF)
KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)


def tea_np(v0, v1, cycles, key):
    """TEA encryption of blocks (v0, v1) (uint32 arrays); returns (v0, v1)."""
    v0 = np.asarray(v0, np.uint64) & M32
    v1 = np.asarray(v1, np.uint64) & M32
    k0, k1, k2, k3 = (np.uint64(k) for k in key)
    s = np.uint64(0)
    for _ in range(cycles):
        s = (s + np.uint64(0x9E3779B9)) & M32
        v0 = (v0 + ((((v1 << np.uint64(4)) & M32) + k0) ^ ((v1 + s) & M32) ^ ((v1 >> np.uint64(5)) + k1))) & M32
        v1 = (v1 + ((((v0 << np.uint64(4)) & M32) + k2) ^ ((v0 + s) & M32) ^ ((v0 >> np.uint64(5)) + k3))) & M32
    return v0.astype(np.uint32), v1.astype(np.uint32)


def lcg_np(state, n):
    """n LCG steps from `state`; returns (states after each step, final)."""
    out = []
    s = int(state)
    for _ in range(n):
        s = (1664525 * s + 1013904223) & 0xFFFFFFFF
        out.append(s)
    return out, s


def self_check():
    a, b = tea_np([0], [0], 32, (0, 0, 0, 0))
    assert (int(a[0]), int(b[0])) == (0x41EA3A0A, 0x94BAA940), "TEA restatement fails the published vector"
    seq, _ = lcg_np(0, 4)
    assert seq == [1013904223, 1196435762, 3519870697, 2868466484], "LCG restatement fails the NR sequence"


def main():
    self_check()
    rng = np.random.default_rng(20240601)
    # (pixel index W*y+x, frame) pairs: corners of a 512x512 and a 1920x1080
    # image, a few frames, and random draws
    pairs = [(0, 0), (1, 0), (0, 1), (511, 0), (512 * 512 - 1, 63), (1920 * 1080 - 1, 255),
             (400 * 267 - 1, 15), (0xFFFFFFFF, 0xFFFFFFFF)]
    while len(pairs) < 32:
        pairs.append((int(rng.integers(0, 1920 * 1080)), int(rng.integers(0, 1024))))
    idx = np.array([p[0] for p in pairs], np.uint64)
    frm = np.array([p[1] for p in pairs], np.uint64)
    seeds, _ = tea_np(idx, frm, 64, KEY)
    vectors = []
    for (i, f), s in zip(pairs, seeds):
        states, _ = lcg_np(int(s), 16)
        vectors.append({
            "index": i, "frame": f, "seed": int(s),
            "lcg_states": states,
            "rnd": [float(np.float32(st & 0xFFFFFF) / np.float32(16777216.0)) for st in states],
        })
    out = {
        "spec": "seed = TEA-64(index, frame) with key A341316C C8013EA4 AD90777D 7E95761E; "
                "state' = 1664525*state + 1013904223 mod 2^32; rnd = (state' & 0xFFFFFF) / 2^24",
        "tea_known_answer": {"key": [0, 0, 0, 0], "cycles": 32, "block": [0, 0],
                             "cipher": [0x41EA3A0A, 0x94BAA940]},
        "vectors": vectors,
    }
    path = Path(__file__).resolve().parent / "rng_vectors.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path} ({len(vectors)} vectors)")


if __name__ == "__main__":
    main()
