"""Image-error measurement and workload cameras shared by the tests and the
benchmark.

relMSE is the error measure of the render path's benchmark: the mean over
pixels and colour channels of (I - R)^2 / (R^2 + 0.01), R a converged
reference.  Converged references are evaluated on a stratified pixel subset:
one pixel per 8 x 8 block (1/64 of the image) at a hashed position inside
the block, so the subset is fixed by the image size alone.
"""
from __future__ import annotations

import math

import numpy as np

RELMSE_EPS = 0.01


def _mix32(x):
    """lowbias32 integer hash (C. Wellons), vectorised over uint32."""
    x = np.asarray(x, np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def stratified_pixels(W: int, H: int, block: int = 8, salt: int = 0x5EED) -> np.ndarray:
    """Launch indices (x + W*y) of one pixel per block x block tile, at a
    hashed offset inside the tile; tiles whose chosen pixel falls outside the
    image are skipped.  Sorted ascending."""
    bx, by = np.meshgrid(np.arange((W + block - 1) // block), np.arange((H + block - 1) // block))
    bx, by = bx.ravel().astype(np.uint64), by.ravel().astype(np.uint64)
    h = _mix32((by * 65536 + bx) ^ np.uint64(salt))
    ox, oy = h % block, (h >> 8) % block
    x, y = bx * block + ox, by * block + oy
    keep = (x < W) & (y < H)
    return np.sort((x[keep] + np.uint64(W) * y[keep]).astype(np.uint32))


def relmse(img: np.ndarray, ref: np.ndarray, eps: float = RELMSE_EPS) -> float:
    """mean over pixels and channels of (I - R)^2 / (R^2 + eps)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    if img.shape != ref.shape:
        raise ValueError(f"relmse: shapes {img.shape} and {ref.shape} differ")
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))


def relmse_trimmed(img: np.ndarray, ref: np.ndarray, trim: float = 1e-3, eps: float = RELMSE_EPS) -> float:
    """relMSE without the pixels of largest error: per pixel the mean over
    channels of (I - R)^2 / (R^2 + eps), the ceil(trim * n) largest dropped,
    the rest averaged.  A scene lit by a small, very bright source (a sun
    lobe of peak 5e4) gives rare samples of huge value (caustics through
    glass); a handful of them decides the plain relMSE at any sample count,
    the trimmed one follows the bulk of the image."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    if img.shape != ref.shape:
        raise ValueError(f"relmse_trimmed: shapes {img.shape} and {ref.shape} differ")
    e = ((img - ref) ** 2 / (ref * ref + eps)).reshape(len(img), -1).mean(1)
    k = int(np.ceil(trim * len(e)))
    if k >= len(e):
        raise ValueError("relmse_trimmed: nothing left after trimming")
    e = np.sort(e)[: len(e) - k]
    return float(e.mean())


def accum_at(accum: np.ndarray, pix: np.ndarray, W: int, H: int) -> np.ndarray:
    """Rows of an accumulation image (H x W x 3, row 0 = top, as
    pt_read_accum returns it) at launch indices pix (x + W*y, y up)."""
    pix = np.asarray(pix, np.int64)
    x, y = pix % W, pix // W
    return accum[H - 1 - y, x]


def orbit_camera(info, frame: int, frames_per_turn: int = 300, r: float = 20.0, height: float = 4.0,
                 vfov: float = 45.0, focus: float = 10.0):
    """Camera of the radiance-cache workload: an orbit of radius r around the
    origin at the given height, one turn per `frames_per_turn` frames (frame
    0 at (r, height, 0)), looking at the origin, pinhole."""
    from . import Camera

    a = 2 * math.pi * frame / frames_per_turn
    o = np.array([r * math.cos(a), height, r * math.sin(a)])
    w = o / np.linalg.norm(o)
    u = np.cross([0.0, 1.0, 0.0], w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    hh = math.tan(math.radians(vfov) / 2)
    hw = info.width / info.height * hh
    c = Camera()
    c.origin[:] = o.tolist()
    c.lower_left[:] = (o - hw * focus * u - hh * focus * v - focus * w).tolist()
    c.horizontal[:] = (2 * hw * focus * u).tolist()
    c.vertical[:] = (2 * hh * focus * v).tolist()
    c.u[:] = u.tolist()
    c.v[:] = v.tolist()
    c.w[:] = w.tolist()
    c.lens_radius, c.time0, c.time1 = 0.0, 0.0, 1.0
    return c
