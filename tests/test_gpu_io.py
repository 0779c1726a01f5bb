"""pt_save_image: the accumulation average written as PNG (Save_PNG) and
Radiance HDR (Save_HDR), decoded here from the file formats and compared
with the same mapping of pt_read_accum's values."""
import struct
import sys
import zlib

import numpy as np
import pytest

import ptamd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, W = 8, b"", None
    while p < len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        kind, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        assert struct.unpack(">I", data[p + 8 + n:p + 12 + n])[0] == zlib.crc32(kind + body) & 0xFFFFFFFF
        if kind == b"IHDR":
            W, H, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 2)
        elif kind == b"IDAT":
            idat += body
        p += 12 + n
    raw = zlib.decompress(idat)
    rows = [raw[y * (3 * W + 1):(y + 1) * (3 * W + 1)] for y in range(H)]
    assert all(r[0] == 0 for r in rows)
    return np.frombuffer(b"".join(r[1:] for r in rows), np.uint8).reshape(H, W, 3)


def read_flat_rgbe(path):
    data = open(path, "rb").read()
    head, _, rest = data.partition(b"\n\n")
    assert head.startswith(b"#?RADIANCE")
    res, _, pix = rest.partition(b"\n")
    _, H, _, W = res.split()
    H, W = int(H), int(W)
    return np.frombuffer(pix, np.uint8).reshape(H, W, 4)


def test_save_png_and_hdr(tmp_path):
    W, H, frames = 48, 40, 6
    sc = ptamd.Scene(ptamd.PT_SCENE_CORNELL, W, H)
    with ptamd.Context(0) as ctx:
        ctx.set_scene(sc.desc)
        ctx.build_accel()
        ctx.set_camera(sc.camera)
        say("[gpu] render and save cornell 48x40")
        ctx.render(sc.render_params(), 0, frames)
        acc, n = ctx.read_accum(W, H)
        ctx.save_image(tmp_path / "out.png")
        ctx.save_image(tmp_path / "out.hdr")
    avg = acc.astype(np.float32) / np.float32(n)
    want_png = (np.float32(255.99) * np.clip(np.sqrt(avg), 0, 1)).astype(np.int32).astype(np.uint8)
    got_png = read_png(tmp_path / "out.png")
    np.testing.assert_array_equal(got_png, want_png)
    # flat RGBE of c / (1 + c): mantissas (c * frexp(m) * 256 / m) truncated
    c = avg / (np.float32(1) + avg)
    m = c.max(2)
    fr, e = np.frexp(m)
    s = np.where(m > 1e-32, fr.astype(np.float32) * np.float32(256) / np.where(m > 0, m, 1), 0).astype(np.float32)
    want = np.zeros((H, W, 4), np.uint8)
    want[..., :3] = (c * s[..., None]).astype(np.uint8)
    want[..., 3] = np.where(m > 1e-32, e + 128, 0)
    got = read_flat_rgbe(tmp_path / "out.hdr")
    np.testing.assert_array_equal(got, want)
    assert got_png.mean() > 10
