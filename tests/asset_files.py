"""Writers of small asset files for the tests: PNG (every colour type, bit
depths 1..16, each scanline filter), binary PPM, Radiance RGBE with
new-style run-length scanlines, and an OBJ/MTL pair.  They follow the file
format specifications (PNG: ISO/IEC 15948; RGBE: the Radiance picture
format), independently of the library's readers."""
import struct
import zlib

import numpy as np


def _chunk(kind, data):
    body = kind + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_row(ftype, row, prev, bpp):
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ftype]
        out[i] = (x - pred) & 0xFF
    return bytes([ftype]) + bytes(out)


def png_bytes(samples, color_type, depth, palette=None, trns=None, filters=(0, 1, 2, 3, 4)):
    """samples: (H, W, channels) integer array of raw sample values (palette
    indices for colour type 3); rows use the filter types in turn."""
    samples = np.asarray(samples)
    H, W = samples.shape[:2]
    ch = samples.shape[2]
    rows = []
    for y in range(H):
        if depth == 16:
            row = b"".join(struct.pack(">H", int(v)) for v in samples[y].reshape(-1))
        elif depth == 8:
            row = bytes(int(v) for v in samples[y].reshape(-1))
        else:
            bits = "".join(format(int(v), f"0{depth}b") for v in samples[y].reshape(-1))
            bits += "0" * (-len(bits) % 8)
            row = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
        rows.append(row)
    bpp = max(1, ch * depth // 8)
    raw = b""
    prev = bytes(len(rows[0]))
    for y, row in enumerate(rows):
        raw += _filter_row(filters[y % len(filters)], row, prev, bpp)
        prev = row
    out = b"\x89PNG\r\n\x1a\n"
    out += _chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, color_type, 0, 0, 0))
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    comp = zlib.compress(raw, 9)
    out += _chunk(b"IDAT", comp[: len(comp) // 2]) + _chunk(b"IDAT", comp[len(comp) // 2:])  # split stream
    out += _chunk(b"IEND", b"")
    return out


def ppm_bytes(rgb8, comment=True):
    H, W, _ = rgb8.shape
    head = b"P6\n" + (b"# test image\n" if comment else b"") + f"{W} {H}\n255\n".encode()
    return head + bytes(np.asarray(rgb8, np.uint8).reshape(-1))


def rgbe_bytes(rgbe):
    """(H, W, 4) uint8 RGBE pixels as a Radiance file with new-style
    run-length scanlines (W in [8, 32767]): each channel of a scanline in
    runs (count > 128: one value repeated count - 128 times) and dumps."""
    H, W, _ = rgbe.shape
    out = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {H} +X {W}\n".encode()
    for y in range(H):
        out += bytes([2, 2, W >> 8, W & 0xFF])
        for c in range(4):
            vals = [int(v) for v in rgbe[y, :, c]]
            i = 0
            while i < W:
                j = i
                while j < W and j - i < 127 and vals[j] == vals[i]:
                    j += 1
                if j - i >= 3:
                    out += bytes([128 + (j - i), vals[i]])
                    i = j
                else:
                    k = i
                    while k < W and k - i < 128 and not (k + 2 < W and vals[k] == vals[k + 1] == vals[k + 2]):
                        k += 1
                    out += bytes([k - i]) + bytes(vals[i:k])
                    i = k
    return out


def rgbe_to_float(rgbe):
    """Radiance decoding as lib/HDRloader.cpp:99-116 does it: (mantissa / 256)
    * 2^(exponent - 128) (exact in float32 for every byte pair)."""
    e = rgbe[..., 3].astype(np.int32)
    scale = np.ldexp(1.0, e - 128) / 256.0
    return (rgbe[..., :3].astype(np.float64) * scale[..., None]).astype(np.float32)
