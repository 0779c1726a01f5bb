"""Writers of small asset files for the tests: PNG (every colour type, bit
depths 1..16, each scanline filter), binary PPM, Radiance RGBE with
new-style run-length scanlines, and an OBJ/MTL pair.  They follow the file
format specifications (PNG: ISO/IEC 15948; RGBE: the Radiance picture
format), independently of the library's readers."""
import math
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


# ------------------------------------------------------------------ JPEG
ZIGZAG = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27,
                   20, 13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58,
                   59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])
_C = np.array([[(math.sqrt(0.5) if u == 0 else 1.0) * math.cos((2 * x + 1) * u * math.pi / 16) for x in range(8)]
               for u in range(8)]) / 2.0  # orthonormal DCT-II basis, C[u, x]


def _category(v):
    return 0 if v == 0 else int(abs(int(v))).bit_length()


def _bits_of(v, s):
    return v if v >= 0 else v + (1 << s) - 1


class _BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, code, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((code >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)  # byte stuffing
                self.acc = self.n = 0

    def flush(self):
        while self.n:
            self.put(1, 1)  # pad with 1-bits


def _table(symbols):
    """canonical code of equal-length words for the used symbols (never the
    all-ones word): (BITS[1..16], HUFFVAL, {symbol: (code, length)})"""
    syms = sorted(set(symbols))
    L = max(1, math.ceil(math.log2(len(syms) + 1)))
    bits = [0] * 16
    bits[L - 1] = len(syms)
    return bits, syms, {s: (i, L) for i, s in enumerate(syms)}


def jpeg_bytes(img, sampling=(1, 1), quality_scale=1.0, restart=0, separate_scans=False, seed=0):
    """Baseline JPEG of an (H, W) grey or (H, W, 3) RGB uint8 image per ITU-T
    T.81 (JFIF colour conversion, luma sampling `sampling`, chroma 1x1,
    restart interval `restart` MCUs, interleaved or one scan per component).
    Returns (file bytes, the quantized coefficients per component and block,
    the quantization tables) for a reference decode."""
    img = np.asarray(img, np.float64)
    grey = img.ndim == 2
    H, W = img.shape[:2]
    rng = np.random.default_rng(seed)
    q = [np.clip(np.round(rng.uniform(2, 40, 64) * quality_scale), 1, 255).astype(int) for _ in range(2)]
    if grey:
        planes = [img]
        samp = [sampling]
    else:
        R, G, B = img[..., 0], img[..., 1], img[..., 2]
        planes = [0.299 * R + 0.587 * G + 0.114 * B,
                  -0.168736 * R - 0.331264 * G + 0.5 * B + 128, 0.5 * R - 0.418688 * G - 0.081312 * B + 128]
        samp = [sampling, (1, 1), (1, 1)]
    hmax, vmax = max(s[0] for s in samp), max(s[1] for s in samp)
    mcux, mcuy = -(-W // (8 * hmax)), -(-H // (8 * vmax))
    comps = []
    for c, (pl, (h, v)) in enumerate(zip(planes, samp)):
        # pad to whole MCUs by edge replication, then subsample by averaging
        full = np.pad(pl, ((0, mcuy * 8 * vmax - H), (0, mcux * 8 * hmax - W)), mode="edge")
        fy, fx = vmax // v, hmax // h
        sub = full.reshape(full.shape[0] // fy, fy, full.shape[1] // fx, fx).mean((1, 3))
        bh, bw = sub.shape[0] // 8, sub.shape[1] // 8
        blocks = sub.reshape(bh, 8, bw, 8).transpose(0, 2, 1, 3) - 128.0
        coef = np.einsum("ux,abxy,vy->abuv", _C, blocks, _C)  # F[u (vertical), v (horizontal)]
        tq = q[0 if c == 0 else 1]
        qnat = np.zeros(64)
        qnat[ZIGZAG] = tq
        quant = np.round(coef / qnat.reshape(8, 8)).astype(int)
        comps.append(dict(h=h, v=v, tq=0 if c == 0 else 1, quant=quant, bw=bw, bh=bh))
    # entropy coding: blocks in scan order
    def scan_units(cids):
        units = []
        if len(cids) == 1:
            c = comps[cids[0]]
            cw = -(-(-(-W * c["h"] // hmax)) // 8)
            ch = -(-(-(-H * c["v"] // vmax)) // 8)
            for by in range(ch):
                for bx in range(cw):
                    units.append([(cids[0], by, bx)])
        else:
            for my in range(mcuy):
                for mx in range(mcux):
                    u = []
                    for ci in cids:
                        c = comps[ci]
                        for vv in range(c["v"]):
                            for hh in range(c["h"]):
                                u.append((ci, my * c["v"] + vv, mx * c["h"] + hh))
                    units.append(u)
        return units

    scans = [[i] for i in range(len(comps))] if (separate_scans or grey) else [list(range(len(comps)))]
    # symbols per table class (0 luma, 1 chroma), for the Huffman tables
    dcs, acs = {0: [], 1: []}, {0: [], 1: []}
    streams = []
    for sc in scans:
        units = scan_units(sc)
        events = []
        pred = {ci: 0 for ci in sc}
        for n, u in enumerate(units):
            if restart and n and n % restart == 0:
                events.append(("rst", (n // restart - 1) % 8))
                pred = {ci: 0 for ci in sc}
            for ci, by, bx in u:
                cls = comps[ci]["tq"]
                zz = comps[ci]["quant"][by, bx].reshape(64)[ZIGZAG]
                d = int(zz[0]) - pred[ci]
                pred[ci] = int(zz[0])
                s = _category(d)
                events.append(("dc", cls, s, _bits_of(d, s)))
                dcs[cls].append(s)
                run = 0
                last = max([k for k in range(1, 64) if zz[k] != 0], default=0)
                for k in range(1, last + 1):
                    if zz[k] == 0:
                        run += 1
                        if run == 16:
                            events.append(("ac", cls, 0xF0, 0, 0))
                            acs[cls].append(0xF0)
                            run = 0
                        continue
                    s = _category(zz[k])
                    events.append(("ac", cls, (run << 4) | s, s, _bits_of(int(zz[k]), s)))
                    acs[cls].append((run << 4) | s)
                    run = 0
                if last < 63:
                    events.append(("ac", cls, 0x00, 0, 0))
                    acs[cls].append(0)
        streams.append((sc, events))
    classes = [0] if grey else [0, 1]
    tabs = {}
    out = bytearray(b"\xff\xd8")
    seg = lambda m, body: bytes([0xFF, m]) + struct.pack(">H", len(body) + 2) + bytes(body)
    out += seg(0xDB, b"".join(bytes([t]) + bytes(int(x) for x in q[t]) for t in classes))
    body = bytes([8]) + struct.pack(">HH", H, W) + bytes([len(comps)])
    for i, c in enumerate(comps):
        body += bytes([i + 1, (c["h"] << 4) | c["v"], c["tq"]])
    out += seg(0xC0, body)
    dht = b""
    for t in classes:
        for tc, syms in ((0, dcs[t]), (1, acs[t])):
            bits, vals, codes = _table(syms or [0])
            tabs[(tc, t)] = codes
            dht += bytes([(tc << 4) | t]) + bytes(bits) + bytes(vals)
    out += seg(0xC4, dht)
    if restart:
        out += seg(0xDD, struct.pack(">H", restart))
    for sc, events in streams:
        body = bytes([len(sc)])
        for ci in sc:
            t = comps[ci]["tq"]
            body += bytes([ci + 1, (t << 4) | t])
        body += bytes([0, 63, 0])
        out += seg(0xDA, body)
        bw = _BitWriter()
        for e in events:
            if e[0] == "rst":
                bw.flush()
                bw.out += bytes([0xFF, 0xD0 + e[1]])
                continue
            if e[0] == "dc":
                code, L = tabs[(0, e[1])][e[2]]
                bw.put(code, L)
                if e[2]:
                    bw.put(e[3], e[2])
            else:
                code, L = tabs[(1, e[1])][e[2]]
                bw.put(code, L)
                if e[3]:
                    bw.put(e[4], e[3])
        bw.flush()
        out += bw.out
    out += b"\xff\xd9"
    return bytes(out), comps, q, (hmax, vmax)


def jpeg_reference_decode(H, W, comps, q, hv):
    """Dequantize, inverse DCT (float64), +128, round, clamp; replicate
    chroma; JFIF YCbCr -> RGB, round, clamp: (H, W, 3) uint8."""
    hmax, vmax = hv
    planes = []
    for c in comps:
        qnat = np.zeros(64)
        qnat[ZIGZAG] = q[c["tq"]]
        F = c["quant"] * qnat.reshape(8, 8)
        pix = np.einsum("ux,abuv,vy->abxy", _C, F, _C)
        pix = np.clip(np.floor(pix + 128.0 + 0.5), 0, 255)
        bh, bw = pix.shape[:2]
        plane = pix.transpose(0, 2, 1, 3).reshape(bh * 8, bw * 8)
        ys = (np.arange(H) * c["v"]) // vmax
        xs = (np.arange(W) * c["h"]) // hmax
        planes.append(plane[ys][:, xs])
    if len(planes) == 1:
        return np.repeat(planes[0][..., None], 3, 2).astype(np.uint8)
    Y, cb, cr = planes[0], planes[1] - 128.0, planes[2] - 128.0
    rgb = np.stack([Y + 1.402 * cr, Y - 0.344136 * cb - 0.714136 * cr, Y + 1.772 * cb], 2)
    return np.clip(np.floor(rgb + 0.5), 0, 255).astype(np.uint8)


# ------------------------------------------------------------- TGA, BMP
def tga_bytes(img, rle=False, top_down=False):
    """(H, W) grey or (H, W, 3|4) RGB(A) uint8 as a TGA file"""
    img = np.asarray(img, np.uint8)
    grey = img.ndim == 2
    H, W = img.shape[:2]
    ch = 1 if grey else img.shape[2]
    px = img.reshape(H, W, 1) if grey else img[..., [2, 1, 0, 3][:ch]]
    rows = px if top_down else px[::-1]
    flat = rows.reshape(-1, ch)
    typ = (11 if rle else 3) if grey else (10 if rle else 2)
    head = bytes([0, 0, typ]) + bytes(9) + struct.pack("<HH", W, H) + bytes([8 * ch, 0x20 if top_down else 0])
    if not rle:
        return head + flat.tobytes()
    out = bytearray()
    i = 0
    while i < len(flat):
        j = i
        while j + 1 < len(flat) and j - i < 127 and np.array_equal(flat[j + 1], flat[i]):
            j += 1
        if j > i:
            out += bytes([0x80 | (j - i)]) + flat[i].tobytes()
            i = j + 1
        else:
            k = i
            while k + 1 < len(flat) and k - i < 127 and not np.array_equal(flat[k + 1], flat[k]):
                k += 1
            out += bytes([k - i]) + flat[i:k + 1].tobytes()
            i = k + 1
    return head + bytes(out)


def bmp_bytes(img, top_down=False):
    """(H, W, 3|4) RGB(A) uint8 as an uncompressed BMP"""
    img = np.asarray(img, np.uint8)
    H, W, ch = img.shape
    row = (W * ch + 3) // 4 * 4
    rows = img if top_down else img[::-1]
    data = b"".join(r[:, [2, 1, 0, 3][:ch]].tobytes() + bytes(row - W * ch) for r in rows)
    info = struct.pack("<IiiHHIIiiII", 40, W, -H if top_down else H, 1, 8 * ch, 0, len(data), 2835, 2835, 0, 0)
    return b"BM" + struct.pack("<IHHI", 14 + len(info) + len(data), 0, 0, 14 + len(info)) + info + data
