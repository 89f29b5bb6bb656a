"""Image decoding oracle: test infrastructure only.

Only tests/ may import this module, as the checker of the device decode; the
product never calls it.  It restates, independently of the product's C++ and
HIP code, what decoding an encoded image must produce:

* png_decode: a PNG to unpremultiplied RGBA8, with Python's zlib as the
  inflater, the PNG specification's five row filters and Adam7 passes, and
  libpng's transformations as DecodePng applies them
  (decoders/src/decode_png.cpp:87-116: png_set_expand for palette, low-depth
  gray and tRNS; png_set_strip_16; png_set_gray_to_rgb).
* jpeg_reconstruct: the samples of a sequential JPEG from its quantized
  coefficients in exact arithmetic: every component's cosine series
  (T.81 A.3.3) evaluated at the component's output density (8 samples per
  block side, or 16 where libjpeg 7+ upsamples chroma through the transform,
  jdmaster.c), rounded; integer replication for the rest of the sampling
  ratio (jdsample.c); and YCbCr -> RGB with libjpeg 9's fixed-point tables
  (jdcolor.c build_ycc_rgb_table).  libjpeg itself computes the 8-point
  transform in 13-bit fixed point (jidctint.c), which may differ from the
  exact value by one step; the tests say how far the device may be.
* premultiply: simd::div255 rounding, as Bitmap::pixelFormat(RGBAPremul)
  (decoders/src/bitmap_decoder.cpp:66-90).
* mip: the next mip level as the rounded mean of each 2x2 footprint.
"""
import zlib

import numpy as np

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunks(data):
    pos = 8
    while pos + 12 <= len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        yield data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        pos += 12 + n


def _unfilter(raw, h, row_bytes, bpp):
    out = np.zeros((h, row_bytes), np.int64)
    prev = np.zeros(row_bytes, np.int64)
    pos = 0
    for y in range(h):
        f = raw[pos]
        x = np.frombuffer(raw, np.uint8, row_bytes, pos + 1).astype(np.int64)
        pos += 1 + row_bytes
        r = np.zeros(row_bytes, np.int64)
        for i in range(row_bytes):
            a = r[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            r[i] = (x[i] + p) & 0xFF
        out[y] = r
        prev = r
    return out


def _samples(rows, w, ch, depth):
    h = rows.shape[0]
    if depth == 16:
        b = rows[:, :w * ch * 2].reshape(h, w * ch, 2)
        return (b[..., 0] << 8 | b[..., 1]).reshape(h, w, ch)
    if depth == 8:
        return rows[:, :w * ch].reshape(h, w, ch)
    per = 8 // depth
    out = np.zeros((h, w * ch), np.int64)
    for i in range(w * ch):
        byte = rows[:, i // per]
        out[:, i] = (byte >> (8 - depth * (i % per + 1))) & ((1 << depth) - 1)
    return out.reshape(h, w, ch)


def png_decode(data):
    data = bytes(data)
    ihdr = next(body for kind, body in _chunks(data) if kind == b"IHDR")
    w, h = int.from_bytes(ihdr[0:4], "big"), int.from_bytes(ihdr[4:8], "big")
    depth, ctype, interlace = ihdr[8], ihdr[9], ihdr[12]
    ch = CHANNELS[ctype]
    idat = b"".join(body for kind, body in _chunks(data) if kind == b"IDAT")
    plte = next((body for kind, body in _chunks(data) if kind == b"PLTE"), b"")
    trns = next((body for kind, body in _chunks(data) if kind == b"tRNS"), None)
    raw = zlib.decompress(idat)
    bpp = max(1, ch * depth // 8)
    passes = [(0, 0, 1, 1)] if interlace == 0 else ADAM7
    samples = np.zeros((h, w, ch), np.int64)
    pos = 0
    for x0, y0, dx, dy in passes:
        pw = (w - x0 + dx - 1) // dx if w > x0 else 0
        ph = (h - y0 + dy - 1) // dy if h > y0 else 0
        if pw == 0 or ph == 0:
            continue
        row_bytes = (pw * ch * depth + 7) // 8
        rows = _unfilter(raw[pos:pos + ph * (row_bytes + 1)], ph, row_bytes, bpp)
        pos += ph * (row_bytes + 1)
        samples[y0::dy, x0::dx][:ph, :pw] = _samples(rows, pw, ch, depth)
    out = np.zeros((h, w, 4), np.int64)

    def to8(v):
        if depth == 16:
            return v >> 8
        return v * (255 // ((1 << depth) - 1))

    if ctype == 3:
        pal = np.frombuffer(plte, np.uint8).reshape(-1, 3).astype(np.int64)
        alpha = np.full(256, 255, np.int64)
        if trns is not None:
            alpha[:len(trns)] = np.frombuffer(trns, np.uint8)
        idx = samples[..., 0]
        out[..., :3] = pal[idx]
        out[..., 3] = alpha[idx]
    elif ctype in (0, 4):
        g = to8(samples[..., 0])
        out[..., 0] = out[..., 1] = out[..., 2] = g
        if ctype == 4:
            out[..., 3] = to8(samples[..., 1])
        else:
            key = int.from_bytes(trns[0:2], "big") if trns is not None else -1
            out[..., 3] = np.where(samples[..., 0] == key, 0, 255)
    else:
        out[..., :3] = to8(samples[..., :3])
        if ctype == 6:
            out[..., 3] = to8(samples[..., 3])
        else:
            if trns is not None:
                key = [int.from_bytes(trns[2 * i:2 * i + 2], "big") for i in range(3)]
                hit = (samples[..., 0] == key[0]) & (samples[..., 1] == key[1]) & (samples[..., 2] == key[2])
                out[..., 3] = np.where(hit, 0, 255)
            else:
                out[..., 3] = 255
    return out.astype(np.uint8)


def premultiply(rgba):
    p = rgba.astype(np.int64)
    a = p[..., 3:4]
    x = p[..., :3] * a + 128
    out = p.copy()
    out[..., :3] = (x + (x >> 8)) >> 8
    return out.astype(np.uint8)


def mip(level):
    h, w = level.shape[:2]
    dh, dw = max(h // 2, 1), max(w // 2, 1)
    y0 = np.minimum(2 * np.arange(dh), h - 1)
    y1 = np.minimum(2 * np.arange(dh) + 1, h - 1)
    x0 = np.minimum(2 * np.arange(dw), w - 1)
    x1 = np.minimum(2 * np.arange(dw) + 1, w - 1)
    l = level.astype(np.int64)
    s = l[y0][:, x0] + l[y0][:, x1] + l[y1][:, x0] + l[y1][:, x1]
    return ((s + 2) >> 2).astype(np.uint8)


def _basis(n):
    """(n, 8) matrix of C(u) cos((2x+1) u pi / (2n)) / 2."""
    x = np.arange(n)[:, None]
    u = np.arange(8)[None, :]
    c = np.where(u == 0, np.sqrt(0.5), 1.0)
    return 0.5 * c * np.cos((2 * x + 1) * u * np.pi / (2 * n))


def jpeg_planes(info, coef):
    """Each component's samples (exact transform, rounded), at its output
    density before replication."""
    planes = []
    for k in range(info.component_count):
        c = info.components[k]
        q = np.array(info.quant[c.quant_table], np.float64).reshape(8, 8)
        blocks = coef[c.coef_offset:c.coef_offset + c.blocks_w * c.blocks_h * 64].reshape(
            c.blocks_h, c.blocks_w, 8, 8).astype(np.float64) * q
        nh, nv = 8 * c.scale_h, 8 * c.scale_v
        bv, bh = _basis(nv), _basis(nh)
        s = np.einsum("yv,abvu,xu->abyx", bv, blocks, bh)
        s = np.clip(np.floor(s + 128.5), 0, 255)
        planes.append(s.transpose(0, 2, 1, 3).reshape(c.blocks_h * nv, c.blocks_w * nh))
    return planes


def jpeg_color(info, planes):
    """RGB (h, w, 3) uint8 from component planes, libjpeg 9's conversion."""
    h, w = info.height, info.width
    comps = []
    for k in range(info.component_count):
        c = info.components[k]
        ys = np.arange(h) // c.rep_v
        xs = np.arange(w) // c.rep_h
        comps.append(planes[k][ys][:, xs].astype(np.int64))
    if info.component_count == 1:
        r = g = b = comps[0]
    elif not info.ycc:
        r, g, b = comps
    else:
        y, cb, cr = comps[0], comps[1] - 128, comps[2] - 128
        r = y + ((91881 * cr + 32768) >> 16)
        b = y + ((116130 * cb + 32768) >> 16)
        g = y + ((-46802 * cr - 22553 * cb + 32768) >> 16)
    return np.clip(np.stack([r, g, b], -1), 0, 255).astype(np.uint8)


def jpeg_reconstruct(info, coef):
    return jpeg_color(info, jpeg_planes(info, coef))
