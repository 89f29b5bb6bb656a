"""Host side of image decoding (no GPU): PNG to RGBA8 and JPEG entropy
decoding, against independent restatements and the reference's own assets.

PNG: the reference's test images (tests/unit_tests/assets, see
tests/golden/make_image_fixtures.py) decode bit-exactly as the oracle's
independent decoder (Python's zlib, the specification's filters and Adam7,
libpng's transformations) decodes them; synthetic PNGs of every colour type
and bit depth, interlaced or not, with tRNS, random row filters, several IDAT
chunks and every zlib block type decode to their source pixels; damaged and
truncated files fail, and the reference's oversized bad.png fails as its
DecodePng does (image_decoders_test.cpp:49-70).

JPEG: images written by a test encoder (tests/jpeg_test_encoder.py) give back
exactly the quantized coefficients the encoder wrote, for 4:2:0, 4:2:2, 4:4:4
and gray, interleaved or one component per scan, with restart intervals, and
for progressive files (libjpeg's simple progression: DC and AC successive
approximation, spectral bands, end-of-band runs); a
file cut short leaves the rest of its blocks zero (libjpeg's behaviour for
insufficient data); the reference's JPEGs report the dimensions its test
asserts (image_decoders_test.cpp:21-47).
"""
import os
import struct
import zlib

import numpy as np
import pytest

import jpeg_test_encoder as enc
import rivehip

HERE = os.path.dirname(os.path.abspath(__file__))
IMAGES = os.path.join(HERE, "golden", "images")
import sys  # noqa: E402

sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import image_oracle as io  # noqa: E402


def asset(name):
    with open(os.path.join(IMAGES, name), "rb") as f:
        return f.read()


# ---------------------------------------------------------------- PNG
def chunk(kind, body):
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def pack_row(samples, depth):
    """One row of samples (ints) packed at the bit depth."""
    if depth == 16:
        return b"".join(int(v).to_bytes(2, "big") for v in samples)
    if depth == 8:
        return bytes(int(v) for v in samples)
    per = 8 // depth
    out = bytearray((len(samples) + per - 1) // per)
    for i, v in enumerate(samples):
        out[i // per] |= int(v) << (8 - depth * (i % per + 1))
    return bytes(out)


def filter_row(row, prev, bpp, f):
    out = bytearray([f])
    for i in range(len(row)):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i] if prev is not None else 0
        c = prev[i - bpp] if (prev is not None and i >= bpp) else 0
        p = [0, a, b, (a + b) // 2, None][f]
        if f == 4:
            pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
            p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
        out.append((row[i] - p) & 0xFF)
    return bytes(out)


def encode_png(samples, ctype, depth, interlace=0, palette=None, trns=None, level=6, rng=None, idat_parts=3):
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    passes = [(0, 0, 1, 1)] if interlace == 0 else io.ADAM7
    raw = bytearray()
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prev = None
        for r in sub:
            row = pack_row(r.reshape(-1), depth)
            raw += filter_row(row, prev, bpp, int(rng.integers(0, 5)))
            prev = row
    z = zlib.compress(bytes(raw), level)
    cut = np.linspace(0, len(z), idat_parts + 1).astype(int)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if palette is not None:
        out += chunk(b"PLTE", bytes(palette.reshape(-1).astype(np.uint8)))
    if trns is not None:
        out += chunk(b"tRNS", trns)
    out += chunk(b"tEXt", b"Comment\x00made by the decoder tests")
    for a, b in zip(cut[:-1], cut[1:]):
        out += chunk(b"IDAT", z[a:b])
    return out + chunk(b"IEND", b"")


def expected_rgba(samples, ctype, depth, palette=None, trns=None):
    """RGBA8 per libpng's expand / strip_16 / gray_to_rgb, from source samples."""
    s = samples.astype(np.int64)
    to8 = (lambda v: v >> 8) if depth == 16 else (lambda v: v * (255 // ((1 << depth) - 1)))
    h, w = s.shape[:2]
    out = np.zeros((h, w, 4), np.int64)
    if ctype == 3:
        alpha = np.full(256, 255)
        if trns is not None:
            alpha[:len(trns)] = list(trns)
        out[..., :3] = palette[s[..., 0]]
        out[..., 3] = alpha[s[..., 0]]
    elif ctype in (0, 4):
        out[..., :3] = to8(s[..., :1])
        out[..., 3] = to8(s[..., 1]) if ctype == 4 else 255
        if ctype == 0 and trns is not None:
            out[..., 3] = np.where(s[..., 0] == int.from_bytes(trns, "big"), 0, 255)
    else:
        out[..., :3] = to8(s[..., :3])
        out[..., 3] = to8(s[..., 3]) if ctype == 6 else 255
        if ctype == 2 and trns is not None:
            key = [int.from_bytes(trns[2 * i:2 * i + 2], "big") for i in range(3)]
            out[..., 3] = np.where((s[..., 0] == key[0]) & (s[..., 1] == key[1]) & (s[..., 2] == key[2]), 0, 255)
    return out.astype(np.uint8)


PNG_CASES = [(0, d) for d in (1, 2, 4, 8, 16)] + [(2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8), (4, 8), (4, 16),
                                                   (6, 8), (6, 16)]


@pytest.mark.parametrize("ctype,depth", PNG_CASES)
@pytest.mark.parametrize("interlace", [0, 1])
def test_png_synthetic(ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth * 2 + interlace)
    for (w, h), level in [((1, 1), 0), ((7, 3), 1), ((33, 17), 9)]:
        ch = io.CHANNELS[ctype]
        top = (1 << depth) - 1
        palette = trns = None
        if ctype == 3:
            n = min(top + 1, 40)
            samples = rng.integers(0, n, (h, w, 1))
            palette = rng.integers(0, 256, (n, 3))
            trns = bytes(rng.integers(0, 256, n // 2).astype(np.uint8))
        else:
            samples = rng.integers(0, top + 1, (h, w, ch))
            if ctype == 0:
                key = int(samples[0, 0, 0])
                trns = key.to_bytes(2, "big")
            elif ctype == 2:
                trns = b"".join(int(v).to_bytes(2, "big") for v in samples[-1, -1])
        data = encode_png(samples, ctype, depth, interlace, palette, trns, level, rng)
        got = rivehip.png_decode(data)
        want = expected_rgba(samples, ctype, depth, palette, trns)
        assert got.shape == (h, w, 4)
        assert np.array_equal(got, want), f"ctype {ctype} depth {depth} {w}x{h} interlace {interlace}"
        assert np.array_equal(io.png_decode(data), want)


@pytest.mark.parametrize("name,size", [("placeholder.png", (226, 128)), ("batdude.png", (442, 412)),
                                       ("nomoon.png", (319, 320))])
def test_png_reference_assets(name, size):
    data = asset(name)
    assert rivehip.image_probe(data) == ("png", *size)
    got = rivehip.png_decode(data)
    assert got.shape == (size[1], size[0], 4)
    assert np.array_equal(got, io.png_decode(data)), name


def test_png_damage_is_refused():
    rng = np.random.default_rng(7)
    samples = rng.integers(0, 256, (9, 11, 4))
    good = encode_png(samples, 6, 8, rng=rng)
    assert np.array_equal(rivehip.png_decode(good), expected_rgba(samples, 6, 8))
    # a flipped byte inside IDAT: CRC mismatch of a critical chunk
    i = good.index(b"IDAT") + 6
    bad = good[:i] + bytes([good[i] ^ 0x40]) + good[i + 1:]
    with pytest.raises(RuntimeError):
        rivehip.png_decode(bad)
    # the image data cut short (chunk lengths consistent, stream truncated)
    raw = zlib.compress(b"\x00" * 10, 6)
    short = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 64, 64, 8, 6, 0, 0, 0)) +
             chunk(b"IDAT", raw) + chunk(b"IEND", b""))
    with pytest.raises(RuntimeError):
        rivehip.png_decode(short)
    # an unknown row filter
    rows = b"".join(b"\x07" + bytes(12) for _ in range(3))
    badf = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 3, 3, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(rows)) + chunk(b"IEND", b""))
    with pytest.raises(RuntimeError):
        rivehip.png_decode(badf)


def test_png_reference_bad_png():
    # image_decoders_test.cpp "bad png file doesn't cause an overflow": a
    # 58278x24566 header over far too little data decodes to nothing.
    data = asset("bad_head.png")
    assert rivehip.image_probe(data) == ("png", 58278, 24566)
    with pytest.raises(RuntimeError):
        rivehip.png_decode(data)


# ---------------------------------------------------------------- JPEG
def test_image_probe_formats():
    assert rivehip.image_probe(asset("open_source.jpg")) == ("jpeg", 350, 200)
    # image_decoders_test.cpp "bad jpeg file doesn't cause an overflow"
    assert rivehip.image_probe(asset("bad.jpg")) == ("jpeg", 24566, 58278)
    webp = b"RIFF\x24\x00\x00\x00WEBPVP8 " + bytes(32)
    with pytest.raises(RuntimeError, match="WebP"):
        rivehip.image_probe(webp)
    with pytest.raises(RuntimeError):
        rivehip.image_probe(b"GIF89a" + bytes(20))


def _flat(written):
    return np.concatenate([w.reshape(-1) for w in written])


JPEG_CASES = [
    dict(sampling=(2, 2)), dict(sampling=(2, 1)), dict(sampling=(1, 2)), dict(sampling=(1, 1)),
    dict(sampling=(2, 2), restart=3), dict(sampling=(2, 1), separate_scans=True),
    dict(sampling=(1, 1), ids=[ord("R"), ord("G"), ord("B")], jfif=False),
    dict(sampling=(2, 2), jfif=False, adobe=1), dict(sampling=(1, 1), jfif=False, adobe=0, ids=[0, 1, 2]),
    dict(gray=True), dict(gray=True, restart=1),
]


@pytest.mark.parametrize("case", JPEG_CASES, ids=[str(c) for c in JPEG_CASES])
def test_jpeg_coefficients_round_trip(case):
    case = dict(case)
    rng = np.random.default_rng(len(str(case)))
    gray = case.pop("gray", False)
    for w, h in [(8, 8), (37, 21), (64, 48)]:
        base = rng.integers(0, 256, (h, w) if gray else (h, w, 3))
        img = np.clip(base + np.linspace(0, 60, w)[None, :, None if not gray else None].reshape(
            (1, w) if gray else (1, w, 1)), 0, 255).astype(np.uint8)
        data, written = enc.encode(img, quality=80, **case)
        info, coef = rivehip.jpeg_decode_info(data)
        assert (info.width, info.height) == (w, h)
        assert info.component_count == (1 if gray else 3)
        # libjpeg 9's colour-space rule (jdapimin.c): component IDs first,
        # then JFIF, then the Adobe transform flag
        rgb = case.get("ids") == [ord("R"), ord("G"), ord("B")] or case.get("adobe") == 0
        assert gray or info.ycc == (0 if rgb else 1)
        assert np.array_equal(coef, _flat(written)), f"{case} {w}x{h}"


@pytest.mark.parametrize("sampling,restart", [((2, 2), 0), ((2, 1), 3), ((1, 1), 0), (None, 0), (None, 5)])
def test_jpeg_progressive_round_trip(sampling, restart):
    # libjpeg's simple progression: DC first/refinement, spectral bands and
    # successive-approximation AC refinement (T.81 G.1.2), on noise (short
    # end-of-band runs) and on a smooth image (runs of many blocks)
    rng = np.random.default_rng(17)
    gray = sampling is None
    for w, h, smooth in [(37, 21, False), (64, 48, False), (160, 120, True)]:
        if smooth:
            yy, xx = np.mgrid[0:h, 0:w]
            base = np.stack([128 + 60 * np.sin(xx / 23.0), 128 + 50 * np.cos(yy / 17.0), 90 + 0 * xx], -1)
        else:
            base = rng.integers(0, 256, (h, w, 3))
        img = np.clip(base, 0, 255).astype(np.uint8)
        if gray:
            img = img[..., 0]
        kw = {} if gray else {"sampling": sampling}
        data, written = enc.encode(img, quality=50 if smooth else 85, progressive=True, restart=restart, **kw)
        info, coef = rivehip.jpeg_decode_info(data)
        assert np.array_equal(coef, _flat(written)), f"{sampling} {w}x{h} restart {restart}"


def test_jpeg_truncated_leaves_zero_blocks():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (64, 64, 3)).astype(np.uint8)
    data, written = enc.encode(img, quality=95, sampling=(1, 1))
    full_info, full = rivehip.jpeg_decode_info(data)
    cut = data[:len(data) // 2]
    info, coef = rivehip.jpeg_decode_info(cut)
    c0 = info.components[0]
    n = c0.blocks_w * c0.blocks_h * 64
    y_full, y_cut = full[:n].reshape(-1, 64), coef[:n].reshape(-1, 64)
    same = np.all(y_full == y_cut, axis=1)
    zero = np.all(y_cut == 0, axis=1)
    # a prefix of the blocks decodes as before, the tail is all zero
    k = int(np.argmin(same)) if not same.all() else len(same)
    assert 0 < k < len(same)
    assert zero[k + 1:].all()


def test_jpeg_reference_asset_layout():
    info, coef = rivehip.jpeg_decode_info(asset("open_source.jpg"))
    assert (info.width, info.height, info.component_count, info.ycc) == (350, 200, 3, 1)
    comps = [(info.components[k].h, info.components[k].v, info.components[k].scale_h,
              info.components[k].scale_v, info.components[k].rep_h) for k in range(3)]
    assert comps == [(2, 2, 1, 1, 1), (1, 1, 2, 2, 1), (1, 1, 2, 2, 1)]
    assert int(info.coef_count) == coef.size and np.abs(coef).max() > 0


def test_jpeg_oracle_transform_is_the_dct_inverse():
    # The oracle's 8-point reconstruction inverts the orthonormal DCT (the
    # encoder's) exactly, and its 16-point evaluation is the same cosine
    # series at twice the density: its even/odd samples straddle the 8-point
    # samples (a linear ramp block reproduces the ramp at half steps).
    d = enc.dct_matrix()
    rng = np.random.default_rng(5)
    blk = rng.uniform(-100, 100, (8, 8))
    coef = d @ blk @ d.T
    b8 = io._basis(8)
    assert np.allclose(b8 @ coef @ b8.T, blk)
    ramp = np.tile(np.arange(8) * 8.0 - 28.0, (8, 1))
    c = d @ ramp @ d.T
    b16 = io._basis(16)
    up = b16 @ c @ b16.T
    # the series of a ramp, sampled at x = (n - 0.5) / 2 in 8-sample units
    assert np.allclose(up[:, 4:12], np.tile((np.arange(4, 12) - 0.5) / 2 * 8.0 - 28.0, (16, 1)), atol=1.5)
