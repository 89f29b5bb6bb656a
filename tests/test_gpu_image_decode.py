"""Device image decoding (rive_hip_image_decode): the decoded, premultiplied
image and its mip chain in HBM.

PNG: level 0 equals the oracle's independent decode, premultiplied with the
reference's div255 rounding (Bitmap::pixelFormat, bitmap_decoder.cpp:66-90),
bit for bit, for the reference's test images and a synthetic image with
partial alpha; every further level is the rounded 2x2 mean of the one above.

JPEG: level 0 is within 3/255 per channel of the oracle's exact-arithmetic
reconstruction from the same coefficients (and within 1/255 on at least 99%
of channel values): the device computes full-resolution components with
libjpeg's 13-bit fixed-point transform, which may differ from the exact value
by one step before colour conversion spreads it.  Synthetic images decode to
within the quantization error of their source (PSNR >= 30 dB at quality 92),
for 4:2:0, 4:2:2, 4:4:0, 4:4:4, gray, RGB-coded and progressive files.  Images beyond the
16384-texel limit (the reference's bad.jpg) and WebP are refused with
RIVE_HIP_ERROR_UNSUPPORTED; the reference's bad.png is refused as its
DecodePng refuses it.
"""
import os
import sys

import numpy as np
import pytest

import jpeg_test_encoder as enc
import rivehip
from test_image_decode import asset, encode_png

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
import image_oracle as io  # noqa: E402

pytestmark = pytest.mark.gpu


def decode(ctx, data):
    h = ctx.image_decode(data)
    try:
        levels = []
        lvl = 0
        while True:
            try:
                levels.append(ctx.image_read(h, lvl))
            except RuntimeError:
                break
            lvl += 1
        return levels
    finally:
        ctx.free_image(h)


def check_mips(levels):
    for a, b in zip(levels[:-1], levels[1:]):
        assert np.array_equal(io.mip(a), b)
    assert levels[-1].shape[:2] == (1, 1)


@pytest.mark.parametrize("name", ["placeholder.png", "batdude.png", "nomoon.png"])
def test_png_assets_on_device(hip_ctx, name):
    print(f"[image] PNG {name}", flush=True)
    data = asset(name)
    levels = decode(hip_ctx, data)
    assert np.array_equal(levels[0], io.premultiply(io.png_decode(data)))
    check_mips(levels)


def test_png_partial_alpha_on_device(hip_ctx):
    print("[image] PNG 61x45 RGBA16 interlaced", flush=True)
    rng = np.random.default_rng(11)
    samples = rng.integers(0, 65536, (45, 61, 4))
    data = encode_png(samples, 6, 16, interlace=1, rng=rng)
    levels = decode(hip_ctx, data)
    want = io.premultiply(io.png_decode(data))
    assert np.array_equal(levels[0], want)
    assert (want[..., 3] < 255).mean() > 0.9
    check_mips(levels)


def psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 10 * np.log10(255.0 ** 2 / max(mse, 1e-12))


def check_jpeg(levels, data):
    info, coef = rivehip.jpeg_decode_info(data)
    want = io.jpeg_reconstruct(info, coef)
    got = levels[0]
    assert got.shape == (info.height, info.width, 4)
    assert (got[..., 3] == 255).all()
    d = np.abs(got[..., :3].astype(np.int32) - want.astype(np.int32))
    print(f"[image]   vs exact reconstruction: max {int(d.max())}, within 1: {(d <= 1).mean():.5f}", flush=True)
    assert int(d.max()) <= 3 and (d <= 1).mean() >= 0.99
    check_mips(levels)
    return got


JPEG_CASES = [dict(sampling=(2, 2)), dict(sampling=(2, 1)), dict(sampling=(1, 2)), dict(sampling=(1, 1)),
              dict(sampling=(2, 2), restart=4), dict(gray=True), dict(sampling=(2, 2), progressive=True),
              dict(sampling=(1, 1), ids=[ord("R"), ord("G"), ord("B")], jfif=False)]


@pytest.mark.parametrize("case", JPEG_CASES, ids=[str(c) for c in JPEG_CASES])
def test_jpeg_synthetic_on_device(hip_ctx, case):
    case = dict(case)
    gray = case.pop("gray", False)
    print(f"[image] JPEG 101x67 {case} gray={gray}", flush=True)
    # a smooth image with an edge: within JPEG's reach at quality 92
    yy, xx = np.mgrid[0:67, 0:101]
    rgb = np.stack([128 + 100 * np.sin(xx / 9.0), 128 + 90 * np.cos(yy / 7.0), 60 + (xx + yy)], -1)
    rgb[20:40, 30:70] = [230, 40, 50]
    img = np.clip(rgb, 0, 255).astype(np.uint8)
    if gray:
        img = img[..., 1]
    data, _ = enc.encode(img, quality=92, **case)
    got = check_jpeg(decode(hip_ctx, data), data)
    src = np.repeat(img[..., None], 3, -1) if gray else img
    p = psnr(got[..., :3], src)
    print(f"[image]   PSNR vs source {p:.1f} dB", flush=True)
    assert p >= 30.0


def test_jpeg_reference_asset_on_device(hip_ctx):
    print("[image] JPEG open_source.jpg 350x200 4:2:0", flush=True)
    data = asset("open_source.jpg")
    got = check_jpeg(decode(hip_ctx, data), data)
    assert got.shape[:2] == (200, 350) and got[..., :3].std() > 10


def test_refused_images(hip_ctx):
    print("[image] refusals: bad.jpg (24566x58278), bad.png, WebP", flush=True)
    L = rivehip.lib()
    for name in ["bad.jpg", "bad_head.png"]:
        data = asset(name)
        with pytest.raises(RuntimeError):
            hip_ctx.image_decode(data)
    err = L.rive_hip_last_error().decode()
    assert "image" in err
    data = asset("bad.jpg")
    import ctypes
    h = ctypes.c_void_p()
    assert L.rive_hip_image_decode(hip_ctx.h, data, len(data), ctypes.byref(h)) == 4  # UNSUPPORTED
    assert "16384" in L.rive_hip_last_error().decode()
    webp = b"RIFF\x24\x00\x00\x00WEBPVP8 " + bytes(32)
    assert L.rive_hip_image_decode(hip_ctx.h, webp, len(webp), ctypes.byref(h)) == 4
