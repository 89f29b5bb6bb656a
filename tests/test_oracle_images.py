"""Pins the oracle's image paints (drawImage as the unit rect under
scale(width, height), rive_renderer.cpp:383-446; PaintAuxData image matrix and
LOD, gpu.cpp:1006-1034; find_paint_color image modulation, draw_path.vert:
479-493) with closed forms: a pixel-aligned draw reproduces the texels, a
quarter-scale draw samples mip level 1 (the rounded 2x2 means) halfway between
texels, nearest filtering picks texels, and opacity scales the result.
"""
import numpy as np
import pytest

import oracle_binding as ob
from streams import StreamBuilder, rect


def render(b):
    s = b.build()
    f = ob.OracleFlush(s)
    try:
        img, _ = f.render()
    finally:
        f.close()
    return img


def random_image(rng, h, w, opaque=True):
    rgb = rng.integers(0, 256, (h, w, 3))
    a = np.full((h, w, 1), 255) if opaque else rng.integers(0, 256, (h, w, 1))
    rgb = np.minimum(rgb, a)  # premultiplied: channels never exceed alpha
    return np.concatenate([rgb, a], axis=2).astype(np.uint8)


def mip1(img):
    h, w = img.shape[:2]
    t = img.astype(np.int64)
    s = t[0:h:2, 0:w:2] + t[1:h:2, 0:w:2] + t[0:h:2, 1:w:2] + t[1:h:2, 1:w:2]
    return (s + 2) >> 2


@pytest.mark.parametrize("sampler", [0, 9])  # bilinear, nearest
def test_pixel_aligned_image_reproduces_texels(sampler):
    rng = np.random.default_rng(0)
    tex = random_image(rng, 8, 16)
    b = StreamBuilder(24, 12)
    k = b.image(tex)
    b.draw_image(k, matrix=(1, 0, 0, 1, 3, 2), sampler=sampler)
    out = render(b)
    assert np.array_equal(out[2:10, 3:19], tex)
    assert out[0:2].max() == 0 and out[:, 0:3].max() == 0


def test_quarter_scale_samples_mip_level_1():
    rng = np.random.default_rng(1)
    tex = random_image(rng, 16, 16)
    b = StreamBuilder(8, 8)
    b.draw_image(b.image(tex), matrix=(.25, 0, 0, .25, 0, 0))
    out = render(b).astype(float)
    m1 = mip1(tex).astype(float)  # 8x8
    # pixel x samples level 1 at 2x + 0.5: the mean of texels 2x and 2x + 1
    want = (m1[0::2, 0::2] + m1[0::2, 1::2] + m1[1::2, 0::2] + m1[1::2, 1::2]) / 4
    assert np.abs(out[:4, :4] - want).max() <= 1.0


def test_opacity_scales_premultiplied_texels():
    rng = np.random.default_rng(2)
    tex = random_image(rng, 8, 8, opaque=False)
    b = StreamBuilder(8, 8)
    b.draw_image(b.image(tex), opacity=0.5)
    out = render(b).astype(float)
    a = round(255 * .5) / 255
    assert np.abs(out - tex.astype(float) * a).max() <= 1.0


def test_image_under_clip_and_blend():
    rng = np.random.default_rng(3)
    tex = random_image(rng, 8, 8)
    b = StreamBuilder(8, 8)
    b.draw(b.path(rect(0, 0, 8, 8)), color=0xff808080)
    clip = b.path(rect(0, 0, 4, 8))
    b.draw_image(b.image(tex), blend=24, clips=(clip,))  # multiply
    out = render(b).astype(float) / 255
    gray = 0x80 / 255
    assert np.abs(out[:, :4, :3] - tex[:, :4, :3] / 255 * gray).max() <= 1.0 / 255 + 1e-6
    assert np.abs(out[:, 4:, :3] - gray).max() <= 1e-6
