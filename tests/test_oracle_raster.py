"""Pins the oracle's rasteriser with closed forms on hand-built streams.

The reference draws fills as a midpoint fan plus an anti-aliasing border
outset by aa = 0.5 * manhattan_pixel_width (draw_path_common.glsl:268-273,
654-773): for a straight edge with unit normal n the coverage at a pixel
centre at signed distance d (positive inside) is clamp(0.5 + d/(|nx|+|ny|),
0, 1).  Strokes use the same ramp on both sides of the centre line.  Fill
rules (draw_raster_order_path.frag:62-82), clips (:84-152), premultiplied
src-over, the two-stop linear ramp (color_ramp.glsl:40-106) and the dither
bound (common.glsl:269-274, +-1/512) are checked the same way.  Results are
RGBA8, so every closed form is compared within 1 code value.
"""
import math

import numpy as np
import pytest

import oracle_binding as ob
from streams import (BUTT, CLOCKWISE, EVENODD, NONZERO, StreamBuilder, rect)


def render(b):
    s = b.build()
    f = ob.OracleFlush(s)
    try:
        img, _ = f.render()
    finally:
        f.close()
    return img


def halfplane_poly(nx, ny, c0, lo=-20.0, hi=90.0):
    """Polygon of the box [lo,hi]^2 cut to nx*x + ny*y < c0."""
    box = [(lo, lo), (hi, lo), (hi, hi), (lo, hi)]
    inside = lambda p: nx * p[0] + ny * p[1] < c0
    poly = []
    for i in range(4):
        a, b = box[i], box[(i + 1) % 4]
        if inside(a):
            poly.append(a)
        if inside(a) != inside(b):
            fa, fb = nx * a[0] + ny * a[1] - c0, nx * b[0] + ny * b[1] - c0
            t = fa / (fa - fb)
            poly.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
    return [(0,) + tuple(poly[0])] + [(1,) + tuple(q) for q in poly[1:]] + [(5,)]


@pytest.mark.parametrize("deg,c0", [(0, 30.25), (90, 17.5), (30, 30.0), (45, 40.3), (63, 25.7),
                                    (180 + 20, -20.0)])
def test_fill_edge_coverage_ramp(deg, c0):
    a = math.radians(deg)
    nx, ny = math.cos(a), math.sin(a)
    b = StreamBuilder(48, 48)
    b.draw(b.path(halfplane_poly(nx, ny, c0)), color=0xffffffff)
    img = render(b)
    ys, xs = np.mgrid[0:48, 0:48] + 0.5
    d = c0 - (nx * xs + ny * ys)
    cov = np.clip(0.5 + d / (abs(nx) + abs(ny)), 0, 1)
    assert np.abs(img[..., 3].astype(float) - 255 * cov).max() <= 1.0


@pytest.mark.parametrize("deg,c0,far", [(45, 40.3, 1e6), (30, 30.0, 3e6), (200, -20.0, 5e5)])
def test_fill_edge_coverage_ramp_from_far_vertices(deg, c0, far):
    """The same half-plane ramp when the edge runs between vertices up to
    millions of pixels off the canvas (exact integer predicates on the
    unclipped triangles, no coordinate clamp below 2^30 subpixels)."""
    a = math.radians(deg)
    nx, ny = math.cos(a), math.sin(a)
    b = StreamBuilder(48, 48)
    b.draw(b.path(halfplane_poly(nx, ny, c0, lo=-far, hi=far)), color=0xffffffff)
    img = render(b)
    ys, xs = np.mgrid[0:48, 0:48] + 0.5
    d = c0 - (nx * xs + ny * ys)
    cov = np.clip(0.5 + d / (abs(nx) + abs(ny)), 0, 1)
    # float32 vertex positions a million pixels out carry ~0.06 px of error
    assert np.abs(img[..., 3].astype(float) - 255 * cov).max() <= 16.0
    assert np.mean(np.abs(img[..., 3].astype(float) - 255 * cov)) <= 1.0


@pytest.mark.parametrize("x0,x1", [(10, 30.25), (10.5, 30.5), (9.75, 20.125)])
def test_axis_aligned_rect_exact_and_premultiplied(x0, x1):
    b = StreamBuilder(40, 40)
    b.draw(b.path(rect(x0, 8, x1, 30)), color=0xc0ff8040)  # a = 192/255
    img = render(b).astype(float)
    xs = np.arange(40) + 0.5
    covx = np.clip(np.minimum(xs - x0, x1 - xs) + 0.5, 0, 1)
    row = img[20]
    a = 192 / 255
    for ch, c in enumerate((1.0, 128 / 255, 64 / 255, 1.0)):
        want = 255 * c * a * covx
        assert np.abs(row[:, ch] - want).max() <= 1.0


def test_fill_rules_on_overlapping_squares():
    cases = {NONZERO: 255, EVENODD: 0}
    for rule, overlap in cases.items():
        b = StreamBuilder(48, 48)
        b.draw(b.path(rect(5, 5, 30, 30) + rect(15, 15, 40, 40), fill_rule=rule), color=0xff0000ff)
        img = render(b)
        assert img[20, 20, 3] == overlap and img[10, 10, 3] == 255 and img[35, 35, 3] == 255
    # nonzero with opposite orientations cancels; clockwise keeps only cw area
    b = StreamBuilder(48, 48)
    b.draw(b.path(rect(5, 5, 30, 30) + rect(15, 15, 40, 40, clockwise=False)), color=0xff0000ff)
    assert render(b)[20, 20, 3] == 0
    for cw, want in ((True, 255), (False, 0)):
        b = StreamBuilder(48, 48)
        b.draw(b.path(rect(5, 5, 30, 30, clockwise=cw), fill_rule=CLOCKWISE), color=0xff0000ff)
        assert render(b)[10, 10, 3] == want


def test_orientation_and_transform_invariance():
    # Away from the corners the coverage does not depend on the contour's
    # orientation.  Pixels whose AA bands meet at a corner do: there the
    # border quads of the two edges overlap, split along a diagonal chosen by
    # the walking direction.
    b1 = StreamBuilder(64, 64)
    b1.draw(b1.path(rect(12.25, 8.625, 40.75, 50.125)), color=0xff336699)
    b2 = StreamBuilder(64, 64)
    b2.draw(b2.path(rect(12.25, 8.625, 40.75, 50.125, clockwise=False)), color=0xff336699)
    b3 = StreamBuilder(64, 64)  # same rect through a power-of-two scale + translation
    b3.draw(b3.path(rect(4.125, 1.3125, 18.375, 22.0625)), color=0xff336699, matrix=(2, 0, 0, 2, 4, 6))
    i1, i2, i3 = render(b1), render(b2), render(b3)
    assert np.array_equal(i1, i3)
    ys, xs = np.mgrid[0:64, 0:64] + 0.5
    corner = np.zeros((64, 64), bool)
    for cx in (12.25, 40.75):
        for cy in (8.625, 50.125):
            corner |= (np.abs(xs - cx) < 1.5) & (np.abs(ys - cy) < 1.5)
    assert np.array_equal(i1[~corner], i2[~corner])


def test_stroke_side_coverage():
    y, w = 20.25, 4.0
    b = StreamBuilder(48, 48)
    b.draw(b.path([(0, 6, y), (1, 42, y)]), color=0xff00ff00, stroke=w, cap=BUTT)
    img = render(b).astype(float)
    ys = np.arange(48) + 0.5
    cov = np.clip(w / 2 - np.abs(ys - y) + 0.5, 0, 1)
    assert np.abs(img[:, 24, 3] - 255 * cov).max() <= 1.0


def test_linear_gradient_two_stops():
    b = StreamBuilder(64, 4)
    b.draw(b.path(rect(0, 0, 64, 4)), gradient=("linear", 0, 0, 64, 0, [(0, 0xff000000), (1, 0xffffffff)]))
    img = render(b).astype(float)
    t = (np.arange(64) + 0.5) / 64
    assert np.abs(img[2, :, 0] - 255 * t).max() <= 1.0
    assert np.all(img[2, :, 3] == 255)


def test_nested_clips_intersect():
    b = StreamBuilder(48, 48)
    c1 = b.path(rect(10, 10, 30, 40))
    c2 = b.path(rect(20, 0, 48, 25))
    b.draw(b.path(rect(0, 0, 48, 48)), color=0xffff0000, clips=(c1, c2))
    a = render(b)[..., 3]
    inside = np.zeros((48, 48), bool)
    inside[10:25, 20:30] = True
    assert np.all(a[inside] == 255) and np.all(a[~inside] == 0)


def test_clear_color_and_dither_bound():
    b = StreamBuilder(8, 8, clear_color=0xff204080)
    img = render(b)
    assert np.all(img.reshape(-1, 4) == [0x20, 0x40, 0x80, 0xff])
    # dither perturbs each channel by at most 1/512 before rounding
    for dither in (0, 1):
        b = StreamBuilder(32, 4, dither=dither)
        b.draw(b.path(rect(0, 0, 32, 4)), color=0x80408030)
        img = render(b).astype(float)
        a = 128 / 255
        want = np.array([64 / 255 * a, 128 / 255 * a, 48 / 255 * a, a]) * 255
        err = np.abs(img.reshape(-1, 4) - want)
        assert err.max() <= (0.5 if dither == 0 else 1.0)
        if dither:
            assert len(np.unique(img[..., 0])) > 1


def test_empty_and_culled_draws_leave_clear():
    b = StreamBuilder(16, 16, clear_color=0x00000000)
    empty = b.path([])
    offscreen = b.path(rect(100, 100, 120, 120))
    b.draw(empty, color=0xffffffff)
    b.draw(offscreen, color=0xffffffff)
    b.draw(b.path(rect(2, 2, 8, 8)), color=0xffffffff, stroke=0.0)  # zero-width stroke
    assert render(b).max() == 0


def test_clip_rect_half_coverage_on_edge():
    """A clip rect's coverage is 1/2 where a pixel center lies on its edge
    (common.glsl:376-400), 0 and 1 half a pixel either side."""
    b = StreamBuilder(16, 16)
    b.draw(b.path(rect(0, 0, 16, 16)), color=0xffffffff, clips=(b.path(rect(4.5, 0, 12, 16)),))
    a = render(b)[..., 3].astype(int)
    assert np.all(a[:, 5:12] == 255) and np.all(a[:, :4] == 0) and np.all(a[:, 12:] == 0)
    assert np.all(np.abs(a[:, 4] - 128) <= 1)


def test_clip_rect_rotated_quarter_turn_is_exact():
    """Under a quarter turn a pixel-aligned clip rect stays pixel aligned."""
    b = StreamBuilder(20, 20)
    m = (0.0, 1.0, -1.0, 0.0, 20.0, 0.0)  # (x, y) -> (20 - y, x)
    b.draw(b.path(rect(-5, -5, 30, 30)), color=0xffffffff, matrix=m, clips=(b.path(rect(2, 3, 9, 17)),))
    a = render(b)[..., 3]
    inside = np.zeros((20, 20), bool)
    inside[2:9, 3:17] = True  # rows from x, columns from 20 - y
    assert np.all(a[inside] == 255) and np.all(a[~inside] == 0)


def test_clip_rects_intersect_like_one_rect():
    """Two clip rects under one matrix intersect geometrically; a rect with
    no area clips everything away."""
    def frame(clips):
        b = StreamBuilder(40, 30, dither=1)
        cs = tuple(b.path(c) for c in clips)
        b.draw(b.path(rect(0, 0, 40, 30)), color=0xc080ff20, clips=cs)
        return render(b)
    two = frame([rect(3.25, 2.5, 30.5, 20), rect(10.75, 5, 38, 27.5, clockwise=False)])
    one = frame([rect(10.75, 5, 30.5, 20)])
    assert np.array_equal(two, one)
    assert np.all(frame([rect(5, 5, 5, 25)])[..., 3] == 0)
