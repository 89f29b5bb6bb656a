"""Edge cases through the C ABI on the device, against the oracle: empty and
culled draws, odd target sizes, off-canvas and huge coordinates, degenerate
and cusped cubics, every join and cap, clockwise fills under a mirroring
matrix, nested and reused clips, radial and many-stop gradients.

Layout (spans, contours, per-draw vertex ranges) is compared bit for bit;
pixels within 2/255 on all but 0.1% of pixels and 8/255 everywhere.
"""
import math

import numpy as np
import pytest

import oracle_binding as ob
import rivehip
from streams import (BEVEL, BUTT, CLOCKWISE, EVENODD, MITER, NONZERO, ROUND, ROUND_CAP, SQUARE,
                     StreamBuilder, rect)

pytestmark = pytest.mark.gpu


def star(cx, cy, r0, r1, n=7):
    cmds = []
    for i in range(2 * n):
        r = r1 if i % 2 == 0 else r0
        a = math.pi * i / n
        cmds.append((0 if i == 0 else 1, cx + r * math.cos(a), cy + r * math.sin(a)))
    return cmds + [(5,)]


def sc_empty_frame():
    return StreamBuilder(40, 24, clear_color=0xff102030)


def sc_culled_draws():
    b = StreamBuilder(37, 23, clear_color=0x80ffffff)
    b.draw(b.path([]), color=0xffff0000)
    b.draw(b.path(rect(100, 100, 120, 130)), color=0xffff0000)
    b.draw(b.path(rect(2, 2, 9, 9)), color=0xffff0000, stroke=0.0)
    b.draw(b.path(rect(3, 3, 20, 20)), color=0xff00ff00)
    return b


def sc_one_pixel():
    b = StreamBuilder(1, 1)
    b.draw(b.path(rect(0.25, 0.25, 0.75, 0.75)), color=0xffffffff)
    return b


def sc_offcanvas_and_huge():
    b = StreamBuilder(50, 30, dither=1)
    b.draw(b.path(rect(-40, -40, 20, 15)), color=0xff3366ff)
    b.draw(b.path([(0, -1e5, 10), (1, 1e5, 12), (1, 0, 25), (5,)]), color=0xc0ff0000)
    b.draw(b.path(star(45, 25, 8, 30)), color=0xa000ff00)
    return b


def sc_degenerate_cubics():
    b = StreamBuilder(48, 48)
    b.draw(b.path([(0, 10, 10), (4, 10, 10, 10, 10, 10, 10), (1, 30, 12), (4, 30, 12, 30, 12, 20, 40),
                   (5,)]), color=0xffffffff)
    b.draw(b.path([(0, 5, 40), (4, 45, 2, 2, 2, 42, 40)]), color=0xffff8000, stroke=3.0, join=ROUND,
           cap=ROUND_CAP)  # a cusp
    b.draw(b.path([(0, 30, 30), (4, 30, 30, 30, 30, 30, 30)]), color=0xff0000ff, stroke=6.0,
           cap=ROUND_CAP)  # an empty stroke with caps is a dot
    b.draw(b.path([(0, 40, 8), (4, 40, 8, 40, 8, 40, 8), (5,)]), color=0xff00ffff, stroke=5.0,
           join=MITER)
    return b


def sc_clockwise_mirrored():
    b = StreamBuilder(48, 48)
    p = b.path(star(24, 24, 8, 20, 9), fill_rule=CLOCKWISE)
    b.draw(p, color=0xff00ff00)
    b.draw(p, color=0x8000ffff, matrix=(-1, 0, 0, 1, 48, 0))  # det < 0
    q = b.path(rect(4, 4, 20, 20, clockwise=False), fill_rule=CLOCKWISE)
    b.draw(q, color=0xffff0000)
    return b


def sc_clips():
    b = StreamBuilder(64, 64, dither=1)
    c1 = b.path(star(32, 32, 12, 30, 6))
    c2 = b.path(rect(10, 8, 50, 40), fill_rule=EVENODD)
    c3 = b.path(rect(200, 200, 210, 210))  # off canvas: clips everything
    body = b.path(rect(0, 0, 64, 64))
    b.draw(body, color=0xffff0000, clips=(c1,))
    b.draw(body, color=0x800000ff, clips=(c1, c2))
    b.draw(body, color=0x8000ff00, clips=(c1, c2))  # the same clip stack again
    b.draw(body, color=0xffffffff, clips=(c3,))
    b.draw(b.path(star(20, 44, 4, 12)), color=0xffffff00, stroke=2.0, clips=(c2,))
    return b


def sc_gradients():
    b = StreamBuilder(96, 64, dither=1)
    stops8 = [(i / 7, 0xff000000 | (i * 0x243b51 & 0xffffff)) for i in range(8)]
    b.draw(b.path(rect(0, 0, 96, 30)), gradient=("linear", 5, 0, 90, 20, stops8))
    b.draw(b.path(rect(0, 30, 96, 64)), gradient=("radial", 48, 47, 30,
                                                  [(0, 0xffff0000), (.4, 0x8000ff00), (1, 0x000000ff)]))
    b.draw(b.path(star(48, 32, 10, 25)), gradient=("linear", 0, 0, 0, 64, [(0, 0xff00ff00), (1, 0xffff00ff)]),
           stroke=3.0, join=ROUND)
    b.draw(b.path(rect(70, 5, 90, 25)), gradient=("linear", 70, 0, 90, 0, [(.2, 0xffffffff), (.8, 0xff000000)]))
    return b


def sc_blend_modes():
    b = StreamBuilder(100, 66, dither=1)
    b.draw(b.path(rect(0, 0, 100, 66)),
           gradient=("linear", 0, 0, 100, 66, [(0, 0xff2040a0), (.5, 0xa0e08040), (1, 0x60f0c020)]))
    for k, mode in enumerate(range(14, 29)):
        x, y = (k % 5) * 20 + 10, (k // 5) * 22 + 11
        b.draw(b.path(star(x, y, 4, 10, 6)), color=0xc0000000 | (k * 0x10a3c7 & 0xffffff), blend=mode)
    b.draw(b.path(rect(5, 50, 95, 64)), blend=24,
           gradient=("radial", 50, 57, 40, [(0, 0xffffffff), (1, 0x80ff0000)]))
    return b


def random_premul(rng, h, w, opaque):
    rgb = rng.integers(0, 256, (h, w, 3))
    a = np.full((h, w, 1), 255) if opaque else rng.integers(0, 256, (h, w, 1))
    return np.concatenate([np.minimum(rgb, a), a], axis=2).astype(np.uint8)


def sc_images():
    rng = np.random.default_rng(5)
    b = StreamBuilder(96, 80, dither=1)
    t1 = b.image(random_premul(rng, 24, 32, True))
    t2 = b.image(random_premul(rng, 64, 64, False))
    b.draw(b.path(rect(0, 0, 96, 80)), color=0xff305070)
    b.draw_image(t1, matrix=(1, 0, 0, 1, 4, 4))
    b.draw_image(t1, matrix=(.9, .3, -.3, .9, 60, 2), sampler=9, opacity=.7)
    b.draw_image(t2, matrix=(.25, 0, 0, .25, 10, 50), blend=24)
    b.draw_image(t2, matrix=(.6, 0, 0, .45, 40, 40), sampler=4)
    # the clip is drawn under the image's matrix too
    b.draw_image(t1, matrix=(1.5, 0, 0, 1.5, 50, 30), clips=(b.path(star(8, 13, 4, 11)),), opacity=.9)
    return b


def sc_clip_rects():
    """Axis-aligned rectangle clips become clip rects (ClipRectInverseMatrix):
    offset edges, rotation, intersection, a clip rect with a clip path, a
    counterclockwise rect, a rotated quad (a clip path), an empty rect and an
    image under a clip rect."""
    rng = np.random.default_rng(9)
    b = StreamBuilder(90, 70, dither=1)
    b.draw(b.path(rect(0, 0, 90, 70)), color=0xff203040)
    full = b.path(rect(-10, -10, 100, 80))
    b.draw(full, color=0xc0ff8000, clips=(b.path(rect(10.5, 8.25, 40.75, 30.5)),))
    rot = (math.cos(.4), math.sin(.4), -math.sin(.4), math.cos(.4), 55, 5)
    b.draw(b.path(star(12, 14, 5, 16)), color=0xe000c0ff, matrix=rot, clips=(b.path(rect(0, 0, 20.5, 25)),))
    b.draw(full, color=0x9000ff40,
           clips=(b.path(rect(5, 35, 60, 65)), b.path(rect(20.25, 30, 80, 55.5, clockwise=False))))
    b.draw(b.path(star(70, 50, 6, 18)), color=0xffffffff,
           clips=(b.path(rect(55, 38, 86, 68)), b.path(star(70, 52, 8, 16))))
    quad = [(0, 45, 20), (1, 60, 28), (1, 52, 43), (1, 37, 35), (5,)]
    b.draw(full, color=0x60ff00ff, clips=(b.path(quad),))
    b.draw(full, color=0xffff0000, clips=(b.path(rect(30, 30, 30, 50)),))
    img = b.image(random_premul(rng, 16, 16, True))
    b.draw_image(img, matrix=(2, 0, 0, 2, 2, 40), clips=(b.path(rect(2.5, 1.5, 13.5, 12)),), opacity=.8)
    return b


def sc_zero_coverage_windings():
    """Whole tiles where a path's winding is nonzero but its coverage is zero:
    even-odd with winding 2 (as a fill and as a clip) and a counterclockwise
    rect under the clockwise rule, so no primitives and a zero-coverage corner
    winding in those tiles."""
    b = StreamBuilder(80, 72, dither=1)
    b.draw(b.path(rect(0, 0, 80, 72)), color=0xff204060)
    b.draw(b.path(rect(1, 1, 60, 62) + rect(3, 2, 62, 66), fill_rule=EVENODD), color=0xc0ff8000)
    b.draw(b.path(rect(36, 2, 79, 50, clockwise=False), fill_rule=CLOCKWISE), color=0xff00ff00)
    b.draw(b.path(rect(36, 2, 79, 50), fill_rule=CLOCKWISE), color=0x6000ff00)
    b.draw(b.path(rect(0, 0, 80, 72)), color=0x9000ffff,
           clips=(b.path(rect(2, 3, 70, 60) + rect(4, 5, 74, 68), fill_rule=EVENODD),))
    return b


SCENARIOS = {
    "clip_rects": sc_clip_rects, "zero_coverage_windings": sc_zero_coverage_windings,
    "empty_frame": sc_empty_frame, "culled_draws": sc_culled_draws, "one_pixel": sc_one_pixel,
    "offcanvas_and_huge": sc_offcanvas_and_huge, "degenerate_cubics": sc_degenerate_cubics,
    "clockwise_mirrored": sc_clockwise_mirrored, "clips": sc_clips, "gradients": sc_gradients,
    "blend_modes": sc_blend_modes, "images": sc_images,
}
def build_join_cap(join, cap):
    b = StreamBuilder(64, 48, dither=1)
    zig = [(0, 6, 40), (1, 16, 8), (1, 26, 40), (1, 30, 10), (4, 40, 0, 60, 50, 58, 20)]
    b.draw(b.path(zig), color=0xe0204080, stroke=4.5, join=join, cap=cap)
    b.draw(b.path(star(32, 24, 6, 14, 5)), color=0x80ff0000, stroke=1.5, join=join, cap=cap)
    return b


for _j in (MITER, ROUND, BEVEL):
    for _c in (BUTT, ROUND_CAP, SQUARE):
        SCENARIOS[f"join{_j}_cap{_c}"] = (lambda j=_j, c=_c: build_join_cap(j, c))


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_edge_case_matches_oracle(hip_ctx, name):
    print(f"[edge-case] {name}", flush=True)
    b = SCENARIOS[name]()
    s = b.build()
    w, h = b.width, b.height
    f = ob.OracleFlush(s)
    t = hip_ctx.target(w, h)
    try:
        ref, _ = f.render()
        hip_ctx.render_stream(s, t)
        out = hip_ctx.read_pixels(t, w, h)
        spans = hip_ctx.readback(0)
        draws = hip_ctx.readback(3).view(rivehip.DRAW)
        ospans = f.get(0)
        odraws = f.draws()
    finally:
        hip_ctx.free_target(t)
        f.close()
    assert spans.tobytes() == ospans.view(np.uint8).tobytes()
    assert len(draws) == len(odraws)
    if len(draws):
        assert np.array_equal(draws["tess_location"], odraws[:, 1])
        assert np.array_equal(draws["tess_vertex_count"], odraws[:, 2])
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"max {d.max()}, frac>2 {(d > 2).mean()}"
