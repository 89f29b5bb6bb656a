"""Binning limits through the C ABI on the device, against the oracle: a path
grid larger than the LDS grid of the binning kernels (more than 8192 tiles,
so counts, corner crossings and bin slots go through global atomics), and
draws with more than 32768 primitives (beyond the packed LDS count and
corner-crossing word), fills and strokes.  The fills keep their running
coverage counts small: the reference accumulates coverage in a half-precision
plane, which the oracle follows and the device's exact sums match only where
that plane rounds nothing.

Pixels within 2/255 on all but 0.1% of pixels and 8/255 everywhere.
"""
import math

import numpy as np
import pytest

import oracle_binding as ob
import rivehip
from streams import EVENODD, NONZERO, ROUND, ROUND_CAP, StreamBuilder, rect

pytestmark = pytest.mark.gpu


def star(cx, cy, r0, r1, n):
    cmds = []
    for i in range(2 * n):
        r = r1 if i % 2 == 0 else r0
        a = math.pi * i / n
        cmds.append((0 if i == 0 else 1, cx + r * math.cos(a), cy + r * math.sin(a)))
    return cmds + [(5,)]


def sc_wide_grid():
    # 1504 x 1400 px = 94 x 88 = 8272 tiles per canvas-sized path
    b = StreamBuilder(1504, 1400, dither=1)
    b.draw(b.path(star(752, 700, 200, 760, 9)), color=0xc0ff8040)
    b.draw(b.path(star(700, 720, 300, 690, 7), fill_rule=EVENODD), color=0x9030a0ff)
    b.draw(b.path(star(780, 690, 150, 700, 5)), color=0xe0ffffff, stroke=7.0, join=ROUND, cap=ROUND_CAP)
    b.draw(b.path(rect(-20, -20, 1530, 1420)), color=0x8020ff20, clips=(b.path(star(752, 700, 250, 740, 11)),))
    return b


def rings(cx, cy, r0, r1, contours, verts):
    """Concentric circles of alternating orientation: nonzero and even-odd
    windings stay within {0, 1}, so the reference's half-precision coverage
    plane rounds nothing the exact sums keep."""
    cmds = []
    for k in range(contours):
        r = r0 + (r1 - r0) * k / max(contours - 1, 1)
        sgn = 1 if k % 2 == 0 else -1
        for i in range(verts):
            a = sgn * 2 * math.pi * i / verts
            cmds.append((0 if i == 0 else 1, cx + r * math.cos(a), cy + r * math.sin(a)))
        cmds.append((5,))
    return cmds


def spiral(cx, cy, r0, r1, turns, n):
    cmds = []
    for i in range(n):
        u = i / (n - 1)
        a = 2 * math.pi * turns * u
        r = r0 + (r1 - r0) * u
        cmds.append((0 if i == 0 else 1, cx + r * math.cos(a), cy + r * math.sin(a)))
    return cmds


def sc_huge_draws():
    # 256 x 224 px; each draw has more than 32768 primitives
    b = StreamBuilder(256, 224, dither=1)
    b.draw(b.path(rect(0, 0, 256, 224)), color=0xff102030)
    b.draw(b.path(rings(128, 112, 6, 104, 40, 860), fill_rule=NONZERO), color=0xa0ff6020)
    b.draw(b.path(rings(120, 104, 9, 96, 40, 840), fill_rule=EVENODD), color=0x8020a0ff)
    b.draw(b.path(spiral(132, 110, 3, 100, 31, 17000)), color=0x60ffffff, stroke=1.5)
    return b


@pytest.mark.parametrize("name", ["wide_grid", "huge_draws"])
def test_binning_limits_match_oracle(hip_ctx, name):
    print(f"[binning-limits] {name}", flush=True)
    b = {"wide_grid": sc_wide_grid, "huge_draws": sc_huge_draws}[name]()
    s = b.build()
    w, h = b.width, b.height
    f = ob.OracleFlush(s)
    t = hip_ctx.target(w, h)
    try:
        ref, _ = f.render()
        hip_ctx.render_stream(s, t)
        out = hip_ctx.read_pixels(t, w, h)
        st = hip_ctx.last_flush_stats().as_dict()
    finally:
        hip_ctx.free_target(t)
        f.close()
    if name == "wide_grid":
        assert st["tile_count"] > 8192
    else:
        assert st["prim_count"] > 2 * 32768
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"max {d.max()}, frac>2 {(d > 2).mean()}"


def _desc(stats, target):
    d = rivehip.FlushDesc()
    d.target = target
    d.mode = rivehip.FLUSH_GEOMETRY
    d.load_action = stats.load_action
    d.clear_color = stats.clear_color
    d.dither = stats.dither
    d.path_count = stats.path_count
    d.draw_count = stats.draw_count
    d.grad_span_count = stats.grad_span_count
    d.grad_simple_rows = stats.grad_simple_rows
    d.grad_complex_rows = stats.grad_complex_rows
    d.first_path_tess_location = stats.first_path_tess_location
    return d


def test_draws_sharing_a_path_bin_together(hip_ctx):
    """Two draw records naming one path bin into that path's grid (the
    shared-path branch of the scatter: one count in LDS, a global reservation
    per cell): the frame equals, bit for bit, one draw of a path holding both
    contours."""
    print("[binning-limits] shared path", flush=True)
    ra, rb = rect(3, 4, 40, 30), rect(20, 12, 61, 45)
    one = StreamBuilder(64, 48, dither=1)
    one.draw(one.path(ra + rb), color=0xc0ff4000)
    two = StreamBuilder(64, 48, dither=1)
    two.draw(two.path(ra), color=0xc0ff4000)
    two.draw(two.path(rb), color=0x8000a0ff)
    t = hip_ctx.target(64, 48)
    try:
        hip_ctx.render_stream(one.build(), t)
        ref = hip_ctx.read_pixels(t, 64, 48)
        hip_ctx.render_stream(two.build(), t)
        stats = hip_ctx.last_flush_stats()
        draws = hip_ctx.readback(3).view(rivehip.DRAW).copy()
        assert len(draws) == 2 and draws[0]["path_id"] != draws[1]["path_id"]
        draws[1]["path_id"] = draws[0]["path_id"]
        hip_ctx.upload("draw", draws)
        hip_ctx.flush(_desc(stats, t))
        out = hip_ctx.read_pixels(t, 64, 48)
    finally:
        hip_ctx.free_target(t)
    assert np.array_equal(out, ref)
