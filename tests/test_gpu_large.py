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
