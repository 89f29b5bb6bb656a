"""Geometry far outside the canvas and non-finite geometry on the device.

Vertices a million pixels away: the device replaces primitives leaving its
guard box (canvas +- 16384 px) with the pieces the clamp onto the box maps
them to, which keeps every winding number and border inside the box; the
oracle evaluates the unclipped triangles with exact integer predicates.
Non-finite positions drop the triangles through them on both sides (a fill's
fan triangles through a non-finite vertex included).

Closed form: a fill whose one edge, from (-1e6, -1e6 + 30) to
(1e6, 1e6 + 30), crosses the canvas diagonally covers pixel (x, y) by
clamp(0.5 + (y - x - 30) / 2, 0, 1) (the AA ramp of a 45-degree edge,
common.glsl manhattan width sqrt(2)).
"""
import math

import numpy as np
import pytest

import oracle_binding as ob
import rivehip
from streams import ROUND, StreamBuilder

pytestmark = pytest.mark.gpu
W, H = 160, 120
NAN = float("nan")


def sc_diagonal_fill():
    b = StreamBuilder(W, H, clear_color=0xff000000)
    b.draw(b.path([(0, -1e6, -1e6 + 30), (1, 1e6, 1e6 + 30), (1, -1e6, 1e6 + 30), (5,)]), color=0xffffffff)
    return b


def sc_diagonal_stroke():
    b = StreamBuilder(W, H, clear_color=0xff000000)
    b.draw(b.path([(0, -1e6, -1e6 + 60), (1, 1e6, 1e6 + 60)]), color=0xff40c0ff, stroke=6.0)
    b.draw(b.path([(0, 2e6, -3e5), (1, -1e6, 80), (1, 3e5, 3e6)]), color=0xa0ff8000, stroke=3.0,
           join=ROUND)
    return b


def sc_far_star():
    """A nonconvex fill with far vertices on every side, even-odd and nonzero."""
    b = StreamBuilder(W, H, clear_color=0xff202020)
    far = [(0, 80, -7e5), (1, 150, 60), (1, 9e5, 50), (1, 90, 110), (1, 60, 9e5), (1, 20, 70),
           (1, -8e5, 40), (1, 40, 20), (5,)]
    b.draw(b.path(far), color=0xc000ff00)
    b.draw(b.path(far, fill_rule=1), color=0x80ff00ff, matrix=(0.9, 0.1, -0.1, 0.9, 10, 5))
    return b


def sc_far_curves():
    b = StreamBuilder(W, H, clear_color=0xff000000)
    b.draw(b.path([(0, 20, 20), (4, 3e5, -2e5, -2e5, 4e5, 140, 100), (1, 20, 100), (5,)]), color=0xffffc000)
    b.draw(b.path([(0, -5e5, 60), (4, 80, -4e5, 80, 4e5, 5e5, 60)]), color=0xff00c0ff, stroke=4.0)
    return b


def sc_nan_points():
    b = StreamBuilder(W, H, clear_color=0xff000000)
    b.draw(b.path([(0, 10, 10), (1, 100, 10), (1, NAN, 50), (1, 10, 100), (5,)]), color=0xffffffff)
    b.draw(b.path([(0, 30, 30), (1, 140, 30), (1, 140, 110), (5,), (0, 60, 60), (1, NAN, NAN), (1, 50, 90), (5,)]),
           color=0xff00ff00)
    b.draw(b.path([(0, 10, 60), (1, 80, 60), (1, NAN, 70), (1, 150, 100)]), color=0xffff0000, stroke=5.0)
    return b


SCENES = {"diagonal_fill": sc_diagonal_fill, "diagonal_stroke": sc_diagonal_stroke, "far_star": sc_far_star,
          "far_curves": sc_far_curves, "nan_points": sc_nan_points}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_far_and_nonfinite_geometry_matches_oracle(hip_ctx, name):
    print(f"[guard-band] {name}", flush=True)
    b = SCENES[name]()
    s = b.build()
    f = ob.OracleFlush(s)
    t = hip_ctx.target(W, H)
    try:
        ref, _ = f.render()
        hip_ctx.render_stream(s, t)
        out = hip_ctx.read_pixels(t, W, H)
        stats = hip_ctx.last_flush_stats().as_dict()
    finally:
        hip_ctx.free_target(t)
        f.close()
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"max {d.max()}, frac>2 {(d > 2).mean()}"
    if name.startswith("diagonal") or name.startswith("far"):
        assert stats["guard_pieces"] > 0
    if name == "diagonal_fill":
        ys, xs = np.mgrid[0:H, 0:W]
        cov = np.clip(0.5 + ((ys + .5) - (xs + .5) - 30) / 2, 0, 1)
        assert np.abs(out[..., 0].astype(float) - 255 * cov).max() <= 1.0
