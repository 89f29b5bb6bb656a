"""The oracle's outer-cubic layout of interior-triangulated fills
(iterateOuterCubics, draw.cpp:2401-2650), pinned on hand-built paths.

* A large square of lines (move, three lines, close) is four patches: three
  lines and the closing line, each a 16-segment patch with one polar and one
  join segment, flagged to cull excess segments; nothing below 512 x 512 px
  or above 256 verbs, and no stroke or feather, is triangulated.
* A large cubic is cut into n = clamp(ceil(4th root of Wang's formula / 16),
  1, 64) pieces, with n from the reference's compiled wangs_formula (oracle/_ref):
  n patches whose end points are the curve at T = i/n (the reference's
  compiled eval), then ceil((n - 1) / 3) retrofitted strips whose corner points
  are piece end points, the last one the cubic's start.
* Tessellation records total 17 texels per patch and, for the double-sided
  fills of the raster-ordering mode, the texel count is twice that.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_binding as ob
from streams import CLOCKWISE, StreamBuilder
from test_oracle_counts import REF_LIB, fptr, need_ref, records, ref_lib

CULL = 1 << 29
RETRO = 1 << 31
IDENT = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)


def layout(build, triangulate=True):
    sb = StreamBuilder(1024, 1024)
    build(sb)
    s = sb.build()
    ob.set_interior_triangulation(triangulate)
    try:
        f = ob.OracleFlush(s)
    finally:
        ob.set_interior_triangulation(False)
    return f, records(f), sb


def square(sb, size=800.0, stroke=None, feather=0.0):
    a, b = 100.0, 100.0 + size
    if feather:
        # feathered fills are drawn only with the clockwise fill rule
        cmds = [(0, a, a), (1, a, b), (1, b, b), (1, b, a), (5,)]
    else:
        cmds = [(0, a, a), (1, b, a), (1, b, b), (1, a, b), (5,)]
    p = sb.path(cmds, fill_rule=CLOCKWISE if feather else 0)
    sb.draw(p, stroke=stroke, feather=feather)


def test_square_is_four_line_patches():
    f, rec, _ = layout(square)
    try:
        assert len(rec) == 4
        assert np.all(rec["parametric"] == 16) and np.all(rec["polar"] == 1) and np.all(rec["join"] == 1)
        assert np.all((rec["cfi"] & CULL) != 0) and np.all((rec["cfi"] & RETRO) == 0)
        assert np.all(rec["total"] == 17)
        assert f.tess_vertex_count >= 4 * 17 * 2
        # the lines and the closing line, in order
        ends = [tuple(r["pts"][6:8]) for r in rec]
        assert ends == [(900.0, 100.0), (900.0, 900.0), (100.0, 900.0), (100.0, 100.0)]
    finally:
        f.close()


@pytest.mark.parametrize("kw", [dict(size=400.0), dict(stroke=4.0), dict(feather=3.0)])
def test_ineligible_draws_keep_midpoint_fans(kw):
    f, rec, _ = layout(lambda sb: square(sb, **kw))
    try:
        assert len(rec) > 0 and np.all((rec["cfi"] & CULL) == 0)
    finally:
        f.close()


@need_ref
def test_cubic_pieces_and_glue_strips():
    L = ref_lib()
    cubic = np.array([100, 900, 50, -2500, 1100, 3500, 900, 120], np.float32)  # a long, bent cubic

    def build(sb):
        p = sb.path([(0, float(cubic[0]), float(cubic[1])),
                     (4, *[float(v) for v in cubic[2:8]]), (5,)])
        sb.draw(p)
    f, rec, _ = layout(build)
    try:
        a, pa = fptr(cubic)
        m, pm = fptr(np.array(IDENT, np.float32))
        pow4 = np.float32(L.ref_wang_cubic_pow4(pa, 4.0, pm))
        wang = np.sqrt(np.sqrt(pow4, dtype=np.float32), dtype=np.float32)
        n = int(min(max(np.ceil(wang * np.float32(1 / 16)), 1), 64))
        assert n > 2
        strips = (n + 1) // 3
        cubic_rec = rec[:n + strips]
        pieces, glue = cubic_rec[:n], cubic_rec[n:]
        assert np.all((pieces["cfi"] & CULL) != 0) and np.all((glue["cfi"] & RETRO) != 0)
        out = np.zeros(2, np.float32)
        ends = []
        for i in range(1, n + 1):
            L.ref_eval_cubic(pa, ctypes.c_float(i / n), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            ends.append(out.copy())
        got = np.array([r["pts"][6:8] for r in pieces])
        assert np.abs(got - np.array(ends)).max() < 2e-2
        # strips: corner points are piece ends, the strip's last corner the cubic's start
        corners = set(map(tuple, np.vstack([cubic[0:2][None], got]).round(3)))
        for r in glue:
            pts = {tuple(np.round(r["pts"][2 * k:2 * k + 2], 3)) for k in range(4)}
            assert pts <= corners
        # each strip's fifth point (the join-tangent slot) is the fan's anchor
        assert all(tuple(np.round(r["jt"], 3)) == tuple(np.round(cubic[0:2], 3)) for r in glue)
        # the closing line (end != start) follows
        assert len(rec) == n + strips + 1 and (rec[-1]["cfi"] & CULL) != 0
    finally:
        f.close()


@pytest.mark.parametrize("kw,code", [(dict(), 3), (dict(size=400.0), 0), (dict(stroke=4.0), 1),
                                     (dict(feather=3.0), 2)])
def test_draw_export_marks_outer_cubic_draws(kw, code):
    """The draw table's is_stroke code (the front end's mark that the device
    counts by) is 3 exactly for a draw laid out as outer-cubic patches."""
    f, _, _ = layout(lambda sb: square(sb, **kw))
    try:
        assert f.draws()[:, 4].tolist() == [code]
    finally:
        f.close()
    f, _, _ = layout(lambda sb: square(sb, **kw), triangulate=False)
    try:
        assert f.draws()[:, 4].tolist() == [0 if code == 3 else code]
    finally:
        f.close()
