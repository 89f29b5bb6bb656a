"""Interior-triangulated fills laid out by the device itself (GEOMETRY mode:
the context's front end with interior triangulation switched on, the
reference's PathDraw::Make at draw.cpp:510-538 with the triangulation
controller's guards alone).

A plain fill of at least 512 x 512 px on screen and at most 256 verbs is
counted and emitted as outer-cubic patches (iterateOuterCubics,
draw.cpp:2401-2650): the draw ranges, contours and spans of every logical
flush equal the oracle's bit for bit when the oracle makes the same choice,
and the draws the device marks (is_stroke 3) are exactly those the oracle
lays out that way.  The frame then differs from the midpoint-fan frame only
by the two tessellations' difference (each polyline within 1/4 px of the
curve): pixels beyond 2/255 lie within a pixel of an anti-aliased edge of the
midpoint-fan frame, on at most 3 % of the frame, by at most 128/255.
"""
import numpy as np
import pytest

import oracle_binding as ob
import rivehip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,n,w,h", [(1, 60, 1024, 1024), (2, 150, 960, 540)])
def test_outer_cubic_geometry_layout_and_pixels(hip_ctx, cfg, n, w, h):
    print(f"[triangulation-geometry] config {cfg}, {n} draws {w}x{h}", flush=True)
    s = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
    t_fan, t_tri = hip_ctx.target(w, h), hip_ctx.target(w, h)
    fan = tri = None
    assert not hip_ctx.interior_triangulation()
    try:
        fan = ob.OracleFlush(s)
        ob.set_interior_triangulation(True)
        try:
            tri = ob.OracleFlush(s)
        finally:
            ob.set_interior_triangulation(False)
        hip_ctx.render_stream(s, t_fan)
        dev_fan = hip_ctx.read_pixels(t_fan, w, h).astype(np.int32)
        hip_ctx.set_interior_triangulation(True)
        try:
            assert hip_ctx.interior_triangulation()
            for k in range(tri.flush_count):
                nf = hip_ctx.render_stream(s, t_tri, max_flushes=k + 1)
                assert nf == tri.flush_count
                hip_ctx.sync()
                spans = hip_ctx.readback(0).view(rivehip.TESS_SPAN)
                contours = hip_ctx.readback(1).view(rivehip.CONTOUR)
                draws = hip_ctx.readback(3).view(rivehip.DRAW)
                tri.select(k)
                od = tri.draws()
                marked = int((draws["is_stroke"] == 3).sum())
                print(f"[triangulation-geometry]   flush {k}: {marked} of {len(draws)} draws as outer cubics",
                      flush=True)
                assert len(draws) == len(od)
                assert marked > 0
                assert np.array_equal(draws["is_stroke"], od[:, 4])
                assert np.array_equal(draws["path_id"], od[:, 0])
                assert np.array_equal(draws["tess_location"], od[:, 1])
                assert np.array_equal(draws["tess_vertex_count"], od[:, 2])
                ocont = tri.get(1)
                assert len(contours) == len(ocont)
                assert contours.tobytes() == ocont.tobytes()
                ospans = tri.get(0)
                assert len(spans) == len(ospans)
                mism = np.nonzero(spans.view(np.uint32).reshape(-1, 16) != ospans.view(np.uint32).reshape(-1, 16))
                assert mism[0].size == 0, f"flush {k}: first mismatching span {mism[0][:4]}, fields {mism[1][:4]}"
            hip_ctx.render_stream(s, t_tri)
            dev_tri = hip_ctx.read_pixels(t_tri, w, h).astype(np.int32)
        finally:
            hip_ctx.set_interior_triangulation(False)
        ref, edges, _ = fan.render_edges()
        near = edges.copy()
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                near |= np.roll(np.roll(edges, dy, 0), dx, 1)
        for name, other in [("oracle midpoint fans", ref.astype(np.int32)), ("device midpoint fans", dev_fan)]:
            d = np.abs(dev_tri - other).max(axis=2)
            over = d > 2
            print(f"[triangulation-geometry]   vs {name}: max {int(d.max())}/255, >2/255 on {over.mean():.2e}, "
                  f"beyond a pixel of AA edges {int((over & ~near).sum())}", flush=True)
            assert over.mean() <= 0.03 and int(d.max()) <= 128
            assert not np.any(over & ~near)
    finally:
        for f in (fan, tri):
            if f is not None:
                f.close()
        hip_ctx.free_target(t_fan)
        hip_ctx.free_target(t_tri)
        rivehip.free_stream(s)
