"""Interior-triangulated fills handed over in SPANS mode (the reference with
a triangulation budget: draw.cpp:510-538, iterateOuterCubics
draw.cpp:2401-2650).

Large fills (at least 512 x 512 px on screen, at most 256 verbs) arrive as
outer-cubic patches: every line and every cubic (chopped uniformly into
pieces of at most 16 parametric segments) as a 16-segment patch flagged to
cull the segments it does not need (tessellate.glsl:195-211), plus triangle
strips retrofitted into patches between a chopped cubic's chord and its
pieces' chords; the interior triangles come separately and are not passed.
The device tessellates the patches as the oracle's restatement of the
tessellation shader does (forward texels equal, positions within 1e-3 px)
and draws the outline they trace, which covers what the reference's patches,
glue strips and interior triangles cover together (the chords cancel).  The
frame is then the same stream drawn with midpoint fans up to the difference
between the two tessellations: each polyline lies within 1/4 px of the
curve (Wang's formula at precision 4), so the two outlines may be up to
1/2 px apart.  Pixels beyond 2/255 must lie within a pixel of an
anti-aliased edge of the midpoint-fan frame, on at most 3 % of the frame,
and by at most 128/255 (half a pixel of coverage).
"""
import numpy as np
import pytest

import oracle_binding as ob
import rivehip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,n,w,h", [(1, 60, 1024, 1024), (2, 150, 960, 540)])
def test_outer_cubic_spans_match_midpoint_fans(hip_ctx, cfg, n, w, h):
    print(f"[triangulation] config {cfg}, {n} draws {w}x{h}", flush=True)
    s = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
    t_fan, t_tri = hip_ctx.target(w, h), hip_ctx.target(w, h)
    fan = tri = None
    try:
        fan = ob.OracleFlush(s)
        ob.set_interior_triangulation(True)
        try:
            tri = ob.OracleFlush(s)
        finally:
            ob.set_interior_triangulation(False)
        assert fan.flush_count == 1 and tri.flush_count == 1
        draws_tri = tri.draws()
        outer = sum(1 for (_, _, cnt, _, _), (_, _, cnt0, _, _) in zip(draws_tri, fan.draws()) if cnt != cnt0)
        print(f"[triangulation]   {outer} of {len(draws_tri)} draws laid out as outer-cubic patches; "
              f"{tri.tess_vertex_count} texels against {fan.tess_vertex_count}", flush=True)
        assert outer > 0
        fan.flush_spans_on(hip_ctx, t_fan)
        dev_fan = hip_ctx.read_pixels(t_fan, w, h).astype(np.int32)
        tri.flush_spans_on(hip_ctx, t_tri)
        dev_tri = hip_ctx.read_pixels(t_tri, w, h).astype(np.int32)
        # the outer-cubic patches' forward texels, as the oracle tessellates them
        tex = hip_ctx.readback(2).view(rivehip.TESS_TEXEL)
        tref = tri.tessellate()
        n = min(len(tex), len(tref))
        live = tex["w"][:n] != 0
        assert live.sum() > 0 and np.array_equal(tex["w"][:n][live], tref["w"][:n][live])
        assert np.abs(tex["x"][:n][live] - tref["x"][:n][live]).max() <= 1e-3
        assert np.abs(tex["y"][:n][live] - tref["y"][:n][live]).max() <= 1e-3
        ref, edges, _ = fan.render_edges()
        near = edges.copy()
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                near |= np.roll(np.roll(edges, dy, 0), dx, 1)
        for name, other in [("oracle midpoint fans", ref.astype(np.int32)), ("device midpoint fans", dev_fan)]:
            d = np.abs(dev_tri - other).max(axis=2)
            over = d > 2
            print(f"[triangulation]   vs {name}: max {int(d.max())}/255, >2/255 on {over.mean():.2e}, "
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
