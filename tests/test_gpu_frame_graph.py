"""Frames captured in HIP graphs (rive_hip_frame_graph_*): the 1080p
5000-draw frame (six logical flushes) and the 1024^2 200-path frame are
captured once, then replayed several times with unrelated eager work (other
frames rendered into the same and into other targets, synchronisations)
between the replays; every replay's pixels equal the eager frame's bit for
bit.  The capture holds kernels and copies only (buffers are cleared by a
fill kernel, not hipMemsetAsync).
"""
import numpy as np
import pytest

import rivehip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config", [1, 2])
def test_graph_replays_equal_the_eager_frame(hip_ctx, config):
    print(f"[graph] config {config}", flush=True)
    s = rivehip.generate_stream(config)
    other = rivehip.generate_stream(1, draw_count=50)
    w, h = s.contents.width, s.contents.height
    ow, oh = other.contents.width, other.contents.height
    t = hip_ctx.target(w, h)
    t2 = hip_ctx.target(ow, oh)
    g = None
    try:
        hip_ctx.render_stream(s, t)
        frame = hip_ctx.last_frame()
        hip_ctx.flush_frame(frame, t)
        hip_ctx.sync()
        eager = hip_ctx.read_pixels(t, w, h)
        g = hip_ctx.frame_graph(frame, t)
        for k in range(3):
            # unrelated eager work between the replays
            hip_ctx.render_stream(other, t2)
            hip_ctx.sync()
            if k == 1:
                hip_ctx.render_stream(other, t) if (ow, oh) == (w, h) else None
            hip_ctx.frame_graph_launch(g)
            hip_ctx.sync()
            got = hip_ctx.read_pixels(t, w, h)
            assert np.array_equal(got, eager), f"replay {k}"
    finally:
        if g is not None:
            hip_ctx.frame_graph_destroy(g)
        hip_ctx.free_target(t)
        hip_ctx.free_target(t2)
        rivehip.free_stream(s)
        rivehip.free_stream(other)
