"""The front end's logical-flush split on the device, at each of the
reference's per-flush limits (render_context.cpp:43-53,484-492,517-541,
617-654): path IDs (30 719), contour IDs (65 535), tessellation vertices
(2048^2 - 25) and gradient rows (2048).  For each stream the device splits
where the oracle splits, every logical flush's draw ranges, contours and
spans equal the oracle's bit for bit, and the composed frame (flushes after
the first load the target) matches the oracle's pixels.
"""
import numpy as np
import pytest

import oracle_binding as ob
import rivehip
from streams import contour_limit_stream, gradient_limit_stream, pixel_rects_stream, tess_limit_stream

pytestmark = pytest.mark.gpu


def build_case(name):
    if name == "paths":
        b, _ = pixel_rects_stream(30720)
        return b
    return {"contours": contour_limit_stream, "tess": tess_limit_stream,
            "gradients": gradient_limit_stream}[name]()


def same_bits(a, b):
    return a.view(np.uint8).tobytes() == b.view(np.uint8).tobytes()


@pytest.mark.parametrize("name", ["paths", "contours", "tess", "gradients"])
def test_device_splits_where_the_oracle_splits(hip_ctx, name):
    print(f"[flush-limits] {name}", flush=True)
    b = build_case(name)
    s = b.build()
    w, h = b.width, b.height
    f = ob.OracleFlush(s)
    t = hip_ctx.target(w, h)
    try:
        assert f.flush_count >= 2
        for k in range(f.flush_count):
            n = hip_ctx.render_stream(s, t, max_flushes=k + 1)
            assert n == f.flush_count, (n, f.flush_count)
            hip_ctx.sync()
            spans = hip_ctx.readback(0).view(rivehip.TESS_SPAN)
            contours = hip_ctx.readback(1).view(rivehip.CONTOUR)
            draws = hip_ctx.readback(3).view(rivehip.DRAW)
            f.select(k)
            od = f.draws()
            assert len(draws) == len(od), (k, len(draws), len(od))
            assert np.array_equal(draws["path_id"], od[:, 0])
            assert np.array_equal(draws["tess_location"], od[:, 1])
            assert np.array_equal(draws["tess_vertex_count"], od[:, 2])
            assert same_bits(contours, f.get(1)), k
            assert same_bits(spans, f.get(0)), k
        frame = hip_ctx.last_frame()
        assert len(frame) == f.flush_count
        assert [int(d.load_action) for d in frame][1:] == [1] * (f.flush_count - 1)
        hip_ctx.render_stream(s, t)
        out = hip_ctx.read_pixels(t, w, h)
        ref, _ = f.render()
    finally:
        hip_ctx.free_target(t)
        f.close()
    if name == "paths":
        assert np.array_equal(out, ref)
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"max {d.max()}, frac>2 {(d > 2).mean()}"
