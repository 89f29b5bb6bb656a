"""Error paths of the C ABI on the device: records that index outside their
buffers, draws naming paths or images that do not exist, a buffer left mapped,
and malformed canvas tables are rejected with an error code (host checks, or
the bounded checks of the counting and draw-range kernels) -- and the context
keeps working: a valid flush afterwards renders the same pixels as before.
As the reference signals failures (render_context_impl.hpp), nothing throws
across the boundary; rive_hip_last_error() names the problem.
"""
import ctypes

import numpy as np
import pytest

import rivehip
from streams import StreamBuilder, rect

pytestmark = pytest.mark.gpu


def scene():
    b = StreamBuilder(48, 40, clear_color=0xff101010, dither=1)
    b.draw(b.path(rect(4, 4, 30, 30)), color=0xc0ff4000)
    b.draw(b.path(rect(12, 10, 44, 36)), color=0x8000a0ff, stroke=3.0)
    b.draw(b.path(rect(0, 0, 48, 40)), color=0x40ffffff, clips=(b.path(rect(20, 2, 46, 20)),))
    return b


def desc_from(stats, target):
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


def expect_error(ctx, desc, words):
    with pytest.raises(RuntimeError) as e:
        ctx.flush(desc)
    assert words in str(e.value), str(e.value)


def test_invalid_flushes_are_rejected_and_context_recovers(hip_ctx):
    print("[errors] invalid flush descriptors", flush=True)
    b = scene()
    s = b.build()
    t = hip_ctx.target(b.width, b.height)
    try:
        hip_ctx.render_stream(s, t)
        good = hip_ctx.read_pixels(t, b.width, b.height)
        stats = hip_ctx.last_flush_stats()
        draws = hip_ctx.readback(3).view(rivehip.DRAW).copy()
        assert len(draws) == stats.draw_count

        d = desc_from(stats, t)
        d.path_count = stats.path_count + 50
        expect_error(hip_ctx, d, "buffers smaller")
        d = desc_from(stats, t)
        d.draw_count = stats.draw_count + 50
        expect_error(hip_ctx, d, "buffers smaller")

        bad = draws.copy()
        bad[0]["path_id"] = stats.path_count + 7
        hip_ctx.upload("draw", bad)
        expect_error(hip_ctx, desc_from(stats, t), "out of range")
        bad = draws.copy()
        bad[1]["first_verb"] = 1 << 20
        hip_ctx.upload("draw", bad)
        expect_error(hip_ctx, desc_from(stats, t), "outside the VERB/POINT buffers")
        bad = draws.copy()
        bad[0]["image_index"] = 3  # the flush has no images
        hip_ctx.upload("draw", bad)
        expect_error(hip_ctx, desc_from(stats, t), "out of range")
        hip_ctx.upload("draw", draws)

        # a buffer left mapped
        L = hip_ctx.L
        ptr = L.rive_hip_buffer_map(hip_ctx.h, rivehip.BUFFER["path"], 16)
        assert ptr
        expect_error(hip_ctx, desc_from(stats, t), "mapped buffer")
        rivehip.check(L.rive_hip_buffer_unmap(hip_ctx.h, rivehip.BUFFER["path"], 0))

        # malformed canvas tables
        canv = (rivehip.Canvas * 2)()
        canv[0].target, canv[0].first_path, canv[0].path_count = t.value, 1, 2
        canv[1].target, canv[1].first_path, canv[1].path_count = t.value, 2, 1  # overlaps
        d = desc_from(stats, None)
        d.canvases, d.canvas_count = ctypes.cast(canv, ctypes.c_void_p), 2
        expect_error(hip_ctx, d, "ascend")
        canv[1].first_path, canv[1].path_count = 3, stats.path_count  # beyond path_count
        expect_error(hip_ctx, d, "ascend")
        canv[1].target = None
        canv[1].path_count = 1
        expect_error(hip_ctx, d, "canvas target")

        # the context still renders the frame
        hip_ctx.reflush_geometry(t, stats)
        hip_ctx.sync()
        again = hip_ctx.read_pixels(t, b.width, b.height)
        assert np.array_equal(good, again)
    finally:
        hip_ctx.free_target(t)


def test_null_arguments_are_rejected(hip_ctx):
    L = hip_ctx.L
    assert L.rive_hip_flush(hip_ctx.h, None) != 0
    assert L.rive_hip_flush(None, None) != 0
    assert L.rive_hip_sync(None) != 0
    assert L.rive_hip_read_pixels(hip_ctx.h, None, None) != 0


def test_oversized_and_inconsistent_records_are_rejected(hip_ctx):
    """Counts that would overflow the per-path arrays are refused on the host;
    a SPANS flush whose draws have texels but no ContourData, or whose spans
    name contour IDs beyond contour_count, fails with an error instead of
    reading records that do not exist."""
    import oracle_binding as ob

    print("[errors] oversized path_count, missing and out-of-range contours", flush=True)
    b = scene()
    s = b.build()
    f = ob.OracleFlush(s)
    t = hip_ctx.target(b.width, b.height)
    try:
        ref, _ = f.render()
        hip_ctx.render_stream(s, t)
        stats = hip_ctx.last_flush_stats()
        d = desc_from(stats, t)
        d.path_count = (1 << 24) + 1
        expect_error(hip_ctx, d, "above 2^24")
        d.path_count = 0xFFFFFFFF
        expect_error(hip_ctx, d, "above 2^24")

        f.flush_spans_on(hip_ctx, t, clear_color=b.clear_color, dither=b.dither)  # a valid SPANS flush
        spans_desc = rivehip.FlushDesc()
        spans_desc.target = t
        spans_desc.mode = rivehip.FLUSH_SPANS
        spans_desc.dither = b.dither
        spans_desc.clear_color = b.clear_color
        spans_desc.path_count = f.path_count
        spans_desc.contour_count = 0
        spans_desc.tess_span_count = f.span_count
        spans_desc.tess_vertex_count = f.tess_vertex_count
        spans_desc.draw_count = f.draw_count
        spans_desc.grad_span_count = len(f.get(5))
        expect_error(hip_ctx, spans_desc, "no contours")

        spans = f.get(0).copy()
        live = (spans["contour_id_with_flags"] & 0xFFFF) != 0
        spans["contour_id_with_flags"][live] = (spans["contour_id_with_flags"][live] & ~np.uint32(0xFFFF)) | (
            f.contour_count + 5)
        hip_ctx.upload("tess_span", spans)
        spans_desc.contour_count = f.contour_count
        expect_error(hip_ctx, spans_desc, "contour IDs outside")

        f.flush_spans_on(hip_ctx, t, clear_color=b.clear_color, dither=b.dither)
        out = hip_ctx.read_pixels(t, b.width, b.height)
        d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max()
        assert d <= 8
    finally:
        hip_ctx.free_target(t)
        f.close()
