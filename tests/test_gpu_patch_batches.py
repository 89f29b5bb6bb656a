"""SPANS flushes described the way the reference's RenderContext hands a flush
to its backend: TessVertexSpans, ContourData and the FlushDescriptor's
DrawBatch list (runs of midpoint-fan patches, gpu.hpp:1219-1281) instead of
per-draw records.  The device derives each path's tessellation range from
its spans (each span names its contour, whose ContourData names the path);
the batches supply image paints.  The frame must equal the same layout
flushed with per-draw rive_hip_draw records, bit for bit, whether the batch
list is one run or cut into many.
"""
import ctypes

import numpy as np
import pytest

import oracle_binding as ob
import rivehip

pytestmark = pytest.mark.gpu

CASES = [(1, 40, 256, 256), (2, 120, 480, 270), (2, 300, 640, 360)]


def patch_batches(draws, every):
    """DrawBatch-like runs: the draws' patch ranges, merged `every` draws at a
    time (a batch ends where the next draw does not follow on)."""
    out = []
    for k in range(0, len(draws), every):
        grp = draws[k:k + every]
        first = int(grp[0][1]) // 8
        end = (int(grp[-1][1]) + int(grp[-1][2])) // 8
        out.append((first, end - first))
    arr = (rivehip.PatchBatch * len(out))()
    for i, (f, n) in enumerate(out):
        arr[i].first_patch, arr[i].patch_count = f, n
    return arr


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"cfg{c[0]}-n{c[1]}-{c[2]}x{c[3]}")
def test_patch_batches_match_draw_records(hip_ctx, case):
    cfg, n, w, h = case
    print(f"[patch-batches] cfg{cfg} {n} draws {w}x{h}", flush=True)
    s = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
    f = ob.OracleFlush(s)
    t = hip_ctx.target(w, h)
    try:
        assert f.flush_count == 1
        f.flush_spans_on(hip_ctx, t)
        ref = hip_ctx.read_pixels(t, w, h)
        ref_stats = hip_ctx.last_flush_stats()
        draws = f.draws()
        for every in (len(draws), 7, 1):
            batches = patch_batches(draws, every)
            d = rivehip.FlushDesc()
            d.target = t
            d.mode = rivehip.FLUSH_SPANS
            d.dither = 1
            d.path_count = f.path_count
            d.contour_count = f.contour_count
            d.tess_span_count = f.span_count
            d.tess_vertex_count = f.tess_vertex_count
            d.grad_span_count = len(f.get(5))
            d.grad_simple_rows = f.grad_simple_rows
            d.grad_complex_rows = f.grad_complex_rows
            d.patch_batches = ctypes.cast(batches, ctypes.c_void_p)
            d.patch_batch_count = len(batches)
            hip_ctx.flush(d)
            out = hip_ctx.read_pixels(t, w, h)
            st = hip_ctx.last_flush_stats()
            assert np.array_equal(out, ref), f"batches of {every} draws"
            assert st.prim_count == ref_stats.prim_count and st.bin_entries == ref_stats.bin_entries
            assert st.draw_count == len(draws)
    finally:
        hip_ctx.free_target(t)
        f.close()
        rivehip.free_stream(s)


def test_patch_batches_outside_the_texture_are_rejected(hip_ctx):
    s = rivehip.generate_stream(1, width=64, height=64, draw_count=4)
    f = ob.OracleFlush(s)
    t = hip_ctx.target(64, 64)
    try:
        f.flush_spans_on(hip_ctx, t)
        batches = (rivehip.PatchBatch * 1)()
        batches[0].first_patch, batches[0].patch_count = f.tess_vertex_count // 8, 1000
        d = rivehip.FlushDesc()
        d.target = t
        d.mode = rivehip.FLUSH_SPANS
        d.path_count = f.path_count
        d.contour_count = f.contour_count
        d.tess_span_count = f.span_count
        d.tess_vertex_count = f.tess_vertex_count
        d.patch_batches = ctypes.cast(batches, ctypes.c_void_p)
        d.patch_batch_count = 1
        with pytest.raises(RuntimeError) as e:
            hip_ctx.flush(d)
        assert "outside the tessellation texture" in str(e.value)
        f.flush_spans_on(hip_ctx, t)  # the context still renders
    finally:
        hip_ctx.free_target(t)
        f.close()
        rivehip.free_stream(s)
