"""rive_hip_flush_noalloc: the flush variant that allocates nothing, checked
bit for bit against rive_hip_flush on the same resident records.

Two contexts hold two draw streams (the second smaller); their no-alloc
flushes run back to back on the one GPU, so the second reuses the scratch
array the first left dirty.  A flush too large for the scratch (the first
logical flush of the 1920x1080 5000-draw frame) is refused with
RIVE_HIP_ERROR_OUT_OF_MEMORY, and the allocating flushes still re-render the
frame afterwards.
"""
import ctypes

import numpy as np
import pytest

import rivehip

pytestmark = pytest.mark.gpu

OUT_OF_MEMORY = 3


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


def test_noalloc_flush_matches_flush():
    cases = [(2, 640, 360, 400), (1, 256, 256, 40)]  # the second is smaller
    ctxs, streams, targets, expect, stats = [], [], [], [], []  # targets: (ctx, target)
    try:
        for cfg, w, h, n in cases:
            print(f"[noalloc] regular flush cfg{cfg} {w}x{h} {n} draws", flush=True)
            ctx = rivehip.Context(0)
            ctxs.append(ctx)
            s = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
            streams.append(s)
            t = ctx.target(w, h)
            targets.append((ctx, t))
            ctx.render_stream(s, t)
            expect.append(ctx.read_pixels(t, w, h))
            stats.append(ctx.last_flush_stats())
        outs = []
        for (cfg, w, h, n), ctx, st in zip(cases, ctxs, stats):
            print(f"[noalloc] no-alloc flush cfg{cfg} {w}x{h} {n} draws", flush=True)
            t2 = ctx.target(w, h)  # zero-filled: the flush must write every pixel
            try:
                d = desc_from(st, t2)
                rivehip.check(ctx.L.rive_hip_flush_noalloc(ctx.h, ctypes.byref(d)))
                outs.append(ctx.read_pixels(t2, w, h))
                again = ctx.last_flush_stats()
                assert again.prim_count == st.prim_count and again.bin_entries == st.bin_entries
                assert again.span_count == st.span_count
            finally:
                ctx.free_target(t2)
        for e, o in zip(expect, outs):
            assert e.shape == o.shape and np.array_equal(e, o)
            assert o[..., 3].max() > 0
        # intermediates of a no-alloc flush are not kept
        size = ctypes.c_size_t()
        assert ctxs[1].L.rive_hip_debug_readback(ctxs[1].h, 0, None, 0, ctypes.byref(size)) != 0

        # a frame beyond the scratch is refused; the allocating flush still works
        print("[noalloc] 1920x1080 5000-draw frame must be refused", flush=True)
        big = rivehip.generate_stream(2)
        streams.append(big)
        ctx = ctxs[0]
        tb = ctx.target(1920, 1080)
        targets.append((ctx, tb))
        ctx.render_stream(big, tb)
        ref = ctx.read_pixels(tb, 1920, 1080)
        frame = ctx.last_frame()  # six logical flushes sharing the record buffers
        d = frame[0]
        rc = ctx.L.rive_hip_flush_noalloc(ctx.h, ctypes.byref(d))
        assert rc == OUT_OF_MEMORY, rc
        assert b"no-alloc scratch" in ctx.L.rive_hip_last_error()
        ctx.reflush_frame(ctx.last_frame(), tb)
        assert np.array_equal(ctx.read_pixels(tb, 1920, 1080), ref)
    finally:
        for ctx, t in targets:
            ctx.free_target(t)
        for s in streams:
            rivehip.free_stream(s)
        for ctx in ctxs:
            ctx.close()
