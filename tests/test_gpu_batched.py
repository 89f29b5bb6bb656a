"""Batched flush through the C ABI: several canvases rendered by one flush
(rive_hip_render_streams, rive_hip_canvas) must equal each canvas rendered by
its own flush, bit for bit, and the oracle within the pixel tolerance of the
single-canvas tests.  The canvases differ in size (odd sizes, one pixel, no
draws), clear color, clips (clip IDs restart per canvas) and paints.
"""
import numpy as np
import pytest

import oracle_binding as ob
import rivehip
from streams import StreamBuilder, rect
from test_gpu_edge_cases import sc_clip_rects, sc_clips, sc_gradients, star

pytestmark = pytest.mark.gpu


def small_canvas(w, h, clear, seed):
    b = StreamBuilder(w, h, clear_color=clear, dither=1)
    rng = np.random.default_rng(seed)
    for k in range(6):
        x0, y0 = rng.uniform(-5, w * .6), rng.uniform(-5, h * .6)
        b.draw(b.path(star(x0 + w * .2, y0 + h * .2, 2 + w * .05, 4 + w * .2, 5 + k)),
               color=0x80000000 | int(rng.integers(0, 1 << 24)))
    b.draw(b.path(rect(1, 1, w - 1, h - 1)), color=0x60ffffff, stroke=1.5)
    return b


def builders():
    return [small_canvas(37, 29, 0xff000000, 1), sc_clips(), StreamBuilder(5, 3, clear_color=0xff00ff00),
            sc_gradients(), small_canvas(1, 1, 0x40404040, 2), sc_clip_rects(), small_canvas(130, 70, 0, 3)]


def test_batched_equals_single_and_oracle(hip_ctx):
    print("[batched] 7 canvases + 3 generated", flush=True)
    bs = builders()
    streams = [b.build() for b in bs]
    sizes = [(b.width, b.height) for b in bs]
    generated = [rivehip.generate_stream(1, seed=0x5249564500000040 + i, width=w, height=h, draw_count=n)
                 for i, (w, h, n) in enumerate([(96, 64, 12), (64, 200, 8), (160, 90, 20)])]
    streams += generated
    sizes += [(96, 64), (64, 200), (160, 90)]
    targets = [hip_ctx.target(w, h) for w, h in sizes]
    singles = [hip_ctx.target(w, h) for w, h in sizes]
    try:
        hip_ctx.render_streams(streams, targets)
        canvases, count = hip_ctx.last_flush_canvases()
        assert count == len(streams)
        firsts = [canvases[i].first_path for i in range(count)]
        assert firsts == sorted(firsts) and firsts[0] == 1
        batched = [hip_ctx.read_pixels(t, w, h) for t, (w, h) in zip(targets, sizes)]
        for s, t in zip(streams, singles):
            hip_ctx.render_stream(s, t)
        for i, (t, (w, h)) in enumerate(zip(singles, sizes)):
            single = hip_ctx.read_pixels(t, w, h)
            assert np.array_equal(batched[i], single), f"canvas {i} differs from its own flush"
        for i, s in enumerate(streams):
            f = ob.OracleFlush(s)
            try:
                ref, _ = f.render()
            finally:
                f.close()
            d = np.abs(batched[i].astype(np.int32) - ref.astype(np.int32)).max(axis=2)
            assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"canvas {i}: max {d.max()}"
    finally:
        for t in targets + singles:
            hip_ctx.free_target(t)
        for g in generated:
            rivehip.free_stream(g)


def test_batched_reflush_from_resident_records(hip_ctx):
    """Flushing the last batched flush again from HBM-resident records gives
    the same pixels."""
    print("[batched] reflush", flush=True)
    streams = rivehip.generate_batch(8, seed=0x5249564500000041)
    try:
        sizes = [(s.contents.width, s.contents.height) for s in streams]
        targets = [hip_ctx.target(w, h) for w, h in sizes]
        try:
            hip_ctx.render_streams(streams, targets)
            first = [hip_ctx.read_pixels(t, w, h) for t, (w, h) in zip(targets, sizes)]
            stats = hip_ctx.last_flush_stats()
            canv = hip_ctx.last_flush_canvases()
            hip_ctx.reflush_geometry(None, stats, canvases=canv)
            hip_ctx.sync()
            again = [hip_ctx.read_pixels(t, w, h) for t, (w, h) in zip(targets, sizes)]
            for a, b in zip(first, again):
                assert np.array_equal(a, b)
            assert stats.tile_count == sum(((w + 15) // 16) * ((h + 15) // 16) for w, h in sizes)
        finally:
            for t in targets:
                hip_ctx.free_target(t)
    finally:
        for s in streams:
            rivehip.free_stream(s)
