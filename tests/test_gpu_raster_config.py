"""Raster launch configurations and the raster autotuner.

The raster pass can take its tiles in canvas/row/column order or heaviest
first (a per-tile cost estimate from the binning), run one workgroup per tile
or a resident grid pulling tiles from a device queue, and be compiled for 5 or
6 waves per SIMD.  Tiles are rendered independently, so every configuration
must give the same pixels bit for bit: on a 1080p-mix frame, on batched
canvases of uneven size and on a feathered 4K-mix frame (reduced sizes).  The
autotuner (rive_hip_autotune_frame) must keep one of the candidates, report a
time for each, and leave the frame's targets as they were.
"""
import numpy as np
import pytest

import rivehip

pytestmark = pytest.mark.gpu

CONFIGS = [(o, p, w) for o in (0, 1) for p in (0, 1) for w in (6, 5)]


@pytest.fixture
def ctx():
    # a context of its own, so no config outlives the test
    c = rivehip.Context(0)
    yield c
    c.close()


def _frame_under_configs(ctx, stream):
    w, h = stream.contents.width, stream.contents.height
    t = ctx.target(w, h)
    out = {}
    try:
        for cfg in CONFIGS:
            ctx.set_raster_config(*cfg)
            ctx.render_stream(stream, t)
            out[cfg] = ctx.read_pixels(t, w, h)
    finally:
        ctx.free_target(t)
    return out


@pytest.mark.parametrize("config,w,h,n", [(2, 640, 360, 600), (3, 960, 540, 800)])
def test_configs_render_the_same_frame(ctx, config, w, h, n):
    print(f"[raster-config] config-{config} stream {w}x{h} {n} draws under {len(CONFIGS)} configs", flush=True)
    s = rivehip.generate_stream(config, width=w, height=h, draw_count=n)
    try:
        px = _frame_under_configs(ctx, s)
    finally:
        rivehip.free_stream(s)
    base = px[CONFIGS[0]]
    assert base[..., 3].max() > 0
    for cfg, p in px.items():
        assert np.array_equal(p, base), f"config {cfg} differs from {CONFIGS[0]}"


def test_configs_render_the_same_batch(ctx):
    print("[raster-config] 24 batched canvases under every config", flush=True)
    streams = rivehip.generate_batch(24)
    sizes = [(s.contents.width, s.contents.height) for s in streams]
    groups = rivehip.flush_groups(streams)
    targets = [ctx.target(w, h) for w, h in sizes]
    try:
        base = None
        for cfg in CONFIGS:
            ctx.set_raster_config(*cfg)
            for grp in groups:
                ctx.render_streams([streams[i] for i in grp], [targets[i] for i in grp])
            px = [ctx.read_pixels(t, w, h) for t, (w, h) in zip(targets, sizes)]
            if base is None:
                base = px
                assert sum(int(p[..., 3].max() > 0) for p in px) == len(px)
            for i, (a, b) in enumerate(zip(px, base)):
                assert np.array_equal(a, b), f"config {cfg}: canvas {i} {sizes[i]} differs"
    finally:
        for t in targets:
            ctx.free_target(t)
        for s in streams:
            rivehip.free_stream(s)


def test_autotune_frame(ctx):
    print("[raster-config] autotune on a 960x540 config-2 frame", flush=True)
    s = rivehip.generate_stream(2, width=960, height=540, draw_count=1500)
    t = ctx.target(960, 540)
    try:
        ctx.render_stream(s, t)
        before = ctx.read_pixels(t, 960, 540)
        frame = ctx.last_frame()
        chosen, times = ctx.autotune_frame(frame, trials=2)
        assert (chosen["tile_order"], chosen["persistent"], chosen["waves_per_simd"]) in CONFIGS
        assert sorted(times) == sorted(CONFIGS) and all(v > 0 for v in times.values())
        best = min(times, key=times.get)
        assert times[(chosen["tile_order"], chosen["persistent"], chosen["waves_per_simd"])] == times[best]
        assert ctx.raster_config() == chosen
        assert np.array_equal(ctx.read_pixels(t, 960, 540), before), "autotune left the target changed"
        ctx.flush_frame(frame, t)
        assert np.array_equal(ctx.read_pixels(t, 960, 540), before)
    finally:
        ctx.free_target(t)
        rivehip.free_stream(s)


def test_raster_config_rejects_bad_values(ctx):
    for bad in [(2, 0, 6), (0, 2, 6), (0, 0, 7), (0, 0, 4)]:
        with pytest.raises(RuntimeError):
            ctx.set_raster_config(*bad)
    assert ctx.raster_config() == {"tile_order": 0, "persistent": 0, "waves_per_simd": 6}


def test_graph_replays_with_a_persistent_ordered_raster(ctx):
    # The heaviest-first order and the persistent queue inside a captured
    # frame: the queue head and class counters are cleared by fill kernels in
    # the graph, so every replay takes every tile once and equals the eager
    # frame bit for bit.
    print("[raster-config] frame graph with tile order + persistent raster", flush=True)
    s = rivehip.generate_stream(2, width=960, height=540, draw_count=1500)
    t = ctx.target(960, 540)
    g = None
    try:
        ctx.set_raster_config(1, 1, 6)
        ctx.render_stream(s, t)
        frame = ctx.last_frame()
        ctx.flush_frame(frame, t)
        ctx.sync()
        eager = ctx.read_pixels(t, 960, 540)
        g = ctx.frame_graph(frame, t)
        for _ in range(3):
            ctx.render_stream(s, t)  # eager work between replays
            ctx.sync()
            ctx.frame_graph_launch(g)
            ctx.sync()
            assert np.array_equal(ctx.read_pixels(t, 960, 540), eager)
    finally:
        if g is not None:
            ctx.frame_graph_destroy(g)
        ctx.free_target(t)
        rivehip.free_stream(s)
