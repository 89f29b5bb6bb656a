"""Seeded random scenes through the C ABI on the device, against the oracle:
odd canvas sizes, polygons and cubics with fractional coordinates that land on
and around 16-pixel tile boundaries, long near-horizontal and near-vertical
edges, strokes with every join and cap, clips, gradients and transforms -- the
cases where a primitive's tile set and sample rows are trimmed to the tile's
region.

Pixels within 2/255 on all but 0.1% of pixels and 8/255 everywhere.
"""
import numpy as np
import pytest

import oracle_binding as ob
from streams import BEVEL, BUTT, CLOCKWISE, EVENODD, MITER, NONZERO, ROUND, ROUND_CAP, SQUARE, StreamBuilder

pytestmark = pytest.mark.gpu


def random_path(rng, w, h):
    n = int(rng.integers(2, 9))
    cmds = []
    # points snapped near tile lines half of the time
    def pt():
        if rng.random() < .5:
            x = 16 * int(rng.integers(0, w // 16 + 2)) + rng.choice([-1, -.5, -1 / 256, 0, 1 / 256, .5, 1])
            y = 16 * int(rng.integers(0, h // 16 + 2)) + rng.choice([-1, -.5, -1 / 256, 0, 1 / 256, .5, 1])
        else:
            x, y = rng.uniform(-8, w + 8), rng.uniform(-8, h + 8)
        return float(x), float(y)
    x0, y0 = pt()
    cmds.append((0, x0, y0))
    for _ in range(n):
        r = rng.random()
        if r < .2:  # a long near-horizontal or near-vertical edge
            x, y = pt()
            if rng.random() < .5:
                cmds.append((1, x, y0 + rng.uniform(-.7, .7)))
            else:
                cmds.append((1, x0 + rng.uniform(-.7, .7), y))
        elif r < .6:
            cmds.append((1, *pt()))
        else:
            cmds.append((4, *pt(), *pt(), *pt()))
        x0, y0 = cmds[-1][-2], cmds[-1][-1]
    return cmds


def random_scene(seed):
    rng = np.random.default_rng(seed)
    w, h = int(rng.integers(17, 100)), int(rng.integers(17, 90))
    b = StreamBuilder(w, h, clear_color=int(rng.integers(0, 1 << 32)) | 0xff000000, dither=int(seed % 2))
    for _ in range(int(rng.integers(8, 30))):
        cmds = random_path(rng, w, h)
        stroke = rng.random() < .3
        rule = [NONZERO, EVENODD, CLOCKWISE][int(rng.integers(0, 3))]
        p = b.path(cmds + ([] if stroke else [(5,)]), fill_rule=rule)
        color = int(rng.integers(0, 1 << 24)) | (int(rng.integers(96, 256)) << 24)
        kw = {}
        if stroke:
            kw = dict(stroke=float(rng.uniform(.5, 9)), join=[MITER, ROUND, BEVEL][int(rng.integers(0, 3))],
                      cap=[BUTT, ROUND_CAP, SQUARE][int(rng.integers(0, 3))])
        if rng.random() < .25:
            kw["gradient"] = ("linear", float(rng.uniform(0, w)), float(rng.uniform(0, h)), float(rng.uniform(0, w)),
                              float(rng.uniform(0, h)), [(0, color), (1, color ^ 0x00ffffff)])
        if rng.random() < .15:
            kw["clips"] = (b.path(random_path(rng, w, h) + [(5,)]),)
        if rng.random() < .2:
            a = float(rng.uniform(-.6, .6))
            kw["matrix"] = (float(np.cos(a)), float(np.sin(a)), float(-np.sin(a)), float(np.cos(a)),
                            float(rng.uniform(-5, 5)), float(rng.uniform(-5, 5)))
        b.draw(p, color=color, **kw)
    return b


@pytest.mark.parametrize("seed", range(48))
def test_random_scene_matches_oracle(hip_ctx, seed):
    print(f"[random-scene] seed {seed}", flush=True)
    b = random_scene(seed)
    s = b.build()
    w, h = b.width, b.height
    f = ob.OracleFlush(s)
    t = hip_ctx.target(w, h)
    try:
        ref, _ = f.render()
        hip_ctx.render_stream(s, t)
        out = hip_ctx.read_pixels(t, w, h)
    finally:
        hip_ctx.free_target(t)
        f.close()
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"seed {seed}: max {d.max()}, frac>2 {(d > 2).mean()}"


def dense_scene(seed):
    """Many overlapping long paths on a larger canvas: every tile lists more
    paths than one batch holds, so coverage flows through several batches and
    listing windows per tile."""
    rng = np.random.default_rng(1000 + seed)
    w, h = 160, 144
    b = StreamBuilder(w, h, clear_color=0xff202020, dither=1)
    for _ in range(120):
        cmds = [(0, float(rng.uniform(0, w)), float(rng.uniform(0, h)))]
        for _ in range(int(rng.integers(10, 40))):
            if rng.random() < .5:
                cmds.append((1, float(rng.uniform(0, w)), float(rng.uniform(0, h))))
            else:
                cmds.append((4, *(float(v) for v in rng.uniform(0, [w, h, w, h, w, h]))))
        stroke = rng.random() < .25
        p = b.path(cmds + ([] if stroke else [(5,)]), fill_rule=[NONZERO, EVENODD][int(rng.integers(0, 2))])
        color = int(rng.integers(0, 1 << 24)) | (int(rng.integers(128, 256)) << 24)
        if stroke:
            b.draw(p, color=color, stroke=float(rng.uniform(1, 6)), join=ROUND, cap=ROUND_CAP)
        else:
            b.draw(p, color=color)
    return b


@pytest.mark.parametrize("seed", range(4))
def test_dense_random_scene_matches_oracle(hip_ctx, seed):
    print(f"[random-scene] dense seed {seed}", flush=True)
    b = dense_scene(seed)
    s = b.build()
    w, h = b.width, b.height
    f = ob.OracleFlush(s)
    t = hip_ctx.target(w, h)
    try:
        ref, _ = f.render()
        hip_ctx.render_stream(s, t)
        out = hip_ctx.read_pixels(t, w, h)
    finally:
        hip_ctx.free_target(t)
        f.close()
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"dense seed {seed}: max {d.max()}, frac>2 {(d > 2).mean()}"
