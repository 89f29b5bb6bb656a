"""The serving loop on the device (rive_hip_serve).

* A scene frame (the draw stream of an animated scene at time t) matches
  the oracle within the pixel tolerance.
* Alone vs in the mix: a request rendered in a batch of one and the same
  request rendered inside a batch of six mixed requests (other scenes,
  times and sizes, all admitted together) give identical pixels.
* Admission: with a budget of about two requests' estimated HBM and at most
  three slots, every batch of more than one request stays within both,
  requests are admitted in arrival order, and arrival <= admit <= done.
"""
import numpy as np
import pytest

import oracle_binding as ob
import rivehip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes():
    sc = [rivehip.Scene(s) for s in (101, 102, 103)]
    yield sc
    for s in sc:
        s.close()


@pytest.mark.parametrize("si,t,w,h", [(0, 0.0, 256, 256), (1, 1.7, 512, 512), (2, 3.3, 320, 200)])
def test_scene_frame_matches_oracle(hip_ctx, scenes, si, t, w, h):
    print(f"[serving] scene {si} t={t} {w}x{h}", flush=True)
    s = scenes[si].stream(t, w, h)
    f = ob.OracleFlush(s)
    tg = hip_ctx.target(w, h)
    try:
        ref, _ = f.render()
        hip_ctx.render_stream(s, tg)
        out = hip_ctx.read_pixels(tg, w, h)
    finally:
        hip_ctx.free_target(tg)
        f.close()
        rivehip.free_stream(s)
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    assert d.max() <= 8 and (d > 2).mean() <= 1e-3, f"max {d.max()}, frac>2 {(d > 2).mean()}"


def test_request_alone_equals_request_in_mix(hip_ctx, scenes):
    print("[serving] alone vs mix", flush=True)
    target = (1, 1.1, 384, 256)
    alone = np.zeros((256, 384, 4), np.uint8)
    res = hip_ctx.serve(scenes, [target + (0.0,)], 1 << 40, 8, pixels={0: alone})
    assert res[0].batch_size == 1
    mix = [(0, 0.4, 256, 256), (2, 2.0, 512, 300), (1, 3.9, 200, 200), target, (0, 1.1, 384, 256),
           (2, 0.0, 300, 512)]
    got = np.zeros((256, 384, 4), np.uint8)
    res = hip_ctx.serve(scenes, [m + (0.0,) for m in mix], 1 << 40, 8, pixels={3: got})
    assert {r.batch_size for r in res} == {6}
    assert alone.any()
    assert np.array_equal(alone, got)
    # and the single-canvas front-end path renders the same pixels
    s = scenes[1].stream(1.1, 384, 256)
    tg = hip_ctx.target(384, 256)
    try:
        hip_ctx.render_stream(s, tg)
        plain = hip_ctx.read_pixels(tg, 384, 256)
    finally:
        hip_ctx.free_target(tg)
        rivehip.free_stream(s)
    assert np.array_equal(plain, alone)


def test_admission_budget_slots_and_order(hip_ctx, scenes):
    print("[serving] admission", flush=True)
    sizes = [(256, 256), (512, 512), (320, 240)]
    reqs = [(i % 3, 0.25 * i, *sizes[i % 3], 0.002 * i) for i in range(12)]
    budget = 2 * max(scenes[i % 3].request_bytes(*sizes[i % 3]) for i in range(3))
    res = hip_ctx.serve(scenes, reqs, budget, 3)
    batches = {}
    for i, r in enumerate(res):
        assert r.arrival_s <= r.admit_s <= r.done_s
        batches.setdefault(r.batch, []).append(i)
    prev_last = -1
    for b in sorted(batches):
        members = batches[b]
        assert members == list(range(members[0], members[0] + len(members)))  # FIFO, contiguous
        assert members[0] == prev_last + 1
        prev_last = members[-1]
        assert len(members) <= 3
        if len(members) > 1:
            est = sum(scenes[reqs[i][0]].request_bytes(reqs[i][2], reqs[i][3]) for i in members)
            assert est <= budget and res[members[0]].batch_bytes == est
