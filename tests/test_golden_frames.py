"""The stored full-size oracle frames (tests/golden/frame_*.npy.xz, written
by tests/golden/make_frame_fixtures.py) still match the workload generator
and the oracle: the 1024x1024 200-path frame is re-rendered and must equal
its fixture bit for bit (pixels and anti-aliased-edge mask); for the
1920x1080 5000-draw frame, whose render takes minutes, the oracle's layout of
the stream (six logical flushes) must hash to the value recorded with the
fixture -- the raster half of the oracle is the one the 1024x1024 frame
exercises.
"""
import json
import os

import numpy as np

import oracle_binding as ob
import rivehip
from golden.make_frame_fixtures import FRAMES, layout_digest, load_xz

HERE = os.path.dirname(os.path.abspath(__file__))
META = json.load(open(os.path.join(HERE, "golden", "frames.json")))


def fixture(name):
    m = META[name]
    img = load_xz(os.path.join(HERE, "golden", f"frame_{name}.npy.xz"))
    bits = load_xz(os.path.join(HERE, "golden", f"edges_{name}.npy.xz"))
    edges = np.unpackbits(bits)[: m["width"] * m["height"]].reshape(m["height"], m["width"]).astype(bool)
    return m, img, edges


def test_cfg1_fixture_matches_a_fresh_oracle_render():
    m, img, edges = fixture("cfg1")
    cfg, w, h, n = FRAMES["cfg1"]
    s = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
    f = ob.OracleFlush(s)
    try:
        assert f.flush_count == m["logical_flushes"] == 1
        assert layout_digest(f) == m["layout_sha256"]
        out, mask, _ = f.render_edges()
    finally:
        f.close()
        rivehip.free_stream(s)
    assert out.shape == img.shape and np.array_equal(out, img)
    assert np.array_equal(mask, edges)


def test_cfg2_fixture_matches_the_current_layout():
    m, img, edges = fixture("cfg2")
    cfg, w, h, n = FRAMES["cfg2"]
    assert img.shape == (h, w, 4) and edges.shape == (h, w)
    s = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
    f = ob.OracleFlush(s)
    try:
        assert f.flush_count == m["logical_flushes"] == 6
        assert layout_digest(f) == m["layout_sha256"]
    finally:
        f.close()
        rivehip.free_stream(s)
