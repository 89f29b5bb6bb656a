"""Host-side checks that need no GPU: the C ABI library loads and exports
every function include/rive_hip.h declares, the seeded stream generator is
deterministic and follows the workload's distribution, the oracle's flush
layout keeps the reference's invariants (render_context.cpp:1150-1183,
3111-3115, 3259-3301; gpu.hpp:291-375), and bench.py refuses N > 1 GPUs.
"""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_binding as ob
import rivehip
from streams import _DRAW, _PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rive_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rive_(?:hip|stream|scene|ease|cmd|image|png|jpeg)_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = header_functions()
    assert len(names) >= 17
    L = rivehip.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(rivehip.EXPORTED) <= set(names)


def view(ptr, dtype, n):
    if n == 0:
        return np.zeros(0, dtype)
    raw = (ctypes.c_uint8 * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(raw, dtype=dtype).copy()


def stream_arrays(p):
    s = p.contents
    return {
        "verbs": view(s.verbs, np.uint8, s.verb_count),
        "points": view(s.points, np.float32, 2 * s.point_count),
        "paths": view(s.paths, _PATH, s.path_count),
        "draws": view(s.draws, _DRAW, s.draw_count),
        "stop_colors": view(s.stop_colors, np.uint32, s.stop_count),
        "stop_offsets": view(s.stop_offsets, np.float32, s.stop_count),
        "header": (s.width, s.height, s.clear_color, s.dither),
    }


def test_generator_is_deterministic_and_seeded():
    a = stream_arrays(p := rivehip.generate_stream(2, draw_count=300))
    rivehip.free_stream(p)
    b = stream_arrays(p := rivehip.generate_stream(2, draw_count=300))
    rivehip.free_stream(p)
    c = stream_arrays(p := rivehip.generate_stream(2, seed=12345, draw_count=300))
    rivehip.free_stream(p)
    for k in a:
        if k == "header":
            assert a[k] == b[k]
        else:
            assert np.array_equal(a[k], b[k]), k
    assert not np.array_equal(a["points"][:100], c["points"][:100])


def test_generator_prefix_property():
    """The first n draws of the 5000-draw stream are the n-draw stream."""
    full = stream_arrays(p := rivehip.generate_stream(2))
    rivehip.free_stream(p)
    part = stream_arrays(p := rivehip.generate_stream(2, draw_count=100))
    rivehip.free_stream(p)
    assert np.array_equal(full["draws"][:100], part["draws"])
    n_pts = len(part["points"])
    assert np.array_equal(full["points"][:n_pts], part["points"])


def test_config1_distribution():
    s = stream_arrays(p := rivehip.generate_stream(1))
    rivehip.free_stream(p)
    assert s["header"] == (1024, 1024, 0, 1)
    d, paths = s["draws"], s["paths"]
    assert len(d) == 200 and np.all(d["style"] == 0) and np.all(d["paint_type"] == 0)
    cubics = (paths["verb_count"] - 2)  # move + n cubics + close
    assert cubics.min() >= 8 and cubics.max() <= 32
    alpha = d["color"] >> 24
    assert alpha.min() >= 127 and alpha.max() <= 255
    pts = s["points"]
    assert pts.min() >= 0 and pts.max() < 1024


def test_config2_distribution():
    s = stream_arrays(p := rivehip.generate_stream(2))
    rivehip.free_stream(p)
    assert s["header"] == (1920, 1080, 0, 1)
    d, paths = s["draws"], s["paths"]
    assert len(d) == 5000
    main = paths[d["path"]]
    stroke = d["style"] == 1
    nonzero = (~stroke) & (main["fill_rule"] == 0)
    evenodd = (~stroke) & (main["fill_rule"] == 1)
    assert abs(nonzero.mean() - .6) < .03 and abs(evenodd.mean() - .2) < .03
    assert abs(stroke.mean() - .2) < .03
    w = d["thickness"][stroke]
    assert w.min() >= 1 and w.max() <= 20
    assert abs((d["paint_type"] != 0).mean() - .3) < .03
    assert abs((d["clip_count"] == 2).mean() - .1) < .02
    assert set(np.unique(d["clip_count"])) <= {0, 2}
    # gradient stops: sorted, first 0, last 1, 2..8 of them
    for g in d[d["paint_type"] != 0][:200]:
        offs = s["stop_offsets"][g["first_stop"]:g["first_stop"] + g["stop_count"]]
        assert 2 <= len(offs) <= 8 and offs[0] == 0 and offs[-1] == 1
        assert np.all(np.diff(offs) >= 0)


@pytest.mark.parametrize("cfg,n,w,h", [(1, 40, 512, 512), (2, 300, 960, 540)])
def test_oracle_layout_invariants(cfg, n, w, h):
    p = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
    f = ob.OracleFlush(p)
    try:
        draws = f.draws()
        paths = f.get(2)
        contours = f.get(1)
        spans = f.get(0)
        paints = f.get(3)
    finally:
        f.close()
        rivehip.free_stream(p)
    # path record 0 is reserved; draws reference 1..pathCount in order
    assert not paths[0].tobytes().strip(b"\0")
    assert np.array_equal(draws[:, 0], np.arange(1, len(draws) + 1))
    # midpoint-fan vertices start after 8 padding vertices and are contiguous
    loc, cnt = draws[:, 1], draws[:, 2]
    assert loc[0] == 8 and np.array_equal(loc[1:], loc[:-1] + cnt[:-1])
    assert f.tess_vertex_count == loc[-1] + cnt[-1] + 1
    # each contour is padded to a multiple of 8 vertices, doubled when drawn
    # from both directions
    double = (draws[:, 3] == 2) | (draws[:, 3] == 3)
    assert np.all(cnt[~double] % 8 == 0) and np.all(cnt[double] % 16 == 0)
    # contour IDs are dense from 1 and every contour's first vertex lies in
    # its draw's range
    assert np.array_equal(contours["path_id"], np.sort(contours["path_id"]))
    for c in contours:
        i = int(c["path_id"]) - 1
        assert loc[i] <= c["vertex_index0"] < loc[i] + cnt[i]
    # spans: 2 padding spans first, then counts per span packed as
    # join << 20 | polar << 10 | parametric; every span has >= 1 join
    # segment, and only emulated caps have no parametric segments
    segs = spans["segment_counts"][2:]
    flags = spans["contour_id_with_flags"][2:]
    assert np.all((segs >> 20) >= 1)
    assert np.all(((segs & 0x3ff) >= 1) | ((flags & (1 << 25)) != 0))
    cid = spans["contour_id_with_flags"][2:] & 0xffff
    assert cid.min() >= 1 and cid.max() == len(contours)
    # fill-rule flags sit only on fills; clip-update paints have type 0
    types = paints["params"][1:] & 0xf
    assert set(np.unique(types)) <= {0, 1, 2, 3}


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_batch_generator_is_seeded_and_in_range():
    """The batched workload's canvases: log-uniform sides 64..1024 with aspect
    .5..2 and 10..2000 closed cubic fills each, reproducible from the seed."""
    import rivehip

    a = rivehip.generate_batch(24, seed=123)
    b = rivehip.generate_batch(24, seed=123)
    try:
        for sa, sb in zip(a, b):
            x, y = sa.contents, sb.contents
            assert (x.width, x.height, x.draw_count) == (y.width, y.height, y.draw_count)
            assert 45 <= x.width <= 1449 and 45 <= x.height <= 1449
            assert 0.5 * 0.98 <= x.width / x.height <= 2.0 * 1.02
            assert 10 <= x.draw_count <= 2000
            n = x.point_count * 2
            fp = ctypes.POINTER(ctypes.c_float)
            pa = np.ctypeslib.as_array(ctypes.cast(x.points, fp), (n,))
            pb = np.ctypeslib.as_array(ctypes.cast(y.points, fp), (n,))
            assert np.array_equal(pa, pb)
            assert pa[0::2].min() >= 0 and pa[0::2].max() < x.width
        assert len({(s.contents.width, s.contents.height) for s in a}) > 20
    finally:
        for s in a + b:
            rivehip.free_stream(s)
