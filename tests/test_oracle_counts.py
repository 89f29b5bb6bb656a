"""Pins the oracle's integer layout -- per-curve segment counts, join and cap
counts, contour padding, tessellation locations and span row wrapping -- and
its tessellated vertices, which the device is compared against bit for bit.

Each check computes its expectation independently of the oracle's code:
  * segment counts from the reference's own maths compiled from its sources
    (oracle/_ref: wangs_formula::cubic_pow4, simd::fast_acos,
    math::calc_polar_segments_per_radian, Mat2D::findMaxScale, eval_cubic_at)
    and the integer rules of draw.cpp:1195-1342 (ceil of the 4th root,
    clamp to [1, 1023]; miter/bevel joins 5 segments; round caps
    min(ceil(spr*pi)+2, 1023), other caps 5);
  * hand-computed values for fixed radii and angles;
  * layout invariants of render_context.cpp:3079-3424 and draw.cpp:1358-1365
    (contours padded to multiples of 8, contiguous draws after 8 padding
    texels, spans covering exactly their records' texel runs);
  * closed forms of the tessellated geometry (points on the cubic at
    T = i/n, tangent angles, bounded turn per polar segment, mirrored runs
    equal to reversed forward runs).
Scenes are seeded: random cubic/line contours under random affine matrices,
fills with all three fill rules and strokes with every join and cap.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import oracle_binding as ob
import rivehip
from streams import (BEVEL, BUTT, MITER, ROUND, ROUND_CAP, SQUARE, StreamBuilder, contour_limit_stream,
                     gradient_limit_stream, pixel_rects_stream, tess_limit_stream)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_math.so")
need_ref = pytest.mark.skipif(not os.path.exists(REF_LIB),
                              reason="oracle/_ref/libref_math.so not built (no reference tree)")

FP = ctypes.POINTER(ctypes.c_float)
REC = np.dtype([("pts", "<f4", 8), ("jt", "<f4", 2), ("parametric", "<u4"), ("polar", "<u4"),
                ("join", "<u4"), ("cfi", "<u4"), ("total", "<u4"), ("fwd", "<u4"), ("mir", "<u4")])
NONE = 0xFFFFFFFF
JOIN_TYPE_MASK = 7 << 26
ROUND_JOIN_FLAG = 2 << 26
BEVEL_JOIN_FLAG = 3 << 26
MITER_REVERT_JOIN_FLAG = 4 << 26
CAP_FLAG = 1 << 25
MIRRORED_FLAG = 1 << 23
CONTOUR_ID_MASK = 0xFFFF
W = 2048


def ref_lib():
    L = ctypes.CDLL(REF_LIB)
    f = ctypes.c_float
    L.ref_wang_cubic_pow4.restype = f
    L.ref_wang_cubic_pow4.argtypes = [FP, f, FP]
    L.ref_fast_acos.restype = f
    L.ref_fast_acos.argtypes = [f]
    L.ref_polar_segments_per_radian.restype = f
    L.ref_polar_segments_per_radian.argtypes = [f]
    L.ref_find_max_scale.restype = f
    L.ref_find_max_scale.argtypes = [FP]
    L.ref_eval_cubic.argtypes = [FP, f, FP]
    return L


def fptr(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(FP)


def records(f):
    n = ob.lib().oracle_get(f.h, 6, None, 0)
    buf = np.empty(n, np.uint8)
    ob.lib().oracle_get(f.h, 6, buf.ctypes.data, n)
    return buf.view(REC)


def random_scene(seed, n_draws=36, size=512):
    """Fills (three fill rules) and strokes (every join and cap) of random
    cubic/line contours, open and closed, under random affine matrices that
    keep every draw on the canvas; no clips, so path ID k is draw k-1."""
    rng = np.random.default_rng(seed)
    b = StreamBuilder(size, size)
    meta = []
    for _ in range(n_draws):
        cmds = []
        for c in range(int(rng.integers(1, 4))):
            p = rng.uniform(-60, 60, 2)
            cmds.append((0, *p))
            for _ in range(int(rng.integers(1, 7))):
                if rng.random() < 0.3:
                    p = rng.uniform(-60, 60, 2)
                    cmds.append((1, *p))
                else:
                    q = rng.uniform(-60, 60, 6)
                    cmds.append((4, *q))
            if rng.random() < 0.6:
                cmds.append((5,))
        stroke = rng.random() < 0.5
        rule = int(rng.integers(0, 3))
        path = b.path(cmds, fill_rule=rule)
        th = rng.uniform(0, 2 * math.pi)
        sx, sy = math.exp(rng.uniform(math.log(0.3), math.log(3))), math.exp(rng.uniform(math.log(0.3), math.log(3)))
        shear = rng.uniform(-0.5, 0.5)
        c, s = math.cos(th), math.sin(th)
        # M = R * [[sx, shear], [0, sy]] (column-major xx, xy, yx, yy)
        xx, xy = c * sx, s * sx
        yx, yy = c * shear - s * sy, s * shear + c * sy
        if rng.random() < 0.2:  # a mirroring matrix (negative determinant)
            xx, xy = -xx, -xy
        tx, ty = rng.uniform(size * 0.3, size * 0.7, 2)
        m = (xx, xy, yx, yy, tx, ty)
        if stroke:
            width = math.exp(rng.uniform(math.log(0.5), math.log(30)))
            join = int(rng.integers(0, 3))
            cap = int(rng.integers(0, 3))
            b.draw(path, matrix=m, stroke=width, join=join, cap=cap)
            meta.append((m, width / 2, join, cap))
        else:
            b.draw(path, matrix=m)
            meta.append((m, 0.0, None, None))
    return b, meta


SEEDS = [11, 12, 13, 14]


@pytest.fixture(scope="module", params=SEEDS)
def scene(request):
    b, meta = random_scene(request.param)
    s = b.build()
    f = ob.OracleFlush(s)
    yield b, meta, f
    f.close()


def record_draws(f, recs):
    """Index of the draw owning each record (by tessellation location)."""
    d = f.draws()
    starts = d[:, 1].astype(np.int64)
    loc = np.where(recs["fwd"] != NONE, recs["fwd"].astype(np.int64), recs["mir"].astype(np.int64) - 1)
    return np.searchsorted(starts, loc, side="right") - 1


def tangents(p):
    """find_cubic_tan0/tan1 (bezier_utils.hpp:155-164)."""
    p = np.asarray(p, np.float32).reshape(4, 2)
    ne = lambda a, b: bool(np.any(a != b))
    t0 = (p[1] if ne(p[0], p[1]) else p[2] if ne(p[1], p[2]) else p[3]) - p[0]
    t1 = p[3] - (p[2] if ne(p[3], p[2]) else p[1] if ne(p[2], p[1]) else p[0])
    return t0.astype(np.float32), t1.astype(np.float32)


def polar_segments(L, t0, t1, spr):
    """draw.cpp:1219-1238 with the reference's fast_acos: float32 cosine of
    the rotation, ceil(theta * spr), clamped to [1, 1023]."""
    t0 = t0.astype(np.float32)
    t1 = t1.astype(np.float32)
    numer = np.float32(t0[0] * t1[0] + t0[1] * t1[1])
    den = np.float32(np.float32(t0[0] * t0[0] + t0[1] * t0[1]) * np.float32(t1[0] * t1[0] + t1[1] * t1[1]))
    cos = np.float32(numer / np.sqrt(den))
    cos = np.float32(min(max(cos, np.float32(-1)), np.float32(1))) if not np.isnan(cos) else np.float32(-1)
    theta = np.float32(L.ref_fast_acos(float(cos)))
    n = np.ceil(np.float32(theta * np.float32(spr)))
    return int(min(max(n, 1), 1023))


def contour_segments(b, path_index, stroke, L):
    """The pushCubic records a path produces, per contour, from its verbs:
    a list of (kind, start tangent, end tangent, pieces) per segment, with
    pieces = 1 + the reference's convex-180 chops for stroked cubics (twice
    the chops around cusps, draw.cpp:1052-1090), an implicit closing line
    when the contour ends away from its start (fills always close), and
    whether the contour is closed."""
    fv, vc, fp_, pc, rule = b.paths[path_index]
    verbs = b.verbs[fv:fv + vc]
    pts = np.asarray(b.points, np.float32).reshape(-1, 2)[fp_:fp_ + pc]
    contours, cur, k, start, last, closed = [], None, 0, None, None, False

    def finish():
        if cur is None:
            return
        is_closed = closed or not stroke
        segs = list(cur)
        if is_closed and np.any(last != start):
            segs.append(("line", (start - last).astype(np.float32), (start - last).astype(np.float32), 1))
        contours.append((segs, is_closed))

    for v in verbs:
        if v == 0:
            finish()
            start = last = pts[k]
            k += 1
            cur, closed = [], False
        elif v == 1:
            t = (pts[k] - last).astype(np.float32)
            cur.append(("line", t, t, 1))
            last = pts[k]
            k += 1
        elif v == 4:
            c = np.stack([last, pts[k], pts[k + 1], pts[k + 2]]).astype(np.float32)
            t0, t1 = tangents(c.ravel())
            pieces = 1
            if stroke:
                T = np.zeros(2, np.float32)
                cusp = ctypes.c_int()
                ca, cp = fptr(c.ravel())
                n = L.ref_convex_180_chops(cp, T.ctypes.data_as(FP), ctypes.byref(cusp))
                pieces = 1 + (2 * n if cusp.value else n)
            cur.append(("cubic", t0, t1, pieces))
            last = pts[k + 2]
            k += 3
        elif v == 5:
            closed = True
    finish()
    return contours


@need_ref
def test_segment_counts_match_reference_maths(scene):
    """Every pushCubic record, walked from the path's verbs: parametric =
    clamp(ceil(4th root of the reference's cubic_pow4 under the draw
    matrix), 1, 1023); for strokes, polar = polar_segments(the piece's
    tangents, spr) with spr = the reference's
    calc_polar_segments_per_radian(r * findMaxScale(M)); a cubic's chopped
    pieces before the last join with 1 segment; a segment's real join has 5
    segments for miter/bevel and, for round, the polar count of the rotation
    from the segment's unchopped end tangent to the next segment's start
    tangent (which the record carries); an open contour starts with a
    zero-segment cap carrier and ends with a cap of min(ceil(spr*pi)+2,
    1023) segments when round, else 5 (draw.cpp:1195-1342, 2060-2300)."""
    L = ref_lib()
    L.ref_convex_180_chops.argtypes = [FP, FP, ctypes.POINTER(ctypes.c_int)]
    L.ref_convex_180_chops.restype = ctypes.c_int
    b, meta, f = scene
    recs = records(f)
    owner = record_draws(f, recs)
    checked = {"parametric": 0, "polar": 0, "round": 0, "miter_bevel": 0, "cap": 0, "chop": 0}
    for di, (m, radius, join, cap) in enumerate(meta):
        rk = recs[owner == di]
        stroke = radius > 0
        ma, mp = fptr(m)
        spr = None
        if stroke:
            scale = L.ref_find_max_scale(mp)
            spr = np.float32(L.ref_polar_segments_per_radian(np.float32(np.float32(radius) * np.float32(scale))))
        i = 0
        for segs, closed in contour_segments(b, b.draws[di]["path"], stroke, L):
            if stroke and not closed:
                carrier = rk[i]
                assert carrier["parametric"] == 0 and carrier["polar"] == 0 and (carrier["cfi"] & CAP_FLAG)
                i += 1
            for si, (kind, st0, st1, pieces) in enumerate(segs):
                for pi in range(pieces):
                    r = rk[i]
                    i += 1
                    pa, pp = fptr(r["pts"])
                    n4 = np.float32(L.ref_wang_cubic_pow4(pp, 4.0, mp))
                    want = int(min(max(np.ceil(np.sqrt(np.sqrt(n4))), 1), 1023))
                    assert r["parametric"] == want, (r, n4)
                    checked["parametric"] += 1
                    if not stroke:
                        assert r["polar"] == 1 and r["join"] == 1, r
                        continue
                    t0, t1 = tangents(r["pts"])
                    assert r["polar"] == polar_segments(L, t0, t1, spr), (r, spr)
                    checked["polar"] += 1
                    if pi < pieces - 1:
                        assert r["join"] == 1, r
                        checked["chop"] += 1
                        continue
                    if not closed and si == len(segs) - 1:
                        assert r["cfi"] & CAP_FLAG
                        cw = int(min(np.ceil(np.float32(spr * np.float32(math.pi))) + 2, 1023))
                        assert r["join"] == (cw if cap == ROUND_CAP else 5), (r, cap, spr)
                        checked["cap"] += 1
                        continue
                    nxt = segs[(si + 1) % len(segs)]
                    if join == ROUND:
                        assert (r["cfi"] & JOIN_TYPE_MASK) == ROUND_JOIN_FLAG
                        assert np.array_equal(r["jt"], nxt[1]), (r, nxt)
                        assert r["join"] == polar_segments(L, st1, nxt[1], spr), (r, spr)
                        checked["round"] += 1
                    else:
                        assert (r["cfi"] & JOIN_TYPE_MASK) in (MITER_REVERT_JOIN_FLAG, BEVEL_JOIN_FLAG)
                        assert r["join"] == 5, r
                        checked["miter_bevel"] += 1
        assert i == len(rk), (di, i, len(rk))
    assert all(v > 0 for k, v in checked.items() if k != "chop"), checked


def single_stroke(cmds, width, join, cap, matrix=(1, 0, 0, 1, 100, 100)):
    b = StreamBuilder(400, 400)
    b.draw(b.path(cmds), matrix=matrix, stroke=width, join=join, cap=cap)
    s = b.build()
    f = ob.OracleFlush(s)
    try:
        return records(f), f.draws()
    finally:
        f.close()


def spr_closed_form(r):
    """Polar segments per radian for radius r (1/8 px precision):
    .5 / acos(1 - 1/(8 r)) (bezier_utils.hpp:108-113), in float64."""
    return 0.5 / math.acos(1 - 1 / (8 * r))


def test_hand_computed_joins_and_caps():
    """Round joins of 90 and 180 degrees and round caps at fixed radii, by
    hand: a polyline turning 90 degrees, stroked 40 px wide (r = 20):
    spr = .5/acos(1 - 1/160) = 4.4699, so the 90-degree round join has
    ceil(pi/2 * 4.4699) = ceil(7.0213) = 8 segments and a round cap
    ceil(pi * 4.4699) + 2 = 15 + 2 = 17; miter and bevel joins and square or
    butt caps have 5."""
    r = 20.0
    spr = spr_closed_form(r)
    assert abs(spr - 4.4699) < 1e-3
    turn90 = [(0, 0, 0), (1, 100, 0), (1, 100, 100)]
    # records: the start cap carrier (no segments, the cap's join), the first
    # line with the round join, the last line carrying the end cap
    recs, _ = single_stroke(turn90, 2 * r, ROUND, ROUND_CAP)
    assert list(recs["parametric"]) == [0, 1, 1] and list(recs["polar"]) == [0, 1, 1]
    assert list((recs["cfi"] & CAP_FLAG) != 0) == [True, False, True]
    assert recs["join"][1] == math.ceil(math.pi / 2 * spr) == 8
    assert recs["join"][0] == recs["join"][2] == math.ceil(math.pi * spr) + 2 == 17
    # a 180-degree turn (a line doubling back): ceil(pi * spr) = 15
    back = [(0, 0, 0), (1, 100, 0), (1, 0, 0)]
    recs, _ = single_stroke(back, 2 * r, ROUND, BUTT)
    assert recs["join"][1] == math.ceil(math.pi * spr) == 15
    assert recs["join"][0] == recs["join"][2] == 5
    for join in (MITER, BEVEL):
        for cap in (BUTT, SQUARE):
            recs, _ = single_stroke(turn90, 2 * r, join, cap)
            assert list(recs["join"]) == [5, 5, 5]
    # a closed square: four lines, four round joins of 90 degrees, no caps
    square = [(0, 0, 0), (1, 100, 0), (1, 100, 100), (1, 0, 100), (5,)]
    recs, _ = single_stroke(square, 2 * r, ROUND, ROUND_CAP)
    assert list(recs["join"]) == [8, 8, 8, 8] and not np.any(recs["cfi"] & CAP_FLAG)
    # padding: the closed square's 4 x (1 + 1 + 8 - 1) = 36 vertices pad to 40
    assert int(recs["total"].sum()) == 40 and recs["total"][0] == 9 + 4
    # a quarter circle (cubic approximation, radius 100) stroked at r = 20
    # turns pi/2: ceil(pi/2 * spr) polar segments
    k = 100 * 0.5522847498
    arc = [(0, 100, 0), (4, 100, k, k, 100, 0, 100)]
    recs, _ = single_stroke(arc, 2 * r, BEVEL, BUTT)
    real = recs[recs["parametric"] > 0]
    assert real["polar"][0] == math.ceil(math.pi / 2 * spr) == 8
    # Wang's formula for that arc at precision 4 (a closed form):
    # n = sqrt(.75 * 4 * max|p0 - 2p1 + p2|, |p1 - 2p2 + p3|)
    p = np.array([[100, 0], [100, k], [k, 100], [0, 100]], np.float64)
    d = max(np.linalg.norm(p[0] - 2 * p[1] + p[2]), np.linalg.norm(p[1] - 2 * p[2] + p[3]))
    assert real["parametric"][0] == math.ceil(math.sqrt(0.75 * 4 * d))


def test_layout_padding_and_locations(scene):
    """Contours are padded to multiples of 8 texels (the padding sits on the
    contour's first record); records are contiguous from location 8; draws
    are contiguous; double-sided fills mirror their forward run from the end
    of the draw's range; the flush ends with one padding texel."""
    b, meta, f = scene
    recs = records(f)
    d = f.draws()
    core = recs["parametric"].astype(np.int64) + recs["polar"] + recs["join"] - 1
    pad = recs["total"].astype(np.int64) - core
    assert np.all((pad >= 0) & (pad < 8))
    cid = recs["cfi"] & CONTOUR_ID_MASK
    first_of_contour = np.r_[True, cid[1:] != cid[:-1]]
    assert np.all(pad[~first_of_contour] == 0)
    for c in np.unique(cid):
        assert recs["total"][cid == c].sum() % 8 == 0
    assert d[0, 1] == 8
    assert np.all(d[1:, 1] == d[:-1, 1] + d[:-1, 2])
    assert f.tess_vertex_count == d[-1, 1] + d[-1, 2] + 1
    owner = record_draws(f, recs)
    for k, (path_id, loc, count, dirs, is_stroke) in enumerate(d):
        rk = recs[owner == k]
        fwd_total = int(rk["total"].sum())
        double = dirs in (2, 3)
        assert count == (2 * fwd_total if double else fwd_total)
        fbase = loc + (fwd_total if dirs == 3 else 0)
        if dirs != 1:
            starts = fbase + np.r_[0, np.cumsum(rk["total"])[:-1]]
            assert np.array_equal(rk["fwd"].astype(np.int64), starts)
        if double:
            mend = loc + count if dirs == 2 else loc + fwd_total
            ends = mend - np.r_[0, np.cumsum(rk["total"])[:-1]]
            assert np.array_equal(rk["mir"].astype(np.int64), ends)


def test_spans_cover_their_records_exactly(scene):
    """Each record's spans cover exactly its forward texels
    [fwd, fwd + total) and, when mirrored, [mir - total, mir), once each, on
    the 2048-wide grid (a span's texels are row*2048 + [x0, x1) clipped to
    the row, mirrored spans run right to left); every span carries its
    record's points, join tangent, packed segment counts and flags; two
    padding spans cover texels [0, 8) and the final texel."""
    b, meta, f = scene
    recs = records(f)
    spans = f.get(0)
    x0 = (spans["x0x1"].astype(np.int32) << 16) >> 16
    x1 = spans["x0x1"].astype(np.int32) >> 16
    rx0 = (spans["reflection_x0x1"].astype(np.int32) << 16) >> 16
    rx1 = spans["reflection_x0x1"].astype(np.int32) >> 16

    def texels(y, a, c):
        if not (y >= 0):
            return []
        lo, hi = max(min(a, c), 0), min(max(a, c), W)
        return list(int(y) * W + np.arange(lo, hi))

    # the padding spans come first (render_context.cpp:1550-1568): 8 texels
    # at location 0, then the one texel after the last draw
    assert spans["segment_counts"][0] == 1 << 20 and texels(spans["y"][0], x0[0], x1[0]) == list(range(8))
    assert spans["segment_counts"][1] == 1 << 20
    assert texels(spans["y"][1], x0[1], x1[1]) == [f.tess_vertex_count - 1]
    i = 2
    for r in recs:
        counts = (int(r["join"]) << 20) | (int(r["polar"]) << 10) | int(r["parametric"])
        fwd_cov, mir_cov = [], []
        while len(fwd_cov) < (r["total"] if r["fwd"] != NONE else 0) or \
                len(mir_cov) < (r["total"] if r["mir"] != NONE else 0):
            s = spans[i]
            assert np.array_equal(s["pts"], r["pts"]) and np.array_equal(s["join_tangent"], r["jt"])
            assert s["segment_counts"] == counts and s["contour_id_with_flags"] == r["cfi"]
            if r["fwd"] != NONE:
                assert x1[i] >= x0[i]
                fwd_cov += texels(s["y"], x0[i], x1[i])
                if r["mir"] != NONE:
                    assert rx1[i] <= rx0[i]
                    mir_cov += texels(s["reflection_y"], rx0[i], rx1[i])
            else:
                assert x1[i] <= x0[i]
                mir_cov += texels(s["y"], x0[i], x1[i])
            i += 1
        if r["fwd"] != NONE:
            assert sorted(fwd_cov) == list(range(r["fwd"], r["fwd"] + r["total"]))
        if r["mir"] != NONE:
            assert sorted(mir_cov) == list(range(r["mir"] - r["total"], r["mir"]))
    assert i == len(spans)


@need_ref
def test_tessellated_vertices_closed_forms(scene):
    """The oracle's tessellation texels: a fill record's vertices after its
    padding are the cubic at T = i/n (n = parametric segments) within 1e-4
    of the reference's eval_cubic_at (relative to the control-point extent);
    a stroke record's curve vertices lie on its cubic with theta equal to the
    tangent angle there, consecutive ones turning by at most theta/polar +
    1e-3; each mirrored run holds its forward run's positions reversed."""
    L = ref_lib()
    b, meta, f = scene
    recs = records(f)
    tex = f.tessellate()
    owner = record_draws(f, recs)
    out = np.zeros(2, np.float32)
    optr = out.ctypes.data_as(FP)
    fills = strokes = 0
    for r, di in zip(recs, owner):
        if r["fwd"] == NONE or r["parametric"] == 0:
            continue
        radius = meta[di][1]
        pa, pp = fptr(r["pts"])
        ext = max(float(np.abs(r["pts"]).max()), 1.0)
        pad = int(r["total"]) - int(r["parametric"] + r["polar"] + r["join"] - 1)
        base = int(r["fwd"]) + pad
        if radius == 0:
            n = int(r["parametric"])
            for i in range(n):
                L.ref_eval_cubic(pp, np.float32(i / n), optr)
                t = tex[base + i]
                assert abs(t["x"] - out[0]) <= 1e-4 * ext and abs(t["y"] - out[1]) <= 1e-4 * ext, (i, n)
            fills += 1
            if r["mir"] != NONE:
                run = tex[int(r["fwd"]):int(r["fwd"]) + int(r["total"])]
                mrun = tex[int(r["mir"]) - int(r["total"]):int(r["mir"])][::-1]
                assert np.array_equal(run["x"], mrun["x"]) and np.array_equal(run["y"], mrun["y"])
                assert np.all((mrun["w"] & MIRRORED_FLAG) != 0) and np.all((run["w"] & MIRRORED_FLAG) == 0)
        else:
            # curve vertices: parametric + polar - 1 merged vertices
            nv = int(r["parametric"] + r["polar"] - 1)
            p = r["pts"].astype(np.float64).reshape(4, 2)
            ts = np.linspace(0, 1, 4001)
            bern = np.stack([(1 - ts) ** 3, 3 * ts * (1 - ts) ** 2, 3 * ts ** 2 * (1 - ts), ts ** 3], 1)
            curve = bern @ p
            t0, t1 = tangents(r["pts"])
            c = float(np.dot(t0, t1)) / max(float(np.linalg.norm(t0) * np.linalg.norm(t1)), 1e-30)
            theta_total = math.acos(max(-1.0, min(1.0, c)))
            prev = None
            for i in range(nv):
                t = tex[base + i]
                v = np.array([t["x"], t["y"]], np.float64)
                dist = np.hypot(curve[:, 0] - v[0], curve[:, 1] - v[1])
                k = int(np.argmin(dist))
                # refine the foot point: Newton on (B(T) - v) . B'(T) = 0
                T = ts[k]
                for _ in range(8):
                    bt = np.array([(1 - T) ** 3, 3 * T * (1 - T) ** 2, 3 * T ** 2 * (1 - T), T ** 3]) @ p
                    d1 = np.array([-3 * (1 - T) ** 2, 3 * (1 - T) ** 2 - 6 * T * (1 - T),
                                   6 * T * (1 - T) - 3 * T ** 2, 3 * T ** 2]) @ p
                    d2 = np.array([6 * (1 - T), -12 * (1 - T) + 6 * T, 6 * (1 - T) - 12 * T, 6 * T]) @ p
                    g = np.dot(bt - v, d1)
                    h = np.dot(d1, d1) + np.dot(bt - v, d2)
                    if h <= 0:
                        break
                    T = min(max(T - g / h, 0.0), 1.0)
                bt = np.array([(1 - T) ** 3, 3 * T * (1 - T) ** 2, 3 * T ** 2 * (1 - T), T ** 3]) @ p
                assert np.hypot(*(bt - v)) <= 2e-3 * ext + 2e-3, (i, dist[k])
                theta = float(np.array([t["z"]], np.uint32).view(np.float32)[0])
                dv = np.array([-3 * (1 - T) ** 2, 3 * (1 - T) ** 2 - 6 * T * (1 - T),
                               6 * T * (1 - T) - 3 * T ** 2, 3 * T ** 2]) @ p
                if np.hypot(*dv) > 1e-3 * ext:
                    ang = math.atan2(dv[1], dv[0]) % (2 * math.pi)
                    dd = abs(ang - theta) % (2 * math.pi)
                    assert min(dd, 2 * math.pi - dd) <= 2e-2, (i, ang, theta)
                if prev is not None and r["polar"] > 0:
                    dd = abs(theta - prev) % (2 * math.pi)
                    assert min(dd, 2 * math.pi - dd) <= theta_total / int(r["polar"]) + 1e-3
                prev = theta
            strokes += 1
    assert fills > 0 or strokes > 0


# Limits of one logical flush (render_context.cpp:43-53,484-492,517-541,617-654).
MAX_PATH_ID = (0x1F << 10) - 1 - 1023 - 1          # MaxPathID(1) - 1
MAX_CONTOUR_ID = 65535
MAX_TESS_BEFORE_PADDING = 2048 * 2048 - (8 + 16 + 1)
MAX_GRADIENT_ROWS = 2048


def flush_infos(f):
    out = []
    for k in range(f.flush_count):
        f.select(k)
        out.append(dict(paths=f.path_count, contours=f.contour_count, tess=f.tess_vertex_count,
                        simple=f.grad_simple_rows, complex=f.grad_complex_rows, draws=f.draws().copy()))
    f.select(0)
    return out


def test_logical_flush_path_limit_and_preserve():
    """30 720 one-path draws: the first logical flush takes exactly
    MaxPathID(1) - 1 = 30 719 paths, the last draw starts a second flush,
    and that flush loads the first one's pixels (every pixel keeps its own
    draw's color, the last one included)."""
    n = MAX_PATH_ID + 1
    b, colors = pixel_rects_stream(n)
    f = ob.OracleFlush(b.build())
    try:
        info = flush_infos(f)
        assert [i["paths"] for i in info] == [MAX_PATH_ID, 1]
        img, _ = f.render()
    finally:
        f.close()
    px = img.reshape(-1, 4)[:n]
    argb = (0xff000000 | (px[:, 0].astype(np.uint32) << 16) | (px[:, 1].astype(np.uint32) << 8)
            | px[:, 2].astype(np.uint32))
    np.testing.assert_array_equal(argb, colors)


def test_logical_flush_contour_limit():
    """66 draws of a 1000-contour path: 65 fit under the 65 535 contour IDs,
    the 66th starts a new flush."""
    b = contour_limit_stream()
    f = ob.OracleFlush(b.build())
    try:
        info = flush_infos(f)
    finally:
        f.close()
    assert [len(i["draws"]) for i in info] == [65, 1]
    assert [i["contours"] for i in info] == [65000, 1000]


def test_logical_flush_tessellation_limit():
    """Identical wide round-joined strokes: a flush holds as many as fit in
    2048^2 - 25 vertices, each draw's count taken from its own draw record."""
    b = tess_limit_stream()
    f = ob.OracleFlush(b.build())
    try:
        info = flush_infos(f)
    finally:
        f.close()
    v = int(info[0]["draws"][0, 2])
    assert v > 20000, v
    assert all(int(i["draws"][k, 2]) == v for i in info for k in range(len(i["draws"])))
    per_flush = MAX_TESS_BEFORE_PADDING // v
    assert [len(i["draws"]) for i in info][:-1] == [per_flush] * (len(info) - 1)
    assert len(info) >= 2
    assert sum(len(i["draws"]) for i in info) == 120
    for i in info:
        # 8 leading padding vertices, the draws, 1 trailing padding vertex.
        assert i["tess"] == 8 + v * len(i["draws"]) + 1


def test_logical_flush_gradient_row_limit():
    """2049 distinct three-stop gradients need 2049 complex ramp rows; the
    gradient texture holds 2048, so the last draw starts a new flush."""
    b = gradient_limit_stream()
    f = ob.OracleFlush(b.build())
    try:
        info = flush_infos(f)
    finally:
        f.close()
    assert [i["complex"] for i in info] == [MAX_GRADIENT_ROWS, 1]
    assert [len(i["draws"]) for i in info] == [MAX_GRADIENT_ROWS, 1]
