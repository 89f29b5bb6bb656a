"""Feathering in the oracle, pinned against the reference's own data and tests
and against what the mathematics fixes.

* The two fp16 function tables regenerated from their definitions
  (gpu.cpp:1991-2063, 2126-2181) equal the reference's shipped tables bit for
  bit (tests/golden/feather_tables.npz, parsed from gpu.cpp by
  tests/golden/make_feather_tables.py), and satisfy the reference's own table
  tests (gpu_namespace_test.cpp:34-160).
* find_cubic_max_height and measure_cubic_local_curvature satisfy the
  properties bezier_utils_test.cpp:690-745 checks, and the curvature of a
  circular arc matches 2*asin(spread / 2R).
* A feathered fill is the shape convolved with a Gaussian of standard
  deviation feather/2 (find_feather_radius, draw.cpp:313-318): a straight
  edge must follow the normal CDF, a convex right-angle corner the product of
  two CDFs, a concave one its complement.  These exercise the whole chain:
  feather-join counting, tessellation, the center-AA patch, the vertex
  shader's corner packing and the fragment integral.
* The reference's feathercusp case (feather_test.cpp:13-47) renders.
"""
import ctypes
import math
import os

import numpy as np
import pytest
from scipy.special import ndtr

import oracle_binding as ob
from streams import CLOCKWISE, NONZERO, StreamBuilder, rect

HERE = os.path.dirname(os.path.abspath(__file__))
FP = ctypes.POINTER(ctypes.c_float)
U16P = ctypes.POINTER(ctypes.c_uint16)


def _f16(u):
    return np.asarray(u, np.uint16).view(np.float16).astype(np.float32)


def _tables():
    g = np.zeros(512, np.uint16)
    i = np.zeros(512, np.uint16)
    ob.lib().oracle_feather_tables(g.ctypes.data_as(U16P), i.ctypes.data_as(U16P))
    return g, i


def test_tables_equal_the_references_bit_for_bit():
    g, i = _tables()
    ref = np.load(os.path.join(HERE, "golden", "feather_tables.npz"))
    np.testing.assert_array_equal(g, ref["gaussian_integral"])
    np.testing.assert_array_equal(i, ref["inverse_gaussian_integral"])


def test_gaussian_table_properties():
    # gpu_namespace_test.cpp:34-66
    t = _f16(_tables()[0])
    assert t[0] == 0.0 and t[511] == 1.0
    assert t[255] <= .5 <= t[256]
    assert abs((t[255] + t[256]) / 2 - .5) <= 1e-3
    assert np.all(np.diff(t) >= 0)
    assert np.all(np.abs(t[:256] + t[::-1][:256] - 1) <= 1e-3)
    # and it is the normal CDF over +-3 standard deviations
    x = (np.arange(512) + .5) / 512
    assert np.abs(t - ndtr((x - .5) * 6)).max() < 5e-3


def test_inverse_table_properties_and_inversion():
    # gpu_namespace_test.cpp:79-160, with its margins
    L = ob.lib()
    g, i = _tables()
    it = _f16(i)
    assert it[0] == 0 and it[511] == 1
    assert abs((it[255] + it[256]) / 2 - .5) <= 1e-4
    assert np.all(np.diff(it) >= 0)
    for k in range(256 - 4):
        m = 5e-4 if k > 100 else (1e-3 if k > 4 else 1e-2)
        assert abs(it[k] + it[511 - k] - 1) <= m + 1e-6
    gp, ip = g.ctypes.data_as(U16P), i.ctypes.data_as(U16P)
    for x in np.arange(0, 1, 1.0 / (512 * 21), dtype=np.float32):
        y = L.oracle_table_lookup(gp, float(x))
        margin = (1 / 512 if .125 < x < .875 else 1 / 256 if .04 < x < .96 else
                  1 / 128 if .02 < x < .98 else 1 / 90)
        assert abs(L.oracle_table_lookup(ip, y) - x) <= margin + 1e-6, x
    assert L.oracle_table_lookup(ip, -1.0) == 0
    assert L.oracle_table_lookup(ip, 2.0) == 1
    assert L.oracle_table_lookup(ip, float("-inf")) == 0
    assert L.oracle_table_lookup(ip, float("inf")) == 1


def test_flat_edge_feather_is_the_table():
    # A flat edge (cotTheta = HORIZONTAL_COTANGENT_VALUE) feathers y0 alone,
    # signed by coverages.x.
    L = ob.lib()
    for y0 in [0.0, .1, .25, .5, .75, 1.0]:
        c = (ctypes.c_float * 4)(.3, -2.0, 1e6, y0)
        assert L.oracle_eval_feathered_fill(c) == pytest.approx(L.oracle_feather(y0), abs=1e-7)
        c = (ctypes.c_float * 4)(-.3, -2.0, 1e6, y0)
        assert L.oracle_eval_feathered_fill(c) == pytest.approx(-L.oracle_feather(y0), abs=1e-7)


def test_right_angle_corner_is_separable():
    # A corner of pi/2 packs cotTheta = y0 = 0, and its integral is the
    # product of two one-dimensional feathers: coverage at the corner's local
    # coordinate (x, y) in the convolution square is FEATHER(x) * FEATHER(y).
    L = ob.lib()
    for sx, sy in [(1.0, 0.0), (.6, .8), (0.0, 1.0)]:
        for outset in [0.0, .25, .5, 1.0]:
            c = (ctypes.c_float * 4)()
            L.oracle_pack_feathered_fill(math.pi / 2, sx, sy, outset, c)
            x = (1 - sx * outset) * .5
            y = (1 - sy * outset) * .5
            want = ndtr((x - .5) * 6) * ndtr((y - .5) * 6)
            assert L.oracle_eval_feathered_fill(c) == pytest.approx(want, abs=0.02)


def _cubic(p):
    return (ctypes.c_float * 8)(*[float(v) for xy in p for v in xy])


def _eval(p, t):
    p = np.asarray(p, np.float64)
    mt = 1 - t
    return (mt ** 3 * p[0] + 3 * mt * mt * t * p[1] + 3 * mt * t * t * p[2] + t ** 3 * p[3])


def _cubics():
    out = []
    for i in range(256):
        out.append([(((i >> (2 * k)) & 1) * 100.0, ((i >> (2 * k + 1)) & 1) * 100.0) for k in range(4)])
    rng = np.random.default_rng(0)
    for _ in range(100):
        out.append([tuple(v) for v in rng.uniform(-100, 100, (4, 2))])
    return out


def test_cubic_max_height():
    # bezier_utils_test.cpp:690-745
    L = ob.lib()
    for p in _cubics():
        t = ctypes.c_float()
        h = L.oracle_cubic_max_height(_cubic(p), ctypes.byref(t))
        t = t.value
        assert h >= 0 and 0 <= t <= 1
        a, b = np.asarray(p[0], np.float64), np.asarray(p[3], np.float64)
        if np.all(a == b):
            assert h == 0
            continue
        n = (b - a) / np.linalg.norm(b - a)
        n = np.array([-n[1], n[0]])
        k = -n @ a

        def height(tt):
            return abs(n @ _eval(p, tt) + k)
        assert height(t) == pytest.approx(h, abs=max(1e-4, 1e-5 * h))
        eps = max(h, 1) * 1e-3
        ts = np.arange(0, 1.0001, .005137)
        # Newton's three steps find the one maximum of a convex curve; the
        # height function of a general cubic can have two, so only the
        # convex ones are held to the global bound.
        convex = _is_convex(p)
        if convex:
            assert all(h + eps > height(tt) for tt in ts)


def _is_convex(p):
    p = np.asarray(p, np.float64)
    def cross(a, b):
        return a[0] * b[1] - a[1] * b[0]
    d = [cross(p[(i + 1) % 4] - p[i], p[(i + 2) % 4] - p[(i + 1) % 4]) for i in range(4)]
    d = [v for v in d if abs(v) > 1e-9]
    return len(d) > 0 and (all(v > 0 for v in d) or all(v < 0 for v in d))


def test_local_curvature_of_a_circular_arc():
    # A quarter circle of radius R as a cubic: across a spread s along the
    # tangent at the arc's middle, the tangent turns by 2 asin(s / 2R).
    L = ob.lib()
    R = 100.0
    k = 0.5522847498 * R
    p = [(R, 0.0), (R, k), (k, R), (0.0, R)]
    for s in [1.0, 10.0, 50.0, 100.0]:
        got = L.oracle_cubic_local_curvature(_cubic(p), .5, s)
        assert got == pytest.approx(2 * math.asin(s / (2 * R)), abs=2e-3), s
    # the spread is clipped inside T = 0..1: the whole arc turns pi/2
    assert L.oracle_cubic_local_curvature(_cubic(p), .5, 1e6) == pytest.approx(math.pi / 2, abs=3e-3)
    # a straight cubic has no curvature
    assert L.oracle_cubic_local_curvature(_cubic([(0, 0), (10, 0), (20, 0), (30, 0)]), .5, 5.0) == 0


def _soften(cmds_verbs, pts, feather, scale):
    L = ob.lib()
    v = np.asarray(cmds_verbs, np.uint8)
    p = np.asarray(pts, np.float32).reshape(-1)
    cap = 4096
    ov = np.zeros(cap, np.uint8)
    op = np.zeros(cap * 2, np.float32)
    nv = ctypes.c_uint32()
    n = L.oracle_soften_path(v.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(v), p.ctypes.data_as(FP),
                             len(p) // 2, feather, scale, ov.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), cap,
                             op.ctypes.data_as(FP), cap, ctypes.byref(nv))
    return ov[:nv.value].copy(), op[:2 * n].reshape(-1, 2).copy()


def test_softened_path_traces_the_same_curve():
    # makeSoftenedCopyForFeathering only chops: the pieces of each cubic join
    # end to end on the original curve, and each turns by at most the chop
    # angle (rive_render_path.cpp:323-449).
    R = 100.0
    k = 0.5522847498 * R
    verbs = [0, 4, 4, 5]
    pts = [(R, 0), (R, k), (k, R), (0, R), (-k, R), (-R, k), (-R, 0)]
    for feather, scale in [(2.0, 1.0), (10.0, 1.0), (40.0, 2.0)]:
        ov, op = _soften(verbs, pts, feather, scale)
        assert ov[0] == 0 and ov[-1] == 5
        assert np.allclose(op[0], pts[0]) and np.allclose(op[-1], pts[-1], atol=1e-4)
        # every output point lies on the circle (cubic approximation error
        # of the input is < 3e-4 R)
        r = np.hypot(op[:, 0], op[:, 1])
        ends = [0]
        i = 1
        for v in ov[1:]:
            if v == 4:
                i += 3
                ends.append(i - 1)
            elif v == 1:
                i += 1
                ends.append(i - 1)
        assert np.all(np.abs(r[ends] - R) < 0.05)
        assert len(ends) - 1 >= 2


def test_softened_cusp_is_crossed_with_a_line():
    # A cusp gets padded chops around it and a straight line through it.
    verbs = [0, 4]
    pts = [(0, 0), (100, 100), (0, 100), (100, 0)]
    ov, op = _soften(verbs, pts, 8.0, 1.0)
    assert 1 in ov[1:].tolist()


def _render(b):
    of = ob.OracleFlush(b.build())
    img, _ = of.render()
    of.close()
    return img[..., 0].astype(np.float64) / 255


@pytest.mark.parametrize("feather", [2.0, 8.0])
def test_feathered_edge_and_corners_are_gaussian(feather):
    sig = feather / 2
    b = StreamBuilder(256, 256, clear_color=0xff000000)
    b.draw(b.path(rect(64, 64, 600, 600), fill_rule=CLOCKWISE), color=0xffffffff, feather=feather)
    cov = _render(b)
    xs = np.arange(30, 110)
    np.testing.assert_allclose(cov[200, xs], ndtr((xs + .5 - 64) / sig), atol=4e-3)
    np.testing.assert_allclose(cov[xs, 200], ndtr((xs + .5 - 64) / sig), atol=4e-3)
    yy, xx = np.mgrid[52:80, 52:80]
    want = ndtr((xx + .5 - 64) / sig) * ndtr((yy + .5 - 64) / sig)
    assert np.abs(cov[52:80, 52:80] - want).max() < 8e-3


def test_feathered_concave_corner():
    sig = 4.0
    b = StreamBuilder(256, 256, clear_color=0xff000000)
    ell = [(0, 40, 40), (1, 200, 40), (1, 200, 128), (1, 128, 128), (1, 128, 200), (1, 40, 200), (5,)]
    b.draw(b.path(ell, fill_rule=CLOCKWISE), color=0xffffffff, feather=8.0)
    cov = _render(b)
    yy, xx = np.mgrid[116:142, 116:142]
    want = 1 - ndtr((xx + .5 - 128) / sig) * ndtr((yy + .5 - 128) / sig)
    assert np.abs(cov[116:142, 116:142] - want).max() < 8e-3


def test_feathered_circle_is_a_radial_gaussian():
    sig, R, c = 4.0, 90.0, 128.0
    k = 0.5522847498 * R
    circ = [(0, c + R, c), (4, c + R, c + k, c + k, c + R, c, c + R), (4, c - k, c + R, c - R, c + k, c - R, c),
            (4, c - R, c - k, c - k, c - R, c, c - R), (4, c + k, c - R, c + R, c - k, c + R, c), (5,)]
    b = StreamBuilder(256, 256, clear_color=0xff000000)
    b.draw(b.path(circ, fill_rule=CLOCKWISE), color=0xffffffff, feather=8.0)
    cov = _render(b)
    yy, xx = np.mgrid[0:256, 0:256]
    d = np.hypot(xx + .5 - c, yy + .5 - c) - R
    m = np.abs(d) < 20
    err = np.abs(cov - ndtr(-d / sig))[m]
    # the tessellation shader softens curves on purpose (tessellate.glsl:146-193)
    assert err.max() < .04 and err.mean() < .01


def test_feathered_fills_that_are_not_clockwise_are_skipped():
    # rive_renderer.cpp:164-170
    b = StreamBuilder(64, 64, clear_color=0xff000000)
    b.draw(b.path(rect(10, 10, 50, 50), fill_rule=NONZERO), color=0xffffffff, feather=4.0)
    assert _render(b).max() == 0


def test_feathercusp_renders():
    # feather_test.cpp:13-47: a zero-area double cusp with feather 32.
    b = StreamBuilder(1000, 1000, clear_color=0xff000000)
    p = b.path([(0, 95.164070, 799.999939),
                (4, -125.766922, 799.999939, 274.233032, 799.999939, -95.164070, 799.999939)],
               fill_rule=CLOCKWISE)
    b.draw(p, color=0xffffffff, feather=32.0)
    cov = _render(b)
    assert np.isfinite(cov).all()


def test_feathered_stroke_profile_is_a_convolved_band():
    # A straight stroke of half-width h feathered by f is the band |d| <= h
    # convolved with a Gaussian of standard deviation f/2 across it:
    # Phi((d + h)/sigma) - Phi((d - h)/sigma) (draw_path_common.glsl:242-258
    # on the stroke's linear coverages, aaRadius = the feather radius).
    for width, feather in [(6.0, 4.0), (20.0, 8.0), (2.0, 6.0)]:
        h, sig = width / 2, feather / 2
        b = StreamBuilder(256, 128, clear_color=0xff000000)
        b.draw(b.path([(0, -100.0, 64.0), (1, 400.0, 64.0)]), color=0xffffffff, stroke=width, feather=feather)
        cov = _render(b)
        ys = np.arange(30, 98)
        d = ys + .5 - 64
        want = ndtr((d + h) / sig) - ndtr((d - h) / sig)
        np.testing.assert_allclose(cov[ys, 128], want, atol=5e-3)
