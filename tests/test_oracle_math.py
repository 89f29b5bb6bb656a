"""Pins the oracle's curve maths against the reference's own unit-test vectors
and against closed forms.

Vectors and tolerances follow the reference's tests:
  wangs_formula_test.cpp:27-39 (kSerp/kLoop/kQuad), :150-225 (log2 boundaries)
  eval_cubic_test.cpp:35-45 and bezier_utils_test.cpp:560-585 (TEST_CUBICS)
  bezier_utils_test.cpp:34-129 (chop_cubic_at), :451-550 (convex-180 chops)
  simd_test.cpp:586-700 (fast_acos, SIMD_FAST_ACOS_MAX_ERROR)
Closed forms: Wang's formula n = sqrt(0.75 * prec * max|M(p0-2p1+p2)|,
|M(p1-2p2+p3)|), the Bernstein form of a cubic, Mat2D max scale = the largest
singular value, polar density .5/acos(1 - 1/(8 r)).
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_binding as ob

KEPS = 1.0 / (1 << 12)
FAST_ACOS_MAX_ERROR = 0.0167552  # simd.hpp:495
TESS_TOL = 1 / 128.0             # wangs_formula_test.cpp kTessellationTolerance

K_SERP = [(285.625, 499.687), (411.625, 808.188), (1064.62, 135.688), (1042.63, 585.187)]
K_LOOP = [(635.625, 614.687), (171.625, 236.188), (1064.62, 135.688), (516.625, 570.187)]
K_QUAD_AS_CUBIC = None  # kQuad is a quadratic; the renderer only sees cubics

TEST_CUBICS = [
    [(199, 1225), (197, 943), (349, 607), (549, 427)],
    [(549, 427), (349, 607), (197, 943), (199, 1225)],
    [(460, 1060), (403, -320), (60, 660), (1181, 634)],
    [(1181, 634), (60, 660), (403, -320), (460, 1060)],
    [(0, 0), (0, 0), (0, 0), (0, 0)],
    [(0, 0), (0, 0), (0, 0), (100, 100)],
    [(0, 0), (0, 0), (100, 100), (100, 100)],
    [(0, 0), (100, 100), (100, 100), (0, 0)],
    [(-100, -100), (100, 100), (100, -100), (-100, 100)],
    [(0, 0), (-100, -100), (200, 200), (100, 100)],
    [(0, 0), (50 * 2 / 3, 100 * 2 / 3), (100 - 50 * 2 / 3, 100 * 2 / 3), (100, 0)],
    [(0, 0), (50 * 2 / 3, 100 * 2 / 3), (100 - 50 * 2 / 3, 100 * 2 / 3), (100, 100)],
    [(100, 0), (0, 0), (0, 0), (0, 0)],
]

FP = ctypes.POINTER(ctypes.c_float)


def fbuf(values):
    a = np.ascontiguousarray(np.asarray(values, np.float32).ravel())
    return a, a.ctypes.data_as(FP)


def wang(pts, m=None, precision=4.0):
    a, pa = fbuf(pts)
    if m is None:
        return ob.lib().oracle_wang_cubic(pa, precision, None)
    b, pb = fbuf(m)
    return ob.lib().oracle_wang_cubic(pa, precision, pb)


def wang_reference(pts, m=(1, 0, 0, 1, 0, 0), precision=4.0):
    """wangs_formula_cubic_reference_impl with the points mapped by M's 2x2."""
    p = np.asarray(pts, np.float64)
    M = np.array([[m[0], m[2]], [m[1], m[3]]], np.float64)
    d0 = M @ (p[0] - 2 * p[1] + p[2])
    d1 = M @ (p[1] - 2 * p[2] + p[3])
    k = (3 * 2) / 8.0 * precision
    return math.sqrt(k * max(np.linalg.norm(d0), np.linalg.norm(d1)))


def random_matrices(rng):
    yield (1, 0, 0, 1, 0, 0)
    for i in range(-10, 31, 4):
        for j in range(-10, 31, 4):
            yield (math.ldexp(1 + rng.random(), i), 0, 0, math.ldexp(1 + rng.random(), j), 0, 0)
            yield (math.ldexp(1 + rng.random(), i), math.ldexp(1 + rng.random(), (i + j) // 2),
                   math.ldexp(1 + rng.random(), (i + j) // 2), math.ldexp(1 + rng.random(), j), 0, 0)


def test_wang_serp_loop_under_random_matrices():
    rng = np.random.default_rng(0)
    for m in random_matrices(rng):
        for curve in (K_SERP, K_LOOP):
            ref = max(1.0, wang_reference(curve, m))
            got = max(1.0, wang(curve, m))
            assert abs(got / ref - 1) <= TESS_TOL, (m, curve, got, ref)


def test_wang_log2_boundaries():
    """Cubics built so the length term sits just below / above 4^level / k:
    ceil(log2(n)) must land on level and level + 1."""
    k = (3 * 2) / (8 * (1 / 4.0))
    for level in range(0, 15):
        eps = math.ldexp(KEPS, level * 2)
        x = math.ldexp(1, level * 2) / k
        for term, want in ((x - eps, level), (x + eps, level + 1)):
            pts = [(term, 0.0), (0, 0), (0, 0), (term * .75, 0.0)]
            ref = wang_reference(pts)
            got = wang(pts)
            assert math.ceil(math.log2(ref)) == want
            assert abs(got / ref - 1) <= TESS_TOL
            assert math.ceil(math.log2(got)) == want


def test_wang_pow4_is_fourth_power():
    a, pa = fbuf(K_SERP)
    n = ob.lib().oracle_wang_cubic(pa, 4.0, None)
    n4 = ob.lib().oracle_wang_cubic_pow4(pa, 4.0, None)
    assert abs(n ** 4 / n4 - 1) < 1e-5


def eval_bernstein(p, t):
    p = np.asarray(p, np.float64)
    return ((1 - t) ** 3 * p[0] + 3 * (1 - t) ** 2 * t * p[1] + 3 * (1 - t) * t * t * p[2]
            + t ** 3 * p[3])


@pytest.mark.parametrize("idx", range(len(TEST_CUBICS)))
def test_eval_cubic_matches_bernstein(idx):
    cubic = TEST_CUBICS[idx]
    a, pa = fbuf(cubic)
    out = (ctypes.c_float * 2)()
    for t in np.arange(0, 1.0001, 1e-2, dtype=np.float32):
        ob.lib().oracle_eval_cubic(pa, float(t), out)
        ref = eval_bernstein(cubic, float(t))
        assert abs(out[0] - ref[0]) <= 1e-3 * max(1, abs(ref[0])) + 1e-3
        assert abs(out[1] - ref[1]) <= 1e-3 * max(1, abs(ref[1])) + 1e-3


CHOP_TS = [0, 3 / 83, 3 / 79, 3 / 73, 3 / 71, 3 / 67, 3 / 61, 3 / 59, 3 / 53, 3 / 47, 3 / 43,
           3 / 41, 3 / 37, 3 / 31, 3 / 29, 3 / 23, 3 / 19, 3 / 17, 3 / 13, 3 / 11, 3 / 7, 3 / 5, 1]


def test_chop_cubic_at():
    rng = np.random.default_rng(1)
    L = ob.lib()
    for _ in range(5):
        pts = rng.random((4, 2)).astype(np.float32)
        a, pa = fbuf(pts)
        out7 = (ctypes.c_float * 14)()
        out10 = (ctypes.c_float * 20)()
        for t in CHOP_TS:
            t = float(np.float32(t))
            L.oracle_chop_cubic_at(pa, t, out7)
            mid = np.array(out7[6:8])
            ref = eval_bernstein(pts, t)
            assert np.all(np.abs(mid - ref) <= KEPS)
            c = np.array(out7).reshape(7, 2)
            if t == 0:
                assert np.array_equal(c[3], pts[0])
            if t == 1:
                assert np.array_equal(c[3], pts[3])
            # chopping at (t, t) leaves an exactly degenerate middle cubic
            L.oracle_chop_cubic_at2(pa, t, t, out10)
            c = np.array(out10).reshape(10, 2)
            assert np.array_equal(c[3], c[4]) and np.array_equal(c[3], c[5]) and np.array_equal(c[3], c[6])
            if t == 0:
                assert all(np.array_equal(c[i], pts[0]) for i in range(4))
            if t == 1:
                assert all(np.array_equal(c[i], pts[3]) for i in range(6, 10))


def chops(pts):
    a, pa = fbuf(pts)
    T = (ctypes.c_float * 2)()
    cusp = ctypes.c_int()
    n = ob.lib().oracle_convex_180_chops(pa, T, ctypes.byref(cusp))
    return n, [T[0], T[1]][:n], bool(cusp.value)


def inflections(p):
    """Textbook inflection parameters: roots in (0,1) of cross(B', B'')."""
    p = np.asarray(p, np.float64)
    A = p[1] - p[0]
    B = p[2] - 2 * p[1] + p[0]
    C = p[3] + 3 * (p[1] - p[2]) - p[0]
    cross = lambda u, v: u[0] * v[1] - u[1] * v[0]
    a, b, c = cross(B, C), cross(A, C), cross(A, B)
    roots = []
    if abs(a) > 1e-12:
        disc = b * b - 4 * a * c
        if disc >= 0:
            s = math.sqrt(disc)
            roots = [(-b - s) / (2 * a), (-b + s) / (2 * a)]
    elif abs(b) > 1e-12:
        roots = [-c / b]
    return sorted(r for r in roots if 0 < r < 1)


def test_convex_180_chops_reference_cases():
    assert chops([(0, 0), (2, 2), (4, 2), (6, 0)])[0] == 0  # exact quadratic
    n, _, cusp = chops([(0, 0), (1, 1), (1, 0), (0, 1)])
    assert n == 1 and cusp
    # Inflections epsilon = 2^-11 apart in T are the cusp threshold.
    eps = 1.0 / (1 << 11)
    h = (1 - eps * eps) / (3 * eps * eps + 1)
    dy = (1 - h) / 2
    n, _, cusp = chops([(0, 0), (1, 1 - 4 * dy), (1, 0 + 4 * dy), (0, 1)])
    assert n == 2 and not cusp
    n, _, cusp = chops([(0, 0), (1, 1 - .9 * dy), (1, 0 + .9 * dy), (0, 1)])
    assert n == 1 and cusp
    # A quadratic raised to a cubic in fp32 (the "no fast floating point" case)
    p3, p4, p5 = np.float32([460, 460]), np.float32([667, 460]), np.float32([1060, 460])
    c0 = p3 + (p4 - p3) * np.float32(2 / 3)
    c1 = p5 + (p4 - p5) * np.float32(2 / 3)
    assert chops([p3, c0, c1, p5])[0] == 0
    # the convex-180 chop at T = 1 - epsilon
    hexpts = np.array([0x3ee0ac74, 0x3f1e061a, 0x3e0fc408, 0x3f457230, 0x3f42ac7c, 0x3f70d76c,
                       0x3f4e6520, 0x3f6acafa], np.uint32).view(np.float32).reshape(4, 2)
    n, T, _ = chops(hexpts)
    assert n >= 0 and all(0 <= t <= 1 for t in T)


def test_convex_180_chops_lines_are_never_cusps():
    p0, p3 = np.array([123, 200], np.float32), np.array([223, 432], np.float32)
    t0 = 1e-3
    while t0 < 1:
        t1 = t0 + .097
        while t1 < 1:
            line = [p0, p0 + (p3 - p0) * np.float32(t0), p0 + (p3 - p0) * np.float32(t1), p3]
            assert not chops(line)[2]
            t1 += .097
        t0 += .12


def test_convex_180_chops_match_inflections_on_corner_cubics():
    """All 256 cubics on the corners of the unit square: where the cubic has
    inflections (and they are not a cusp pair), the chops are the inflection
    points."""
    for i in range(256):
        p = [((i >> 0) & 1, (i >> 1) & 1), ((i >> 2) & 1, (i >> 3) & 1),
             ((i >> 4) & 1, (i >> 5) & 1), ((i >> 6) & 1, (i >> 7) & 1)]
        infl = inflections(p)
        n, T, cusp = chops(p)
        if infl and not cusp:
            assert n == len(infl), (p, infl, T)
            for a, b in zip(infl, T):
                assert abs(a - b) <= KEPS


def test_fast_acos_error_bound():
    L = ob.lib()
    for x in (-1.0, 0.0, 1.0):
        assert abs(L.oracle_fast_acos(x) - math.acos(x)) <= FAST_ACOS_MAX_ERROR
    for x in np.linspace(-1, 1, 20001, dtype=np.float32):
        assert abs(L.oracle_fast_acos(float(x)) - math.acos(float(x))) <= FAST_ACOS_MAX_ERROR


def test_find_max_scale_is_largest_singular_value():
    rng = np.random.default_rng(2)
    L = ob.lib()
    for _ in range(500):
        m = rng.normal(size=6).astype(np.float32) * np.float32(10 ** rng.uniform(-3, 3))
        a, pa = fbuf(m)
        got = L.oracle_find_max_scale(pa)
        M = np.array([[m[0], m[2]], [m[1], m[3]]], np.float64)
        ref = np.linalg.svd(M, compute_uv=False)[0]
        b = float(m[0]) * m[2] + float(m[3]) * m[1]
        if b * b > KEPS * KEPS:
            assert abs(got - ref) <= 1e-4 * ref + 1e-6
        else:
            # Nearly orthogonal columns: the reference takes the longer column
            # (mat2d_find_max_scale.cpp), which bounds sigma_max from below.
            col = max(np.linalg.norm(M[:, 0]), np.linalg.norm(M[:, 1]))
            assert col * (1 - 1e-5) <= got <= ref * (1 + 1e-5)
    a, pa = fbuf([3, 0, 0, -5, 7, 9])
    assert L.oracle_find_max_scale(pa) == 5.0


def test_polar_segments_per_radian_closed_form():
    L = ob.lib()
    for r in (0.5, 1, 2.5, 10, 100, 1e4):
        # the cosine is formed in fp32, then acos is exact
        cos = float(np.float32(1) - np.float32(1 / 8) / np.float32(r))
        ref = 0.5 / math.acos(max(cos, -1))
        assert abs(L.oracle_polar_segments_per_radian(r) / ref - 1) < 1e-6
    # tiny radii clamp the cosine at -1: exactly .5/pi segments per radian
    assert abs(L.oracle_polar_segments_per_radian(0.01) - 0.5 / math.pi) < 1e-6
