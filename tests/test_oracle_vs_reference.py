"""Pins the oracle's curve maths bit for bit against the reference's own
sources, compiled from the reference tree into oracle/_ref/libref_math.so by
oracle/ref.mk (built by __graft_entry__.build() where the reference tree is
present).  Random and adversarial inputs: cubics from the unit-square corners,
random fp32 cubics over several magnitudes, random matrices, radii across
eight decades, and a fast_acos sweep.  Skipped where the library was not
built.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_binding as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_math.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_LIB),
                                reason="oracle/_ref/libref_math.so not built (no reference tree)")

FP = ctypes.POINTER(ctypes.c_float)


def ref():
    L = ctypes.CDLL(REF_LIB)
    f = ctypes.c_float
    for n in ("ref_wang_cubic_pow4", "ref_wang_cubic"):
        getattr(L, n).restype = f
        getattr(L, n).argtypes = [FP, f, FP]
    L.ref_fast_acos.restype = f
    L.ref_fast_acos.argtypes = [f]
    L.ref_polar_segments_per_radian.restype = f
    L.ref_polar_segments_per_radian.argtypes = [f]
    L.ref_find_max_scale.restype = f
    L.ref_find_max_scale.argtypes = [FP]
    L.ref_eval_cubic.argtypes = [FP, f, FP]
    L.ref_chop_cubic_at.argtypes = [FP, f, FP]
    L.ref_chop_cubic_at2.argtypes = [FP, f, f, FP]
    L.ref_convex_180_chops.argtypes = [FP, FP, ctypes.POINTER(ctypes.c_int)]
    L.ref_convex_180_chops.restype = ctypes.c_int
    return L


def fp(a):
    return a.ctypes.data_as(FP)


def bits(x):
    return np.float32(x).view(np.uint32)


def cubics(rng, n):
    out = []
    for i in range(256):  # every cubic on the corners of the unit square, x100
        out.append(np.array([((i >> 0) & 1, (i >> 1) & 1), ((i >> 2) & 1, (i >> 3) & 1),
                             ((i >> 4) & 1, (i >> 5) & 1), ((i >> 6) & 1, (i >> 7) & 1)],
                            np.float32) * 100)
    for scale in (1e-2, 1.0, 100.0, 2000.0, 1e5):
        for _ in range(n):
            out.append((rng.random((4, 2)) * scale).astype(np.float32))
    return [np.ascontiguousarray(c.ravel()) for c in out]


def matrices(rng, n):
    out = [np.array([1, 0, 0, 1, 0, 0], np.float32)]
    for _ in range(n):
        m = rng.normal(size=6).astype(np.float32) * np.float32(10 ** rng.uniform(-2, 2))
        out.append(m)
        out.append(np.array([m[0], 0, 0, m[3], m[4], m[5]], np.float32))
    return out


def test_wang_bit_exact():
    R, O = ref(), ob.lib()
    rng = np.random.default_rng(7)
    for m in matrices(rng, 20):
        for c in cubics(rng, 40):
            for prec in (4.0, 1.0):
                a = O.oracle_wang_cubic_pow4(fp(c), prec, fp(m))
                b = R.ref_wang_cubic_pow4(fp(c), prec, fp(m))
                assert bits(a) == bits(b), (c, m, prec, a, b)


def test_polar_density_and_fast_acos_bit_exact():
    R, O = ref(), ob.lib()
    for r in np.geomspace(1e-3, 1e5, 4001).astype(np.float32):
        assert bits(O.oracle_polar_segments_per_radian(float(r))) == \
            bits(R.ref_polar_segments_per_radian(float(r)))
    for x in np.linspace(-1, 1, 40001, dtype=np.float32):
        assert bits(O.oracle_fast_acos(float(x))) == bits(R.ref_fast_acos(float(x)))


def test_find_max_scale_bit_exact():
    R, O = ref(), ob.lib()
    rng = np.random.default_rng(8)
    for m in matrices(rng, 3000):
        assert bits(O.oracle_find_max_scale(fp(m))) == bits(R.ref_find_max_scale(fp(m)))


def test_eval_and_chops_bit_exact():
    R, O = ref(), ob.lib()
    rng = np.random.default_rng(9)
    a7, b7 = np.zeros(14, np.float32), np.zeros(14, np.float32)
    a10, b10 = np.zeros(20, np.float32), np.zeros(20, np.float32)
    a2, b2 = np.zeros(2, np.float32), np.zeros(2, np.float32)
    for c in cubics(rng, 30):
        for t in (0.0, 1e-3, 0.25, 1 / 3, 0.5, 0.9, 1.0):
            t = float(np.float32(t))
            O.oracle_eval_cubic(fp(c), t, fp(a2))
            R.ref_eval_cubic(fp(c), t, fp(b2))
            assert a2.tobytes() == b2.tobytes()
            O.oracle_chop_cubic_at(fp(c), t, fp(a7))
            R.ref_chop_cubic_at(fp(c), t, fp(b7))
            assert a7.tobytes() == b7.tobytes()
            t1 = float(np.float32(min(1.0, t + 0.3)))
            O.oracle_chop_cubic_at2(fp(c), t, t1, fp(a10))
            R.ref_chop_cubic_at2(fp(c), t, t1, fp(b10))
            assert a10.tobytes() == b10.tobytes()


def test_convex_180_chops_bit_exact():
    R, O = ref(), ob.lib()
    rng = np.random.default_rng(10)
    ta, tb = np.zeros(2, np.float32), np.zeros(2, np.float32)
    ca, cb = ctypes.c_int(), ctypes.c_int()
    eps = 1.0 / (1 << 11)
    h = (1 - eps * eps) / (3 * eps * eps + 1)
    dy = (1 - h) / 2
    extra = [np.array([0, 0, 1, 1 - k * dy, 1, k * dy, 0, 1], np.float32) for k in (0.5, .9, 1, 1.1, 4)]
    for c in cubics(rng, 300) + extra:
        na = O.oracle_convex_180_chops(fp(c), fp(ta), ctypes.byref(ca))
        nb = R.ref_convex_180_chops(fp(c), fp(tb), ctypes.byref(cb))
        assert na == nb and ca.value == cb.value, c
        assert ta[:na].tobytes() == tb[:nb].tobytes(), c
