"""The oracle against committed outputs of the reference's own curve maths
(tests/golden/ref_math_vectors.npz, written by
tests/golden/make_ref_math_vectors.py from the reference sources compiled by
oracle/ref.mk).  Bit-exact; runs anywhere, including where the reference tree
is absent."""
import ctypes
import os

import numpy as np

import oracle_binding as ob

HERE = os.path.dirname(os.path.abspath(__file__))
V = np.load(os.path.join(HERE, "golden", "ref_math_vectors.npz"))
FP = ctypes.POINTER(ctypes.c_float)


def p(a):
    return a.ctypes.data_as(FP)


def test_wang_pow4_matches_reference_vectors():
    O = ob.lib()
    cubics, mats = V["cubics"], V["mats"]
    got = np.zeros_like(V["wang"])
    for i, c in enumerate(cubics):
        c = np.ascontiguousarray(c)
        for k, m in enumerate(mats[[0, 1 + i % 63]]):
            m = np.ascontiguousarray(m)
            got[i, 2 * k] = O.oracle_wang_cubic_pow4(p(c), 4.0, p(m))
            got[i, 2 * k + 1] = O.oracle_wang_cubic_pow4(p(c), 1.0, p(m))
    assert got.tobytes() == V["wang"].tobytes()


def test_acos_polar_scale_match_reference_vectors():
    O = ob.lib()
    acos = np.array([O.oracle_fast_acos(float(x)) for x in V["acos_x"]], np.float32)
    assert acos.tobytes() == V["acos_y"].tobytes()
    polar = np.array([O.oracle_polar_segments_per_radian(float(r)) for r in V["radii"]], np.float32)
    assert polar.tobytes() == V["polar"].tobytes()
    scale = np.array([O.oracle_find_max_scale(p(np.ascontiguousarray(m))) for m in V["scale_m"]], np.float32)
    assert scale.tobytes() == V["scale"].tobytes()


def test_eval_chop_convex180_match_reference_vectors():
    O = ob.lib()
    out2, out14, T = np.zeros(2, np.float32), np.zeros(14, np.float32), np.zeros(2, np.float32)
    cu = ctypes.c_int()
    for i, c in enumerate(V["cubics"]):
        c = np.ascontiguousarray(c)
        for k, t in enumerate(V["ts"]):
            O.oracle_eval_cubic(p(c), float(t), p(out2))
            assert out2.tobytes() == V["evals"][i, k].tobytes()
            O.oracle_chop_cubic_at(p(c), float(t), p(out14))
            assert out14.tobytes() == V["chops"][i, k].tobytes()
        n = O.oracle_convex_180_chops(p(c), p(T), ctypes.byref(cu))
        assert n == V["n180"][i] and cu.value == V["cusp"][i]
        assert T[:n].tobytes() == V["t180"][i, :n].tobytes()
