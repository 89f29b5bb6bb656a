"""Writes tests/golden/ref_math_vectors.npz: inputs and outputs of the
reference's own curve maths (Wang's formula, fast_acos, polar segment
density, Mat2D::findMaxScale, eval/chop of cubics, convex-180 chops), computed
by oracle/_ref/libref_math.so, which oracle/ref.mk compiles from the reference
sources.  Run where that library is built:

    python tests/golden/make_ref_math_vectors.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_math.so")
FP = ctypes.POINTER(ctypes.c_float)


def main():
    L = ctypes.CDLL(LIB)
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
    L.ref_chop_cubic_at.argtypes = [FP, f, FP]
    L.ref_convex_180_chops.argtypes = [FP, FP, ctypes.POINTER(ctypes.c_int)]
    L.ref_convex_180_chops.restype = ctypes.c_int
    p = lambda a: a.ctypes.data_as(FP)

    rng = np.random.default_rng(20261202)
    corners = np.array([[((i >> 0) & 1, (i >> 1) & 1), ((i >> 2) & 1, (i >> 3) & 1),
                         ((i >> 4) & 1, (i >> 5) & 1), ((i >> 6) & 1, (i >> 7) & 1)]
                        for i in range(256)], np.float32).reshape(256, 8) * 100
    rand = np.concatenate([(rng.random((200, 8)) * s).astype(np.float32)
                           for s in (1e-2, 1.0, 1e2, 2e3, 1e5)])
    cubics = np.ascontiguousarray(np.concatenate([corners, rand]))
    mats = np.concatenate([np.array([[1, 0, 0, 1, 0, 0]], np.float32),
                           (rng.normal(size=(63, 6)) * 10 ** rng.uniform(-2, 2, (63, 1))).astype(np.float32)])
    wang = np.zeros((len(cubics), 4), np.float32)
    for i, c in enumerate(cubics):
        for k, m in enumerate(mats[[0, 1 + i % 63]]):
            m = np.ascontiguousarray(m)
            wang[i, 2 * k] = L.ref_wang_cubic_pow4(p(c), 4.0, p(m))
            wang[i, 2 * k + 1] = L.ref_wang_cubic_pow4(p(c), 1.0, p(m))
    acos_x = np.linspace(-1, 1, 4001, dtype=np.float32)
    acos_y = np.array([L.ref_fast_acos(float(x)) for x in acos_x], np.float32)
    radii = np.geomspace(1e-3, 1e5, 801).astype(np.float32)
    polar = np.array([L.ref_polar_segments_per_radian(float(r)) for r in radii], np.float32)
    scale_m = np.ascontiguousarray((rng.normal(size=(500, 6)) * 10 ** rng.uniform(-3, 3, (500, 1))).astype(np.float32))
    scale = np.array([L.ref_find_max_scale(p(np.ascontiguousarray(m))) for m in scale_m], np.float32)
    ts = np.array([0, 1e-3, .25, 1 / 3, .5, .9, 1], np.float32)
    evals = np.zeros((len(cubics), len(ts), 2), np.float32)
    chops = np.zeros((len(cubics), len(ts), 14), np.float32)
    out2, out14 = np.zeros(2, np.float32), np.zeros(14, np.float32)
    for i, c in enumerate(cubics):
        for k, t in enumerate(ts):
            L.ref_eval_cubic(p(c), float(t), p(out2))
            evals[i, k] = out2
            L.ref_chop_cubic_at(p(c), float(t), p(out14))
            chops[i, k] = out14
    n180 = np.zeros(len(cubics), np.int32)
    t180 = np.zeros((len(cubics), 2), np.float32)
    cusp = np.zeros(len(cubics), np.int32)
    T, cu = np.zeros(2, np.float32), ctypes.c_int()
    for i, c in enumerate(cubics):
        n180[i] = L.ref_convex_180_chops(p(c), p(T), ctypes.byref(cu))
        t180[i] = T
        cusp[i] = cu.value
    np.savez_compressed(os.path.join(HERE, "ref_math_vectors.npz"), cubics=cubics, mats=mats, wang=wang,
                        acos_x=acos_x, acos_y=acos_y, radii=radii, polar=polar, scale_m=scale_m,
                        scale=scale, ts=ts, evals=evals, chops=chops, n180=n180, t180=t180, cusp=cusp)


if __name__ == "__main__":
    main()
