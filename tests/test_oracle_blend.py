"""Pins the oracle's advanced blend modes (advanced_blend.glsl) against the
reference's own test cases (tests/unit_tests/renderer/advanced_blend_test.cpp:
colordodge/colorburn special values, random closed-form checks, HSL modes
against the NV_blend_equation_advanced spec functions) and the KHR closed
forms of the separable modes, then checks blended pixels end to end.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_binding as ob
from streams import StreamBuilder, rect

INF, NAN = float("inf"), float("nan")
FP = ctypes.POINTER(ctypes.c_float)
MODES = {"screen": 1, "overlay": 2, "darken": 3, "lighten": 4, "colordodge": 5, "colorburn": 6,
         "hardlight": 7, "softlight": 8, "difference": 9, "exclusion": 10, "multiply": 11,
         "hue": 12, "saturation": 13, "color": 14, "luminosity": 15}
RIVE_MODES = {"screen": 14, "overlay": 15, "darken": 16, "lighten": 17, "colordodge": 18,
              "colorburn": 19, "hardlight": 20, "softlight": 21, "difference": 22, "exclusion": 23,
              "multiply": 24, "hue": 25, "saturation": 26, "color": 27, "luminosity": 28}


def lib():
    L = ob.lib()
    L.oracle_blend_coeffs.argtypes = [FP, FP, ctypes.c_uint32, FP]
    L.oracle_blend_coeffs.restype = None
    L.oracle_shader_blend_mode.argtypes = [ctypes.c_uint32]
    L.oracle_shader_blend_mode.restype = ctypes.c_uint32
    return L


def coeffs(src, dst_premul, mode):
    s = np.asarray(src, np.float32)
    d = np.asarray(dst_premul, np.float32)
    out = np.zeros(3, np.float32)
    lib().oracle_blend_coeffs(s.ctypes.data_as(FP), d.ctypes.data_as(FP), MODES[mode], out.ctypes.data_as(FP))
    return out


def with_alpha(src, dst, a, mode):
    return coeffs(src, [dst[0] * a, dst[1] * a, dst[2] * a, a], mode)


def test_shader_mode_mapping():
    L = lib()
    assert L.oracle_shader_blend_mode(3) == 0
    for name, rive in RIVE_MODES.items():
        assert L.oracle_shader_blend_mode(rive) == MODES[name]


def test_colordodge_special_values():
    for s in ([0, 0, 0], [1, 1.001, INF], [-.0001, -1, -INF]):
        for d in ([0, 0, 0, 0], [.0001, 1, INF, 0], [-.0001, -1, -INF, 0]):
            assert np.all(coeffs(s, d, "colordodge") == 0)
    for a in (0, 1 / 255, .25, .5, 254 / 255, 1, 256 / 255):
        for s in ([0, 0, 0], [1, 1, 1], [-INF, INF, NAN]):
            for d in ([0, -.001, -1], [-10, -100, -INF]):
                assert np.all(with_alpha(s, d, a, "colordodge") == 0)
        if a != 0:
            for s in ([1, 1.00001, 2], [10, 100, INF]):
                for d in (1e-10, 1, 1.0001, INF):
                    assert np.all(with_alpha(s, [d] * 3, a, "colordodge") == 1)
        rng = np.random.default_rng(0)
        for _ in range(100):
            s = rng.random(3).astype(np.float32)
            d = np.zeros(3, np.float32) if a == 0 else rng.random(3).astype(np.float32)
            c = with_alpha(s, d, a, "colordodge")
            for i in range(3):
                if d[i] <= 0:
                    assert c[i] == 0
                elif s[i] >= 1:
                    assert c[i] == 1
                else:
                    assert c[i] == pytest.approx(min(1, d[i] / (1 - s[i])), rel=1e-4, abs=1e-5)


def test_colorburn_special_values():
    for a in (0, 1 / 255, .25, .5, 254 / 255, 1, 256 / 255):
        if a != 0:
            for s in ([0, 0, 0], [1, 1, 1], [-INF, INF, NAN]):
                for d in ([1, 1.001, 2], [10, 100, INF]):
                    assert np.all(with_alpha(s, d, a, "colorburn") == 1)
        for s in ([0, -1e-1, -1], [-10, -100, -INF]):
            for d in (1 - 1e-6, 0, -1e-6, -INF):
                assert np.all(with_alpha(s, [d] * 3, a, "colorburn") == 0)
        rng = np.random.default_rng(1)
        for _ in range(100):
            s = rng.random(3).astype(np.float32)
            d = np.zeros(3, np.float32) if a == 0 else rng.random(3).astype(np.float32)
            c = with_alpha(s, d, a, "colorburn")
            for i in range(3):
                if d[i] >= 1 and a != 0:
                    assert c[i] == 1
                elif s[i] <= 0:
                    assert c[i] == 0
                else:
                    assert c[i] == pytest.approx(1 - min(1, (1 - d[i]) / s[i]), rel=1e-4, abs=1e-4)


# NV_blend_equation_advanced spec functions (what the reference test compares to)
def lum(c):
    return float(np.dot(c, [.30, .59, .11]))


def clip_color(c):
    l, mn, mx = lum(c), c.min(), c.max()
    if mn < 0:
        c = l + ((c - l) * l) / (l - mn)
    if mx > 1:
        c = l + ((c - l) * (1 - l)) / (mx - l)
    return c


def set_lum(base, lumc):
    return clip_color(base + (lum(lumc) - lum(base)))


def set_lum_sat(base, sat, lumc):
    sb = base.max() - base.min()
    c = (base - base.min()) * (sat.max() - sat.min()) / sb if sb > 0 else np.zeros(3)
    return set_lum(c, lumc)


SEPARABLE = {
    "multiply": lambda s, d: s * d,
    "screen": lambda s, d: s + d - s * d,
    "overlay": lambda s, d: np.where(d <= .5, 2 * s * d, 1 - 2 * (1 - s) * (1 - d)),
    "darken": np.minimum,
    "lighten": np.maximum,
    "hardlight": lambda s, d: np.where(s <= .5, 2 * s * d, 1 - 2 * (1 - s) * (1 - d)),
    "softlight": lambda s, d: np.where(
        s <= .5, d - (1 - 2 * s) * d * (1 - d),
        np.where(d <= .25, d + (2 * s - 1) * d * ((16 * d - 12) * d + 3), d + (2 * s - 1) * (np.sqrt(d) - d))),
    "difference": lambda s, d: np.abs(d - s),
    "exclusion": lambda s, d: s + d - 2 * s * d,
    "hue": lambda s, d: set_lum_sat(s, d, d),
    "saturation": lambda s, d: set_lum_sat(d, s, d),
    "color": lambda s, d: set_lum(s, d),
    "luminosity": lambda s, d: set_lum(d, s),
}


@pytest.mark.parametrize("mode", sorted(SEPARABLE))
def test_modes_match_spec_on_color_grid(mode):
    steps = np.floor(np.linspace(0, 1, 6) * 255) / 255
    grid = np.array(np.meshgrid(steps, steps, steps, indexing="ij")).reshape(3, -1).T
    rng = np.random.default_rng(2)
    for s in grid:
        for d in grid[rng.choice(len(grid), 12, replace=False)]:
            got = with_alpha(s, d, 1.0, mode)
            want = SEPARABLE[mode](s.astype(np.float64), d.astype(np.float64))
            assert np.abs(got - want).max() <= 1e-4, (mode, s, d, got, want)


def render(b):
    s = b.build()
    f = ob.OracleFlush(s)
    try:
        img, _ = f.render()
    finally:
        f.close()
    return img


@pytest.mark.parametrize("mode", ["multiply", "screen", "difference", "darken", "lighten"])
def test_blended_pixels_closed_form(mode):
    """Opaque destination D, then an opaque rect S in `mode`: pixel = f(S, D)."""
    D = np.array([0x40, 0x90, 0xe0]) / 255
    S = np.array([0xc0, 0x60, 0x20]) / 255
    b = StreamBuilder(16, 16)
    b.draw(b.path(rect(0, 0, 16, 16)), color=0xff4090e0)
    b.draw(b.path(rect(4, 4, 12, 12)), color=0xffc06020)
    b.draws[-1]["blend_mode"] = RIVE_MODES[mode]
    img = render(b).astype(float) / 255
    want = SEPARABLE[mode](S, D)
    assert np.abs(img[8, 8, :3] - want).max() <= 1 / 255 + 1e-6
    assert np.abs(img[1, 1, :3] - D).max() <= 1 / 255 + 1e-6
    assert img[8, 8, 3] == 1.0
