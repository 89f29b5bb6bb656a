"""Hand-built draw streams (rive_stream, include/rive_hip.h) for edge-case
and closed-form tests.

    b = StreamBuilder(64, 64, dither=0)
    sq = b.path([(0, 10, 10), (1, 30, 10), (1, 30, 30), (1, 10, 30), (5,)])
    b.draw(sq, color=0xffff0000)
    s = b.build()          # ctypes pointer, valid while `b` lives

Verbs: (0, x, y) move, (1, x, y) line, (4, x1, y1, x2, y2, x3, y3) cubic,
(5,) close.
"""
import ctypes

import numpy as np

import rivehip

NONZERO, EVENODD, CLOCKWISE = 0, 1, 2
FILL, STROKE = 0, 1
MITER, ROUND, BEVEL = 0, 1, 2
BUTT, ROUND_CAP, SQUARE = 0, 1, 2

_PATH = np.dtype([("first_verb", "<u4"), ("verb_count", "<u4"), ("first_point", "<u4"),
                  ("point_count", "<u4"), ("fill_rule", "<u4")])
_DRAW = np.dtype([("path", "<u4"), ("matrix", "<f4", 6), ("style", "<u4"), ("thickness", "<f4"),
                  ("join", "<u4"), ("cap", "<u4"), ("feather", "<f4"), ("blend_mode", "<u4"),
                  ("paint_type", "<u4"), ("color", "<u4"), ("first_stop", "<u4"),
                  ("stop_count", "<u4"), ("grad", "<f4", 4), ("clip_count", "<u4"),
                  ("clip_paths", "<u4", 2), ("image", "<u4"), ("image_sampler", "<u4"),
                  ("opacity", "<f4")])
_IMAGE = np.dtype([("width", "<u4"), ("height", "<u4"), ("rgba", "<u8")])
assert _PATH.itemsize == 20 and _DRAW.itemsize == 108 and _IMAGE.itemsize == 16

IDENTITY = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)


def rect(x0, y0, x1, y1, clockwise=True):
    """Closed rectangle; clockwise in y-down screen space by default."""
    if clockwise:
        return [(0, x0, y0), (1, x1, y0), (1, x1, y1), (1, x0, y1), (5,)]
    return [(0, x0, y0), (1, x0, y1), (1, x1, y1), (1, x1, y0), (5,)]


class StreamBuilder:
    def __init__(self, width, height, clear_color=0, dither=0):
        self.width, self.height = width, height
        self.clear_color, self.dither = clear_color, dither
        self.verbs, self.points, self.paths, self.draws = [], [], [], []
        self.stop_colors, self.stop_offsets = [], []
        self.images = []

    def image(self, rgba_premul):
        """Adds a premultiplied RGBA8 image (h, w, 4); returns its 1-based index."""
        self.images.append(np.ascontiguousarray(rgba_premul, dtype=np.uint8))
        return len(self.images)

    def draw_image(self, image, matrix=IDENTITY, blend=3, sampler=0, opacity=1.0, clips=()):
        d = np.zeros((), _DRAW)
        d["matrix"] = matrix
        d["blend_mode"] = blend
        d["image"] = image
        d["image_sampler"] = sampler
        d["opacity"] = opacity
        d["clip_count"] = len(clips)
        for i, c in enumerate(clips):
            d["clip_paths"][i] = c
        self.draws.append(d)

    def path(self, cmds, fill_rule=NONZERO):
        first_verb, first_point = len(self.verbs), len(self.points) // 2
        for c in cmds:
            self.verbs.append(c[0])
            self.points.extend(float(v) for v in c[1:])
        self.paths.append((first_verb, len(self.verbs) - first_verb, first_point,
                           len(self.points) // 2 - first_point, fill_rule))
        return len(self.paths) - 1

    def draw(self, path, color=0xff000000, matrix=IDENTITY, stroke=None, join=MITER, cap=BUTT,
             gradient=None, clips=(), feather=0.0, blend=3):
        """stroke: width or None.  gradient: ("linear", sx, sy, ex, ey, stops)
        or ("radial", cx, cy, r, stops) with stops [(offset, ARGB), ...]."""
        d = np.zeros((), _DRAW)
        d["path"] = path
        d["matrix"] = matrix
        d["style"] = STROKE if stroke is not None else FILL
        d["thickness"] = stroke if stroke is not None else 0.0
        d["join"], d["cap"], d["feather"], d["blend_mode"] = join, cap, feather, blend
        d["color"] = color
        if gradient is not None:
            kind, *geo, stops = gradient
            d["paint_type"] = 1 if kind == "linear" else 2
            d["first_stop"], d["stop_count"] = len(self.stop_colors), len(stops)
            for off, col in stops:
                self.stop_offsets.append(off)
                self.stop_colors.append(col)
            g = list(geo) + [0.0] * (4 - len(geo))
            d["grad"] = g
        d["clip_count"] = len(clips)
        for i, c in enumerate(clips):
            d["clip_paths"][i] = c
        self.draws.append(d)

    def build(self):
        self._verbs = np.asarray(self.verbs, np.uint8)
        self._points = np.asarray(self.points, np.float32)
        self._paths = np.array(self.paths, dtype=_PATH) if self.paths else np.zeros(0, _PATH)
        self._draws = np.zeros(len(self.draws), _DRAW)
        for i, d in enumerate(self.draws):
            self._draws[i] = d
        self._colors = np.asarray(self.stop_colors or [0], np.uint32)
        self._offsets = np.asarray(self.stop_offsets or [0.0], np.float32)
        s = rivehip.Stream()
        s.width, s.height = self.width, self.height
        s.clear_color, s.dither = self.clear_color, self.dither
        s.verbs, s.verb_count = self._verbs.ctypes.data, len(self._verbs)
        s.points, s.point_count = self._points.ctypes.data, len(self._points) // 2
        s.paths, s.path_count = self._paths.ctypes.data, len(self._paths)
        s.draws, s.draw_count = self._draws.ctypes.data, len(self._draws)
        s.stop_colors, s.stop_offsets = self._colors.ctypes.data, self._offsets.ctypes.data
        s.stop_count = len(self.stop_colors)
        self._images = np.zeros(len(self.images), _IMAGE)
        for i, img in enumerate(self.images):
            self._images[i]["width"] = img.shape[1]
            self._images[i]["height"] = img.shape[0]
            self._images[i]["rgba"] = img.ctypes.data
        s.images = self._images.ctypes.data if len(self.images) else None
        s.image_count = len(self.images)
        self._struct = s
        return ctypes.pointer(s)


# Streams that cross one logical flush's resource limits
# (render_context.cpp:43-53,484-492,517-541,617-654).

def pixel_rects_stream(n, width=256):
    """n opaque 1x1-px rectangles, draw i on pixel (i % width, i // width),
    each with its own color: past 30 719 draws the path IDs run out.
    Returns (builder, ARGB colors)."""
    height = (n + width - 1) // width
    b = StreamBuilder(width, height, clear_color=0xff000000)
    colors = []
    for i in range(n):
        x, y = i % width, i // width
        p = b.path([(0, x, y), (1, x + 1, y), (1, x + 1, y + 1), (1, x, y + 1), (5,)])
        c = 0xff000000 | ((i * 2654435761) & 0xffffff)
        b.draw(p, color=c)
        colors.append(c)
    return b, np.asarray(colors, np.uint32)


def contour_limit_stream(draws=66):
    """A 1000-contour path drawn `draws` times: 65 draws fit under the
    65 535 contour IDs of a flush."""
    b = StreamBuilder(128, 128)
    cmds = []
    for k in range(1000):
        x, y = (k % 40) * 3, (k // 40) * 3
        cmds += [(0, x, y), (1, x + 2, y), (1, x + 2, y + 2), (1, x, y + 2), (5,)]
    p = b.path(cmds)
    for i in range(draws):
        b.draw(p, color=0x80000000 | (i * 0x030507))
    return b


def tess_limit_stream(draws=120):
    """Identical wide round-joined strokes of a 300-cubic zigzag, ~38 k
    tessellation vertices each: a flush holds about 109 of them."""
    b = StreamBuilder(512, 512)
    cmds = [(0, 10, 10)]
    for k in range(300):
        y = 10 + 1.6 * k
        cmds.append((4, 500, y, 10, y + 0.4, 500, y + 0.8))
    p = b.path(cmds)
    for i in range(draws):
        b.draw(p, color=0x20000000 | (i * 0x010203), stroke=300.0, join=ROUND, cap=ROUND_CAP)
    return b


def gradient_limit_stream(draws=2049):
    """`draws` distinct three-stop gradients, one complex ramp row each: the
    gradient texture holds 2048 rows."""
    b = StreamBuilder(64, 64)
    p = b.path([(0, 1, 1), (1, 9, 1), (1, 9, 9), (1, 1, 9), (5,)])
    for k in range(draws):
        stops = [(0.0, 0xff000000 | k), (0.5, 0xffff0000), (1.0, 0xff0000ff)]
        b.draw(p, gradient=("linear", 1.0, 1.0, 9.0, 9.0, stops))
    return b
