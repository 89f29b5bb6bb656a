"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test use only."""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "rive-runtime_amd"))
import rivehip  # noqa: E402

ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/liboracle.so is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(ORACLE_PATH)
        vp, f, fp = ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(ctypes.c_float)
        L.oracle_build.restype = vp
        L.oracle_build.argtypes = [ctypes.POINTER(rivehip.Stream)]
        L.oracle_free.argtypes = [vp]
        L.oracle_free.restype = None
        L.oracle_set_interior_triangulation.argtypes = [ctypes.c_int]
        L.oracle_set_interior_triangulation.restype = None
        L.oracle_get.restype = ctypes.c_size_t
        L.oracle_get.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
        L.oracle_tessellate.restype = ctypes.c_size_t
        L.oracle_tessellate.argtypes = [vp, vp, ctypes.c_size_t]
        L.oracle_ramps.restype = ctypes.c_size_t
        L.oracle_ramps.argtypes = [vp, vp, ctypes.c_size_t]
        L.oracle_render.restype = ctypes.c_double
        L.oracle_render.argtypes = [vp, vp]
        L.oracle_render_edges.restype = ctypes.c_double
        L.oracle_render_edges.argtypes = [vp, vp, vp]
        L.oracle_flush_count.restype = ctypes.c_uint32
        L.oracle_flush_count.argtypes = [vp]
        L.oracle_select.restype = ctypes.c_int
        L.oracle_select.argtypes = [vp, ctypes.c_uint32]
        for name in ["oracle_wang_cubic", "oracle_wang_cubic_pow4"]:
            getattr(L, name).restype = f
            getattr(L, name).argtypes = [fp, f, fp]
        L.oracle_fast_acos.restype = f
        L.oracle_fast_acos.argtypes = [f]
        L.oracle_polar_segments_per_radian.restype = f
        L.oracle_polar_segments_per_radian.argtypes = [f]
        L.oracle_find_max_scale.restype = f
        L.oracle_find_max_scale.argtypes = [fp]
        L.oracle_eval_cubic.argtypes = [fp, f, fp]
        L.oracle_eval_cubic.restype = None
        L.oracle_chop_cubic_at.argtypes = [fp, f, fp]
        L.oracle_chop_cubic_at.restype = None
        L.oracle_chop_cubic_at2.argtypes = [fp, f, f, fp]
        L.oracle_chop_cubic_at2.restype = None
        L.oracle_convex_180_chops.argtypes = [fp, fp, ctypes.POINTER(ctypes.c_int)]
        L.oracle_convex_180_chops.restype = ctypes.c_int
        u16p = ctypes.POINTER(ctypes.c_uint16)
        L.oracle_feather_tables.argtypes = [u16p, u16p]
        L.oracle_feather_tables.restype = None
        L.oracle_table_lookup.argtypes = [u16p, f]
        L.oracle_table_lookup.restype = f
        for name in ["oracle_feather", "oracle_inverse_feather"]:
            getattr(L, name).argtypes = [f]
            getattr(L, name).restype = f
        for name in ["oracle_eval_feathered_fill", "oracle_eval_feathered_stroke"]:
            getattr(L, name).argtypes = [fp]
            getattr(L, name).restype = f
        L.oracle_pack_feathered_fill.argtypes = [f, f, f, f, fp]
        L.oracle_pack_feathered_fill.restype = None
        L.oracle_cubic_max_height.argtypes = [fp, fp]
        L.oracle_cubic_max_height.restype = f
        L.oracle_cubic_local_curvature.argtypes = [fp, f, f]
        L.oracle_cubic_local_curvature.restype = f
        L.oracle_soften_path.argtypes = [ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32, fp, ctypes.c_uint32, f, f,
                                         ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32, fp, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_soften_path.restype = ctypes.c_uint32
        _lib = L
    return _lib


RECORD_DTYPES = {0: rivehip.TESS_SPAN, 1: rivehip.CONTOUR, 2: rivehip.PATH_DATA,
                 3: rivehip.PAINT_DATA, 4: rivehip.PAINT_AUX, 5: rivehip.GRAD_SPAN}


def set_interior_triangulation(enabled):
    """Lay eligible fills out as outer-cubic patches (draw.cpp:510-538, the
    triangulation controller's guards alone) in frames built from now on."""
    lib().oracle_set_interior_triangulation(1 if enabled else 0)


class OracleFlush:
    """A frame laid out by the oracle's restatement of RiveRenderer and
    RenderContext: one or more logical flushes (the reference splits a frame
    where its resource limits require).  Record getters read the selected
    logical flush (the first by default); render() renders all of them."""

    def __init__(self, stream_ptr):
        self.L = lib()
        self.h = self.L.oracle_build(stream_ptr)
        self.flush_count = int(self.L.oracle_flush_count(self.h))
        self.select(0)

    def select(self, index):
        """Points the record getters at logical flush `index`."""
        if self.L.oracle_select(self.h, index) != 0:
            raise IndexError(index)
        self.index = index
        info = np.zeros(10, np.uint32)
        self.L.oracle_get(self.h, 8, info.ctypes.data, info.nbytes)
        (self.path_count, self.contour_count, self.span_count, self.tess_vertex_count,
         self.grad_simple_rows, self.grad_complex_rows, self.grad_height, self.draw_count,
         self.width, self.height) = [int(v) for v in info]
        return self

    def close(self):
        if self.h:
            self.L.oracle_free(self.h)
            self.h = None

    def get(self, what):
        n = self.L.oracle_get(self.h, what, None, 0)
        buf = np.empty(n, np.uint8)
        self.L.oracle_get(self.h, what, buf.ctypes.data, n)
        if what in RECORD_DTYPES:
            return buf.view(RECORD_DTYPES[what])
        return buf.view(np.uint32)

    def draws(self):
        """(pathID, tessLocation, tessVertexCount, dirs, isStroke) per draw."""
        return self.get(7).reshape(-1, 5)

    def tessellate(self):
        n = self.L.oracle_tessellate(self.h, None, 0)
        buf = np.empty(n, np.uint8)
        self.L.oracle_tessellate(self.h, buf.ctypes.data, n)
        return buf.view(rivehip.TESS_TEXEL)

    def ramps(self):
        n = self.L.oracle_ramps(self.h, None, 0)
        buf = np.empty(n, np.uint8)
        self.L.oracle_ramps(self.h, buf.ctypes.data, n)
        return buf.view(np.uint32)

    def render(self):
        """The whole frame (every logical flush): (RGBA8 image, seconds)."""
        img = np.zeros((self.height, self.width, 4), np.uint8)
        secs = self.L.oracle_render(self.h, img.ctypes.data)
        return img, secs

    def render_edges(self):
        """The frame plus a mask of pixels some draw covers fractionally
        (its final coverage strictly between 0 and 1): (image, mask, secs)."""
        img = np.zeros((self.height, self.width, 4), np.uint8)
        mask = np.zeros((self.height, self.width), np.uint8)
        secs = self.L.oracle_render_edges(self.h, img.ctypes.data, mask.ctypes.data)
        return img, mask.astype(bool), secs

    def draw_records(self):
        """rive_hip_draw records for a SPANS-mode flush of this layout."""
        d = self.draws()
        out = np.zeros(len(d), rivehip.DRAW)
        out["path_id"] = d[:, 0]
        out["tess_location"] = d[:, 1]
        out["tess_vertex_count"] = d[:, 2]
        out["contour_directions"] = d[:, 3]
        out["is_stroke"] = d[:, 4]
        return out

    def flush_spans_on(self, ctx, target, load_action=0, clear_color=0, dither=1):
        """Uploads this layout to a HIP context and flushes it in SPANS mode."""
        ctx.upload("path", self.get(2))
        ctx.upload("paint", self.get(3))
        ctx.upload("paint_aux", self.get(4))
        ctx.upload("contour", self.get(1))
        ctx.upload("tess_span", self.get(0))
        ctx.upload("grad_span", self.get(5))
        ctx.upload("draw", self.draw_records())
        desc = rivehip.FlushDesc()
        desc.target = target
        desc.mode = rivehip.FLUSH_SPANS
        desc.load_action = load_action
        desc.clear_color = clear_color
        desc.dither = dither
        desc.path_count = self.path_count
        desc.contour_count = self.contour_count
        desc.tess_span_count = self.span_count
        desc.tess_vertex_count = self.tess_vertex_count
        desc.draw_count = self.draw_count
        desc.grad_span_count = len(self.get(5))
        desc.grad_simple_rows = self.grad_simple_rows
        desc.grad_complex_rows = self.grad_complex_rows
        ctx.flush(desc)

    def flush_all_spans_on(self, ctx, target, clear_color=0, dither=1):
        """The reference's multi-flush model in SPANS mode: every logical
        flush's records concatenated into shared buffers (one upload per
        buffer, as RenderContext maps them once per frame), then one
        rive_hip_flush per logical flush with first_* offsets into those
        buffers; flushes after the first load the target
        (render_context.cpp:1218-1222; gpu.hpp:1386-1395)."""
        keep = self.index
        parts = {w: [] for w in ("path", "paint", "paint_aux", "contour", "tess_span", "grad_span", "draw")}
        descs = []
        offs = {w: 0 for w in parts}
        for k in range(self.flush_count):
            self.select(k)
            recs = {"path": self.get(2), "paint": self.get(3), "paint_aux": self.get(4),
                    "contour": self.get(1), "tess_span": self.get(0), "grad_span": self.get(5),
                    "draw": self.draw_records()}
            d = rivehip.FlushDesc()
            d.target = target
            d.mode = rivehip.FLUSH_SPANS
            d.load_action = 0 if k == 0 else 1
            d.clear_color = clear_color
            d.dither = dither
            d.path_count = self.path_count
            d.contour_count = self.contour_count
            d.tess_span_count = self.span_count
            d.tess_vertex_count = self.tess_vertex_count
            d.draw_count = self.draw_count
            d.grad_span_count = len(recs["grad_span"])
            d.grad_simple_rows = self.grad_simple_rows
            d.grad_complex_rows = self.grad_complex_rows
            d.first_path, d.first_paint, d.first_paint_aux = offs["path"], offs["paint"], offs["paint_aux"]
            d.first_contour, d.first_tess_span = offs["contour"], offs["tess_span"]
            d.first_grad_span, d.first_draw = offs["grad_span"], offs["draw"]
            descs.append(d)
            for w, r in recs.items():
                parts[w].append(r)
                offs[w] += len(r)
        for w, rs in parts.items():
            ctx.upload(w, np.concatenate(rs))
        for d in descs:
            ctx.flush(d)
        self.select(keep)

