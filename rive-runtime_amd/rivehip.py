"""ctypes binding of librive_hip.so (include/rive_hip.h).

The library is built in-tree by __graft_entry__.build() (or `make` in this
directory).  Loading fails loudly if it is missing: there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librive_hip.so")

OK = 0
BUFFER = {
    "flush_uniform": 0, "path": 1, "paint": 2, "paint_aux": 3, "contour": 4,
    "grad_span": 5, "tess_span": 6, "triangle_vertex": 7, "image_draw": 8,
    "verb": 9, "point": 10, "draw": 11,
}
FLUSH_SPANS = 0
FLUSH_GEOMETRY = 1

# numpy mirrors of the record structs (byte-compatible with include/rive_hip.h)
TESS_SPAN = np.dtype([("pts", "<f4", 8), ("join_tangent", "<f4", 2), ("y", "<f4"),
                      ("reflection_y", "<f4"), ("x0x1", "<i4"), ("reflection_x0x1", "<i4"),
                      ("segment_counts", "<u4"), ("contour_id_with_flags", "<u4")])
CONTOUR = np.dtype([("midpoint", "<f4", 2), ("path_id", "<u4"), ("vertex_index0", "<u4")])
PATH_DATA = np.dtype([("matrix", "<f4", 6), ("stroke_radius", "<f4"), ("feather_radius", "<f4"),
                      ("z_index", "<u4"), ("atlas_transform", "<f4", 3), ("coverage_range", "<u4", 4)])
PAINT_DATA = np.dtype([("params", "<u4"), ("value", "<u4")])
PAINT_AUX = np.dtype([("paint_matrix", "<f4", 6), ("grad_span", "<f4", 2),
                      ("clip_rect_inverse_matrix", "<f4", 6), ("inverse_fwidth", "<f4", 2),
                      ("image_matrix", "<f4", 6), ("image_lod", "<f4"), ("pad", "<f4", 9)])
GRAD_SPAN = np.dtype([("horizontal_span", "<u4"), ("y_with_flags", "<u4"), ("color0", "<u4"),
                      ("color1", "<u4")])
DRAW = np.dtype([("path_id", "<u4"), ("first_verb", "<u4"), ("verb_count", "<u4"),
                 ("first_point", "<u4"), ("point_count", "<u4"), ("contour_directions", "<u4"),
                 ("contour_flags", "<u4"), ("join", "<u4"), ("cap", "<u4"), ("is_stroke", "<u4"),
                 ("matrix", "<f4", 6), ("stroke_radius", "<f4"), ("matrix_max_scale", "<f4"),
                 ("polar_segments_per_radian", "<f4"), ("parametric_precision", "<f4"),
                 ("tess_location", "<u4"), ("tess_vertex_count", "<u4"), ("image_index", "<u4"),
                 ("image_sampler", "<u4")])
TESS_TEXEL = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<u4"), ("w", "<u4")])
assert TESS_SPAN.itemsize == 64 and CONTOUR.itemsize == 16 and PATH_DATA.itemsize == 64
assert PAINT_DATA.itemsize == 8 and PAINT_AUX.itemsize == 128 and GRAD_SPAN.itemsize == 16
assert DRAW.itemsize == 96


class Stream(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("clear_color", ctypes.c_uint32), ("dither", ctypes.c_uint32),
                ("verbs", ctypes.c_void_p), ("verb_count", ctypes.c_uint32),
                ("points", ctypes.c_void_p), ("point_count", ctypes.c_uint32),
                ("paths", ctypes.c_void_p), ("path_count", ctypes.c_uint32),
                ("draws", ctypes.c_void_p), ("draw_count", ctypes.c_uint32),
                ("stop_colors", ctypes.c_void_p), ("stop_offsets", ctypes.c_void_p),
                ("stop_count", ctypes.c_uint32), ("images", ctypes.c_void_p),
                ("image_count", ctypes.c_uint32)]


class FlushDesc(ctypes.Structure):
    _fields_ = [("target", ctypes.c_void_p), ("mode", ctypes.c_uint32),
                ("load_action", ctypes.c_uint32), ("clear_color", ctypes.c_uint32),
                ("dither", ctypes.c_uint32), ("path_count", ctypes.c_uint32),
                ("contour_count", ctypes.c_uint32), ("tess_span_count", ctypes.c_uint32),
                ("tess_vertex_count", ctypes.c_uint32), ("draw_count", ctypes.c_uint32),
                ("grad_span_count", ctypes.c_uint32), ("grad_simple_rows", ctypes.c_uint32),
                ("grad_complex_rows", ctypes.c_uint32),
                ("first_path_tess_location", ctypes.c_uint32), ("images", ctypes.c_void_p),
                ("image_count", ctypes.c_uint32), ("canvases", ctypes.c_void_p),
                ("canvas_count", ctypes.c_uint32),
                ("first_path", ctypes.c_uint64), ("first_paint", ctypes.c_uint64),
                ("first_paint_aux", ctypes.c_uint64), ("first_contour", ctypes.c_uint64),
                ("first_grad_span", ctypes.c_uint64), ("first_tess_span", ctypes.c_uint64),
                ("first_draw", ctypes.c_uint64), ("patch_batches", ctypes.c_void_p),
                ("patch_batch_count", ctypes.c_uint32)]


class PatchBatch(ctypes.Structure):
    """rive_hip_patch_batch: one DrawBatch of midpoint-fan patches."""
    _fields_ = [("first_patch", ctypes.c_uint32), ("patch_count", ctypes.c_uint32),
                ("image_index", ctypes.c_uint32), ("image_sampler", ctypes.c_uint32)]


class Canvas(ctypes.Structure):
    """rive_hip_canvas: one canvas of a batched flush."""
    _fields_ = [("target", ctypes.c_void_p), ("first_path", ctypes.c_uint32),
                ("path_count", ctypes.c_uint32), ("clear_color", ctypes.c_uint32),
                ("load_action", ctypes.c_uint32)]


class FlushStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "width", "height", "load_action", "mode", "path_count", "draw_count", "contour_count",
        "span_count", "tess_vertex_count", "curve_record_count", "prim_count", "bin_entries",
        "tile_count", "grad_simple_rows", "grad_complex_rows", "aux_path_count", "grad_span_count",
        "first_path_tess_location", "grad_simple_ramps", "clear_color", "dither", "guard_pieces")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class SceneInfo(ctypes.Structure):
    _fields_ = [("node_count", ctypes.c_uint32), ("path_count", ctypes.c_uint32),
                ("cubic_count", ctypes.c_uint64), ("duration", ctypes.c_float), ("extent", ctypes.c_float)]


class SceneKey(ctypes.Structure):
    _fields_ = [("time", ctypes.c_float), ("value", ctypes.c_float * 5), ("cubic", ctypes.c_uint32),
                ("ease", ctypes.c_float * 4)]


class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint32), ("time", ctypes.c_float), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("arrival_s", ctypes.c_double)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_uint32)]


class RequestResult(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("batch", ctypes.c_uint32), ("batch_size", ctypes.c_uint32), ("batch_bytes", ctypes.c_uint64)]


class JpegComponentInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("h", ctypes.c_uint32), ("v", ctypes.c_uint32),
                ("quant_table", ctypes.c_uint32), ("scale_h", ctypes.c_uint32), ("scale_v", ctypes.c_uint32),
                ("rep_h", ctypes.c_uint32), ("rep_v", ctypes.c_uint32), ("blocks_w", ctypes.c_uint32),
                ("blocks_h", ctypes.c_uint32), ("coef_offset", ctypes.c_uint64)]


class JpegInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("component_count", ctypes.c_uint32),
                ("ycc", ctypes.c_uint32), ("hmax", ctypes.c_uint32), ("vmax", ctypes.c_uint32),
                ("coef_count", ctypes.c_uint64), ("components", JpegComponentInfo * 4),
                ("quant", (ctypes.c_uint16 * 64) * 4)]


class RasterConfig(ctypes.Structure):
    _fields_ = [("tile_order", ctypes.c_uint32), ("persistent", ctypes.c_uint32),
                ("waves_per_simd", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_ if n != "reserved"}


EXPORTED = [
    "rive_hip_set_interior_triangulation", "rive_hip_get_interior_triangulation",
    "rive_hip_last_error", "rive_hip_create", "rive_hip_destroy", "rive_hip_buffer_resize",
    "rive_hip_buffer_map", "rive_hip_buffer_unmap", "rive_hip_target_create",
    "rive_hip_target_destroy", "rive_hip_flush", "rive_hip_sync", "rive_hip_read_pixels",
    "rive_hip_debug_readback", "rive_hip_pass_times", "rive_stream_generate", "rive_stream_free",
    "rive_hip_render_stream", "rive_hip_last_flush_stats", "rive_hip_image_create",
    "rive_hip_image_destroy", "rive_hip_render_streams", "rive_hip_last_flush_canvases",
    "rive_stream_generate_batch", "rive_stream_contour_bound", "rive_hip_flush_noalloc",
    "rive_hip_count_draws", "rive_hip_render_stream_ex", "rive_hip_last_frame", "rive_hip_flush_frame",
    "rive_scene_generate", "rive_scene_free", "rive_scene_info", "rive_scene_node",
    "rive_scene_world_transforms", "rive_scene_stream", "rive_ease_cubic", "rive_scene_request_bytes",
    "rive_hip_serve", "rive_cmd_record_stream", "rive_cmd_free", "rive_cmd_bytes", "rive_cmd_from_bytes",
    "rive_hip_render_commands", "rive_hip_frame_graph_create", "rive_hip_frame_graph_launch",
    "rive_hip_frame_graph_destroy", "rive_hip_set_raster_config", "rive_hip_get_raster_config",
    "rive_hip_autotune_frame", "rive_image_probe", "rive_png_decode", "rive_jpeg_decode_info",
    "rive_hip_image_decode", "rive_hip_image_read",
]

_lib = None


def lib():
    """Loads librive_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    L.rive_hip_last_error.restype = ctypes.c_char_p
    L.rive_hip_create.argtypes = [i32, ctypes.POINTER(vp)]
    L.rive_hip_destroy.argtypes = [vp]
    L.rive_hip_destroy.restype = None
    L.rive_hip_buffer_resize.argtypes = [vp, i32, ctypes.c_size_t]
    L.rive_hip_buffer_map.argtypes = [vp, i32, ctypes.c_size_t]
    L.rive_hip_buffer_map.restype = vp
    L.rive_hip_buffer_unmap.argtypes = [vp, i32, ctypes.c_size_t]
    L.rive_hip_target_create.argtypes = [vp, u32, u32, ctypes.POINTER(vp)]
    L.rive_hip_target_destroy.argtypes = [vp]
    L.rive_hip_target_destroy.restype = None
    L.rive_hip_flush.argtypes = [vp, ctypes.POINTER(FlushDesc)]
    L.rive_hip_flush_noalloc.argtypes = [vp, ctypes.POINTER(FlushDesc)]
    L.rive_hip_sync.argtypes = [vp]
    L.rive_hip_read_pixels.argtypes = [vp, vp, vp]
    L.rive_hip_debug_readback.argtypes = [vp, i32, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.rive_hip_pass_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), i32]
    L.rive_stream_generate.argtypes = [i32, ctypes.c_uint64, u32, u32, u32,
                                       ctypes.POINTER(ctypes.POINTER(Stream))]
    L.rive_stream_free.argtypes = [ctypes.POINTER(Stream)]
    L.rive_stream_free.restype = None
    L.rive_hip_render_stream.argtypes = [vp, ctypes.POINTER(Stream), vp]
    L.rive_hip_render_stream_ex.argtypes = [vp, ctypes.POINTER(Stream), vp, u32, ctypes.POINTER(u32)]
    L.rive_hip_count_draws.argtypes = [vp, u32, vp, vp]
    L.rive_hip_last_frame.argtypes = [vp, ctypes.POINTER(FlushDesc), u32, ctypes.POINTER(u32)]
    L.rive_hip_flush_frame.argtypes = [vp, ctypes.POINTER(FlushDesc), u32]
    L.rive_hip_render_streams.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(Stream)),
                                          ctypes.POINTER(vp), u32]
    L.rive_hip_last_flush_canvases.argtypes = [vp, ctypes.POINTER(Canvas), u32, ctypes.POINTER(u32)]
    L.rive_stream_generate_batch.argtypes = [ctypes.c_uint64, u32, ctypes.POINTER(ctypes.POINTER(Stream))]
    L.rive_stream_contour_bound.argtypes = [ctypes.POINTER(Stream)]
    L.rive_stream_contour_bound.restype = u32
    L.rive_hip_last_flush_stats.argtypes = [vp, ctypes.POINTER(FlushStats)]
    L.rive_image_probe.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                   ctypes.POINTER(u32)]
    L.rive_png_decode.argtypes = [vp, ctypes.c_size_t, vp, ctypes.c_size_t]
    L.rive_jpeg_decode_info.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(JpegInfo), vp, ctypes.c_size_t]
    L.rive_hip_image_decode.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    L.rive_hip_image_read.argtypes = [vp, u32, vp, ctypes.c_size_t, ctypes.POINTER(u32), ctypes.POINTER(u32)]
    L.rive_hip_set_raster_config.argtypes = [vp, ctypes.POINTER(RasterConfig)]
    L.rive_hip_get_raster_config.argtypes = [vp, ctypes.POINTER(RasterConfig)]
    L.rive_hip_set_interior_triangulation.argtypes = [vp, ctypes.c_uint32]
    L.rive_hip_get_interior_triangulation.argtypes = [vp]
    L.rive_hip_get_interior_triangulation.restype = ctypes.c_uint32
    L.rive_hip_autotune_frame.argtypes = [vp, ctypes.POINTER(FlushDesc), u32, u32, ctypes.POINTER(RasterConfig),
                                          ctypes.POINTER(ctypes.c_float)]
    L.rive_hip_image_create.argtypes = [vp, u32, u32, u32, vp, ctypes.POINTER(vp)]
    L.rive_hip_image_destroy.argtypes = [vp]
    L.rive_hip_image_destroy.restype = None
    L.rive_scene_generate.argtypes = [ctypes.c_uint64, ctypes.POINTER(vp)]
    L.rive_scene_free.argtypes = [vp]
    L.rive_scene_free.restype = None
    L.rive_scene_info.argtypes = [vp, ctypes.POINTER(SceneInfo)]
    L.rive_scene_node.argtypes = [vp, u32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(SceneKey), u32,
                                  ctypes.POINTER(u32)]
    L.rive_scene_world_transforms.argtypes = [vp, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
    L.rive_scene_stream.argtypes = [vp, ctypes.c_float, u32, u32, ctypes.POINTER(ctypes.POINTER(Stream))]
    L.rive_ease_cubic.argtypes = [ctypes.c_float] * 5
    L.rive_ease_cubic.restype = ctypes.c_float
    L.rive_scene_request_bytes.argtypes = [vp, u32, u32]
    L.rive_scene_request_bytes.restype = ctypes.c_uint64
    L.rive_hip_serve.argtypes = [vp, ctypes.POINTER(vp), u32, ctypes.POINTER(Request), u32,
                                 ctypes.POINTER(ServeConfig), ctypes.POINTER(RequestResult),
                                 ctypes.POINTER(ctypes.c_void_p)]
    L.rive_cmd_record_stream.argtypes = [ctypes.POINTER(Stream), ctypes.POINTER(vp)]
    L.rive_cmd_free.argtypes = [vp]
    L.rive_cmd_free.restype = None
    L.rive_cmd_bytes.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(vp),
                                 ctypes.POINTER(ctypes.c_uint64)]
    L.rive_cmd_from_bytes.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint64, u32, u32, u32, u32,
                                      ctypes.POINTER(vp)]
    L.rive_hip_render_commands.argtypes = [vp, vp, vp]
    L.rive_hip_frame_graph_create.argtypes = [vp, ctypes.POINTER(FlushDesc), u32, ctypes.POINTER(vp)]
    L.rive_hip_frame_graph_launch.argtypes = [vp]
    L.rive_hip_frame_graph_destroy.argtypes = [vp]
    L.rive_hip_frame_graph_destroy.restype = None
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise RuntimeError(f"rive_hip error {rc}: {lib().rive_hip_last_error().decode()}")


CONFIG_SEEDS = {1: 0x5249564500000001, 2: 0x5249564500000002, 3: 0x5249564500000003, 4: 0x5249564500000004}


def generate_stream(config, seed=None, width=0, height=0, draw_count=0):
    """Seeded synthetic draw stream (see rive_stream_generate)."""
    L = lib()
    p = ctypes.POINTER(Stream)()
    check(L.rive_stream_generate(config, CONFIG_SEEDS[config] if seed is None else seed,
                                 width, height, draw_count, ctypes.byref(p)))
    return p


def free_stream(p):
    lib().rive_stream_free(p)


def contour_bound(stream):
    """Upper bound of the contours a stream produces (rive_stream_contour_bound)."""
    return int(lib().rive_stream_contour_bound(stream))


IMAGE_FORMATS = {0: "unknown", 1: "png", 2: "jpeg", 3: "webp"}


def image_probe(data):
    """(format name, width, height) of an encoded image (rive_image_probe)."""
    buf = bytes(data)
    f, w, h = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    check(lib().rive_image_probe(buf, len(buf), ctypes.byref(f), ctypes.byref(w), ctypes.byref(h)))
    return IMAGE_FORMATS[f.value], w.value, h.value


def png_decode(data):
    """A PNG decoded on the host: unpremultiplied RGBA8, shape (h, w, 4)."""
    buf = bytes(data)
    _, w, h = image_probe(buf)
    out = np.empty((h, w, 4), np.uint8)
    check(lib().rive_png_decode(buf, len(buf), out.ctypes.data, out.nbytes))
    return out


def jpeg_decode_info(data, coefficients=True):
    """(JpegInfo, int16 coefficients or None) of a JPEG after entropy decoding."""
    buf = bytes(data)
    info = JpegInfo()
    check(lib().rive_jpeg_decode_info(buf, len(buf), ctypes.byref(info), None, 0))
    if not coefficients:
        return info, None
    coef = np.empty(max(int(info.coef_count), 1), np.int16)
    check(lib().rive_jpeg_decode_info(buf, len(buf), ctypes.byref(info), coef.ctypes.data, coef.size))
    return info, coef[:int(info.coef_count)]


def flush_groups(streams, limit=65535):
    """Consecutive runs of streams whose contour bounds sum to at most `limit`
    (the canvases rive_hip_render_streams puts in one flush)."""
    groups, cur, total = [], [], 0
    for i, s in enumerate(streams):
        b = contour_bound(s)
        if cur and total + b > limit:
            groups.append(cur)
            cur, total = [], 0
        cur.append(i)
        total += b
    if cur:
        groups.append(cur)
    return groups


def generate_batch(n=256, seed=None):
    """The batched workload's canvases (see rive_stream_generate_batch)."""
    L = lib()
    arr = (ctypes.POINTER(Stream) * n)()
    check(L.rive_stream_generate_batch(CONFIG_SEEDS[4] if seed is None else seed, n, arr))
    return list(arr)


class Scene:
    """An animated scene (rive_scene_generate): a transform tree with keyframed
    affine animation and cubic paths on its nodes."""

    def __init__(self, seed):
        self.L = lib()
        self.h = ctypes.c_void_p()
        check(self.L.rive_scene_generate(seed, ctypes.byref(self.h)))
        info = SceneInfo()
        check(self.L.rive_scene_info(self.h, ctypes.byref(info)))
        self.node_count, self.path_count = int(info.node_count), int(info.path_count)
        self.cubic_count, self.duration, self.extent = int(info.cubic_count), float(info.duration), float(info.extent)

    def close(self):
        if self.h:
            self.L.rive_scene_free(self.h)
            self.h = ctypes.c_void_p()

    def node(self, i):
        """(parent, [SceneKey...]) of node i."""
        parent = ctypes.c_int32()
        n = ctypes.c_uint32()
        check(self.L.rive_scene_node(self.h, i, ctypes.byref(parent), None, 0, ctypes.byref(n)))
        keys = (SceneKey * max(n.value, 1))()
        check(self.L.rive_scene_node(self.h, i, ctypes.byref(parent), keys, n.value, ctypes.byref(n)))
        return int(parent.value), [keys[k] for k in range(n.value)]

    def world_transforms(self, t):
        out = (ctypes.c_float * (6 * self.node_count))()
        check(self.L.rive_scene_world_transforms(self.h, t, out))
        return list(out)

    def stream(self, t, width, height):
        """The scene at time t as a draw stream (free with free_stream)."""
        p = ctypes.POINTER(Stream)()
        check(self.L.rive_scene_stream(self.h, t, width, height, ctypes.byref(p)))
        return p

    def request_bytes(self, width, height):
        return int(self.L.rive_scene_request_bytes(self.h, width, height))


class CommandBuffer:
    """A deferred 2D command stream (rive_cmd_*).  Files written by save():
    b"RIVECMD1", then u32 width, height, clear colour, dither, u64 command
    bytes, u64 blob bytes, then the command and blob bytes."""

    MAGIC = b"RIVECMD1"

    def __init__(self, handle, width, height, clear_color, dither):
        self.L = lib()
        self.h = handle
        self.width, self.height, self.clear_color, self.dither = width, height, clear_color, dither

    @classmethod
    def record(cls, stream):
        h = ctypes.c_void_p()
        check(lib().rive_cmd_record_stream(stream, ctypes.byref(h)))
        st = stream.contents
        return cls(h, st.width, st.height, st.clear_color, st.dither)

    @classmethod
    def from_bytes(cls, commands, blobs, width, height, clear_color=0, dither=1):
        h = ctypes.c_void_p()
        cb = ctypes.create_string_buffer(bytes(commands), max(len(commands), 1))
        bb = ctypes.create_string_buffer(bytes(blobs), max(len(blobs), 1))
        check(lib().rive_cmd_from_bytes(cb, len(commands), bb, len(blobs), width, height, clear_color, dither,
                                        ctypes.byref(h)))
        return cls(h, width, height, clear_color, dither)

    def bytes(self):
        """(command bytes, blob bytes)."""
        c, b = ctypes.c_void_p(), ctypes.c_void_p()
        nc, nb = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.L.rive_cmd_bytes(self.h, ctypes.byref(c), ctypes.byref(nc), ctypes.byref(b), ctypes.byref(nb)))
        return (ctypes.string_at(c, nc.value) if nc.value else b""), (ctypes.string_at(b, nb.value) if nb.value else b"")

    def save(self, path):
        import struct
        c, b = self.bytes()
        with open(path, "wb") as f:
            f.write(self.MAGIC + struct.pack("<4I2Q", self.width, self.height, self.clear_color, self.dither,
                                             len(c), len(b)) + c + b)

    @classmethod
    def load(cls, path):
        import struct
        data = open(path, "rb").read()
        if data[:8] != cls.MAGIC:
            raise ValueError("not a command-stream file")
        w, h, clear, dither, nc, nb = struct.unpack_from("<4I2Q", data, 8)
        off = 8 + struct.calcsize("<4I2Q")
        return cls.from_bytes(data[off:off + nc], data[off + nc:off + nc + nb], w, h, clear, dither)

    def close(self):
        if self.h:
            self.L.rive_cmd_free(self.h)
            self.h = ctypes.c_void_p()


def ease_cubic(x1, y1, x2, y2, f):
    return float(lib().rive_ease_cubic(x1, y1, x2, y2, f))


class Context:
    """One backend context (RenderContextImpl) on one HIP device."""

    def __init__(self, device=0):
        self.L = lib()
        self.h = ctypes.c_void_p()
        check(self.L.rive_hip_create(device, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            self.L.rive_hip_destroy(self.h)
            self.h = ctypes.c_void_p()

    def target(self, w, h):
        t = ctypes.c_void_p()
        check(self.L.rive_hip_target_create(self.h, w, h, ctypes.byref(t)))
        return t

    def free_target(self, t):
        self.L.rive_hip_target_destroy(t)

    def upload(self, kind, array):
        """resize + map + memcpy + unmap of one record buffer."""
        k = BUFFER[kind]
        data = np.ascontiguousarray(array).view(np.uint8).ravel()
        n = max(int(data.size), 16)
        check(self.L.rive_hip_buffer_resize(self.h, k, n))
        ptr = self.L.rive_hip_buffer_map(self.h, k, n)
        if not ptr:
            raise RuntimeError(self.L.rive_hip_last_error().decode())
        if data.size:
            ctypes.memmove(ptr, data.ctypes.data, data.size)
        check(self.L.rive_hip_buffer_unmap(self.h, k, int(data.size)))

    def flush(self, desc):
        check(self.L.rive_hip_flush(self.h, ctypes.byref(desc)))

    def sync(self):
        check(self.L.rive_hip_sync(self.h))

    def read_pixels(self, target, w, h):
        out = np.empty((h, w, 4), np.uint8)
        check(self.L.rive_hip_read_pixels(self.h, target, out.ctypes.data))
        return out

    def readback(self, what):
        size = ctypes.c_size_t()
        check(self.L.rive_hip_debug_readback(self.h, what, None, 0, ctypes.byref(size)))
        buf = np.empty(size.value, np.uint8)
        check(self.L.rive_hip_debug_readback(self.h, what, buf.ctypes.data, buf.size,
                                             ctypes.byref(size)))
        return buf

    def pass_times(self):
        ms = (ctypes.c_float * 6)()
        check(self.L.rive_hip_pass_times(self.h, ms, 6))
        return list(ms)

    def render_stream(self, stream, target, max_flushes=0):
        """Replays a stream through the front end and flushes it; the frame
        may split into several logical flushes.  max_flushes > 0 stops after
        that many.  Returns the frame's logical flush count."""
        n = ctypes.c_uint32()
        check(self.L.rive_hip_render_stream_ex(self.h, stream, target, max_flushes, ctypes.byref(n)))
        return n.value

    def render_streams(self, streams, targets):
        """One batched flush: stream i renders into targets[i]."""
        n = len(streams)
        sa = (ctypes.POINTER(Stream) * n)(*streams)
        ta = (ctypes.c_void_p * n)(*[t.value if isinstance(t, ctypes.c_void_p) else t for t in targets])
        check(self.L.rive_hip_render_streams(self.h, sa, ta, n))

    def last_flush_canvases(self):
        count = ctypes.c_uint32()
        check(self.L.rive_hip_last_flush_canvases(self.h, None, 0, ctypes.byref(count)))
        arr = (Canvas * max(count.value, 1))()
        check(self.L.rive_hip_last_flush_canvases(self.h, arr, count.value, ctypes.byref(count)))
        return arr, count.value

    def last_frame(self):
        """The last frame's logical flush descriptors (FlushDesc list)."""
        n = ctypes.c_uint32()
        check(self.L.rive_hip_last_frame(self.h, None, 0, ctypes.byref(n)))
        arr = (FlushDesc * max(n.value, 1))()
        check(self.L.rive_hip_last_frame(self.h, arr, n.value, ctypes.byref(n)))
        return [arr[k] for k in range(n.value)]

    def flush_frame(self, descs, target=None):
        """rive_hip_flush_frame over a frame's logical flush descriptors."""
        arr = (FlushDesc * len(descs))(*descs)
        if target is not None:
            for k in range(len(descs)):
                arr[k].target = target
        check(self.L.rive_hip_flush_frame(self.h, arr, len(descs)))

    def reflush_frame(self, descs, target=None):
        """Re-issues a frame's logical flushes from the records resident in
        HBM (optionally into another target); returns each flush's stats."""
        out = []
        for d in descs:
            if target is not None:
                d.target = target
            self.flush(d)
            out.append(self.last_flush_stats())
        return out

    def image_decode(self, data):
        """An encoded image decoded into a device image with its mip chain."""
        buf = bytes(data)
        h = ctypes.c_void_p()
        check(self.L.rive_hip_image_decode(self.h, buf, len(buf), ctypes.byref(h)))
        return h

    def free_image(self, image):
        self.L.rive_hip_image_destroy(image)

    def image_read(self, image, level=0):
        """One mip level of a device image, premultiplied RGBA8 (h, w, 4)."""
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        check(self.L.rive_hip_image_read(image, level, None, 0, ctypes.byref(w), ctypes.byref(h)))
        out = np.empty((h.value, w.value, 4), np.uint8)
        check(self.L.rive_hip_image_read(image, level, out.ctypes.data, out.nbytes, ctypes.byref(w),
                                         ctypes.byref(h)))
        return out

    def set_interior_triangulation(self, enabled):
        """Whether the front end lays large plain fills out as outer-cubic
        patches (interior triangulation); off by default."""
        check(self.L.rive_hip_set_interior_triangulation(self.h, 1 if enabled else 0))

    def interior_triangulation(self):
        return bool(self.L.rive_hip_get_interior_triangulation(self.h))

    def set_raster_config(self, tile_order=0, persistent=0, waves_per_simd=6):
        """How the raster pass runs (every choice renders the same pixels)."""
        cfg = RasterConfig(tile_order, persistent, waves_per_simd, 0)
        check(self.L.rive_hip_set_raster_config(self.h, ctypes.byref(cfg)))

    def autotune_frame(self, descs, trials=5):
        """Times every raster config on a frame's logical flushes and keeps the
        fastest; returns (chosen config dict, {config tuple: median ms})."""
        arr = (FlushDesc * len(descs))(*descs)
        cfg = RasterConfig()
        times = (ctypes.c_float * 8)()
        check(self.L.rive_hip_autotune_frame(self.h, arr, len(descs), trials, ctypes.byref(cfg), times))
        keys = [(o, p, w) for o in (0, 1) for p in (0, 1) for w in (6, 5)]
        return cfg.as_dict(), {k: float(times[i]) for i, k in enumerate(keys)}

    def raster_config(self):
        cfg = RasterConfig()
        check(self.L.rive_hip_get_raster_config(self.h, ctypes.byref(cfg)))
        return cfg.as_dict()

    def last_flush_stats(self):
        st = FlushStats()
        check(self.L.rive_hip_last_flush_stats(self.h, ctypes.byref(st)))
        return st

    def reflush_geometry(self, target, stats, canvases=None):
        """Flushes again from the records already resident in HBM (the last
        GEOMETRY flush's PATH/PAINT/PAINT_AUX/DRAW/VERB/POINT/GRAD_SPAN
        buffers), with no host writes.  canvases: a (Canvas array, count)
        pair from last_flush_canvases() for a batched flush."""
        desc = FlushDesc()
        desc.target = target
        if canvases is not None:
            desc.canvases = ctypes.cast(canvases[0], ctypes.c_void_p)
            desc.canvas_count = canvases[1]
        desc.mode = FLUSH_GEOMETRY
        desc.load_action = stats.load_action
        desc.clear_color = stats.clear_color
        desc.dither = stats.dither
        desc.path_count = stats.path_count
        desc.draw_count = stats.draw_count
        desc.grad_span_count = stats.grad_span_count
        desc.grad_simple_rows = stats.grad_simple_rows
        desc.grad_complex_rows = stats.grad_complex_rows
        desc.first_path_tess_location = stats.first_path_tess_location
        self.flush(desc)

    def frame_graph(self, descs, target=None):
        """rive_hip_frame_graph_create over a frame's descriptors (optionally
        retargeted); returns a handle for frame_graph_launch / _destroy."""
        arr = (FlushDesc * len(descs))(*descs)
        if target is not None:
            for d in arr:
                d.target = target
        g = ctypes.c_void_p()
        check(self.L.rive_hip_frame_graph_create(self.h, arr, len(descs), ctypes.byref(g)))
        return g

    def frame_graph_launch(self, g):
        check(self.L.rive_hip_frame_graph_launch(g))

    def frame_graph_destroy(self, g):
        self.L.rive_hip_frame_graph_destroy(g)

    def render_commands(self, cmd, target):
        """Replays a CommandBuffer into target (rive_hip_render_commands)."""
        check(self.L.rive_hip_render_commands(self.h, cmd.h, target))

    def serve(self, scenes, requests, hbm_budget_bytes, max_batch, pixels=None):
        """rive_hip_serve: requests is a list of (scene index, time, width,
        height, arrival seconds) in arrival order; pixels, when given, maps a
        request index to a (height, width, 4) uint8 array receiving its
        result.  Returns a list of RequestResult."""
        n = len(requests)
        sc = (ctypes.c_void_p * len(scenes))(*[x.h.value for x in scenes])
        rq = (Request * max(n, 1))()
        for i, (si, t, w, h, at) in enumerate(requests):
            rq[i] = Request(si, t, w, h, at)
        cfg = ServeConfig(int(hbm_budget_bytes), int(max_batch))
        res = (RequestResult * max(n, 1))()
        px = None
        if pixels:
            px = (ctypes.c_void_p * max(n, 1))()
            for i, a in pixels.items():
                px[i] = a.ctypes.data
        check(self.L.rive_hip_serve(self.h, sc, len(scenes), rq, n, ctypes.byref(cfg), res, px))
        return [res[i] for i in range(n)]

    def image(self, rgba_premul, mip_levels=0):
        """Uploads a premultiplied RGBA8 image (h, w, 4) and builds its mips."""
        a = np.ascontiguousarray(rgba_premul, dtype=np.uint8)
        h, w = a.shape[:2]
        out = ctypes.c_void_p()
        check(self.L.rive_hip_image_create(self.h, w, h, mip_levels, a.ctypes.data, ctypes.byref(out)))
        return out

    def free_image(self, img):
        self.L.rive_hip_image_destroy(img)
