"""Device parity through the C ABI (include/rive_hip.h) against the oracle.

Integer/record work (spans, contours, tessellation ranges) is compared
bit-for-bit; pixels must agree within 2/255 on all but 0.1% of pixels and
within 8/255 everywhere (the AA-edge bound: fp16 coverage rounding and the
dither add up to a few code values on edge pixels).  Tessellated vertex
positions and tangent angles are float results of acos/sin/cos and agree to
2e-3 px and 2e-3 rad.
"""
import numpy as np
import pytest

import oracle_binding as ob
import rivehip

pytestmark = pytest.mark.gpu

PIXEL_TOL = 2          # per-channel code values, bulk
PIXEL_BAD_FRAC = 1e-3  # fraction of pixels allowed above PIXEL_TOL
PIXEL_MAX = 8          # per-channel code values, anywhere
TEXEL_TOL_PX = 2e-3
MIRRORED_CONTOUR_FLAG = 1 << 23
THETA_TOL_RAD = 2e-3

# (config, draw_count, width, height): small enough for the oracle to raster
# in seconds.
RASTER_CASES = [(1, 20, 256, 256), (1, 120, 512, 512), (2, 60, 480, 270), (2, 200, 640, 360)]
# The layout-only comparisons run at the full workload sizes.
LAYOUT_CASES = RASTER_CASES + [(1, 200, 1024, 1024), (2, 5000, 1920, 1080)]


def case_id(c):
    return f"cfg{c[0]}-n{c[1]}-{c[2]}x{c[3]}"


def announce(tag, c):
    print(f"[{tag}] {case_id(c)}", flush=True)


def pixel_check(out, ref):
    d = np.abs(out.astype(np.int32) - ref.astype(np.int32)).max(axis=2)
    bad = float((d > PIXEL_TOL).mean())
    assert bad <= PIXEL_BAD_FRAC and int(d.max()) <= PIXEL_MAX, (
        f"max diff {int(d.max())}, fraction > {PIXEL_TOL}: {bad:.6f}")


def same_bits(a, b):
    return a.tobytes() == b.tobytes()


@pytest.fixture
def stream_case(request):
    c = request.param
    s = rivehip.generate_stream(c[0], width=c[2], height=c[3], draw_count=c[1])
    f = ob.OracleFlush(s)
    yield c, s, f
    f.close()
    rivehip.free_stream(s)


@pytest.mark.parametrize("stream_case", RASTER_CASES, ids=case_id, indirect=True)
def test_spans_mode_matches_oracle(hip_ctx, stream_case):
    c, s, f = stream_case
    announce("spans", c)
    ref, _ = f.render()
    t = hip_ctx.target(c[2], c[3])
    try:
        f.flush_spans_on(hip_ctx, t)
        out = hip_ctx.read_pixels(t, c[2], c[3])
        tex = hip_ctx.readback(2).view(rivehip.TESS_TEXEL)
    finally:
        hip_ctx.free_target(t)
    pixel_check(out, ref)
    tref = f.tessellate()
    n = min(len(tex), len(tref))
    # The device tessellates each span's first rect; the mirrored copies of
    # double-sided fills are read back through the forward texels instead.
    live = tex["w"][:n] != 0
    forward = (tref["w"][:n] != 0) & ((tref["w"][:n] & MIRRORED_CONTOUR_FLAG) == 0)
    assert not np.any(forward & ~live), "forward texels the device did not write"
    assert np.array_equal(tex["w"][:n][live], tref["w"][:n][live])
    # z holds the tangent angle theta in [0, 2pi) as float bits.  acos near
    # +-1 amplifies one-ulp input differences to ~3e-4 rad.
    za = tex["z"][:n][live].view(np.float32).astype(np.float64)
    zb = tref["z"][:n][live].view(np.float32).astype(np.float64)
    dz = np.abs(za - zb) % (2 * np.pi)
    assert np.minimum(dz, 2 * np.pi - dz).max(initial=0) <= THETA_TOL_RAD
    dx = np.abs(tex["x"][:n][live] - tref["x"][:n][live]).max(initial=0)
    dy = np.abs(tex["y"][:n][live] - tref["y"][:n][live]).max(initial=0)
    assert max(dx, dy) <= TEXEL_TOL_PX


@pytest.mark.parametrize("stream_case", LAYOUT_CASES, ids=case_id, indirect=True)
def test_geometry_layout_bit_exact(hip_ctx, stream_case):
    """Every logical flush of the frame -- the full 5000-draw 1080p stream
    splits into six, as the reference's 2048x2048-texel tessellation limit
    requires -- has the oracle's draw ranges, contours and spans bit for bit
    (the device lays each flush out from its own counts)."""
    c, s, f = stream_case
    announce("geometry-layout", c)
    t = hip_ctx.target(c[2], c[3])
    try:
        for k in range(f.flush_count):
            n = hip_ctx.render_stream(s, t, max_flushes=k + 1)
            assert n == f.flush_count, (n, f.flush_count)
            hip_ctx.sync()
            spans = hip_ctx.readback(0).view(rivehip.TESS_SPAN)
            contours = hip_ctx.readback(1).view(rivehip.CONTOUR)
            draws = hip_ctx.readback(3).view(rivehip.DRAW)
            f.select(k)
            od = f.draws()
            assert len(draws) == len(od), k
            assert np.array_equal(draws["path_id"], od[:, 0])
            assert np.array_equal(draws["tess_location"], od[:, 1])
            assert np.array_equal(draws["tess_vertex_count"], od[:, 2])
            ocont = f.get(1)
            assert len(contours) == len(ocont) and same_bits(contours, ocont)
            ospans = f.get(0)
            assert len(spans) == len(ospans)
            mism = np.nonzero(spans.view(np.uint32).reshape(-1, 16) != ospans.view(np.uint32).reshape(-1, 16))
            assert mism[0].size == 0, f"flush {k}: first mismatching span {mism[0][:4]}, fields {mism[1][:4]}"
    finally:
        hip_ctx.free_target(t)
    if c[1] == 5000:
        assert f.flush_count == 6


@pytest.mark.parametrize("stream_case", RASTER_CASES, ids=case_id, indirect=True)
def test_geometry_mode_pixels(hip_ctx, stream_case):
    c, s, f = stream_case
    announce("geometry-pixels", c)
    ref, _ = f.render()
    t = hip_ctx.target(c[2], c[3])
    try:
        hip_ctx.render_stream(s, t)
        out = hip_ctx.read_pixels(t, c[2], c[3])
    finally:
        hip_ctx.free_target(t)
    pixel_check(out, ref)


def test_full_workload_modes_agree_and_repeat(hip_ctx):
    """1920x1080, 5000 draws: GEOMETRY and SPANS flushes of the same frame give
    identical pixels, and a repeated flush reproduces them bit for bit.  The
    SPANS frame is the reference's model of six logical flushes: every
    flush's records uploaded once into shared buffers, then six rive_hip_flush
    calls with FlushDescriptor-style first_* offsets, the later ones loading
    the target."""
    c = (2, 5000, 1920, 1080)
    announce("full-modes", c)
    s = rivehip.generate_stream(2)
    f = ob.OracleFlush(s)
    t = hip_ctx.target(1920, 1080)
    try:
        hip_ctx.render_stream(s, t)
        a = hip_ctx.read_pixels(t, 1920, 1080)
        hip_ctx.render_stream(s, t)
        b = hip_ctx.read_pixels(t, 1920, 1080)
        # the same six flushes one rive_hip_flush at a time (the frame call
        # counts them in one launch)
        hip_ctx.reflush_frame(hip_ctx.last_frame(), t)
        seq = hip_ctx.read_pixels(t, 1920, 1080)
        f.flush_all_spans_on(hip_ctx, t, clear_color=0, dither=1)
        c2 = hip_ctx.read_pixels(t, 1920, 1080)
    finally:
        hip_ctx.free_target(t)
        f.close()
        rivehip.free_stream(s)
    assert same_bits(a, b)
    assert same_bits(a, seq)
    assert same_bits(a, c2)
    assert a[..., 3].max() > 0
