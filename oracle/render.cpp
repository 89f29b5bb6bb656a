// CPU ORACLE (test infrastructure only): the front end (RiveRenderer clip
// logic, PathDraw records, gradients, one logical flush) and the raster
// (midpoint-fan patch vertex shader, triangle rasterisation, raster-ordered
// pixel-local-storage resolve, color ramps, dither).
//
// Rasterisation is exact on a 1/256-pixel fixed-point grid with a symbolic
// perturbation tie rule (sample point nudged by (+e, +e^2)), so every sample
// lands strictly inside exactly one of two triangles sharing an edge.
#include "oracle.hpp"
#include "oracle_flush.hpp"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstring>

namespace oracle
{
namespace
{
constexpr float PI_ = 3.14159265359f;
constexpr float _2PI_ = 6.28318530718f;
constexpr float AA_RADIUS = .5f;
constexpr uint32_t SOLID_COLOR_PAINT_TYPE = 1, LINEAR_GRADIENT_PAINT_TYPE = 2,
                   RADIAL_GRADIENT_PAINT_TYPE = 3, CLIP_UPDATE_PAINT_TYPE = 0;
constexpr uint32_t PAINT_FLAG_NON_ZERO_FILL = 0x100u, PAINT_FLAG_EVEN_ODD_FILL = 0x200u;
constexpr uint32_t PAINT_FLAG_HAS_CLIP_RECT = 0x400u;
constexpr uint32_t GRAD_SPAN_FLAG_LEFT_BORDER = 0x80000000u;
constexpr uint32_t GRAD_SPAN_FLAG_RIGHT_BORDER = 0x40000000u;
constexpr uint32_t GRAD_SPAN_FLAG_COMPLEX_BORDER = 0x20000000u;
constexpr uint32_t GRAD_SPAN_FLAGS_MASK = 0xe0000000u;

inline uint32_t fbits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
inline float ubits(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// IEEE fp32 -> fp16 -> fp32 with round-to-nearest-even (packHalf2x16 /
// unpackHalf2x16 of the pixel-local-storage coverage plane).
float round_to_f16(float x)
{
    if (!std::isfinite(x))
        return x;
    float ax = fabsf(x);
    if (ax >= 65520.f)
        return copysignf(INFINITY, x);
    if (ax < 6.103515625e-05f)
    {
        // subnormal half: quantum 2^-24
        float q = rintf(ax * 16777216.f) / 16777216.f;
        return copysignf(q, x);
    }
    int e;
    frexpf(ax, &e); // ax = m * 2^e, m in [0.5,1)
    float quantum = ldexpf(1.f, e - 11);
    return copysignf(rintf(ax / quantum) * quantum, x);
}

// ---------------- Mat2D helpers (include/rive/math/mat2d.hpp) -------------
Mat mul(const Mat& a, const Mat& b) // a * b: apply b first
{
    Mat r;
    r.m[0] = a.m[0] * b.m[0] + a.m[2] * b.m[1];
    r.m[1] = a.m[1] * b.m[0] + a.m[3] * b.m[1];
    r.m[2] = a.m[0] * b.m[2] + a.m[2] * b.m[3];
    r.m[3] = a.m[1] * b.m[2] + a.m[3] * b.m[3];
    r.m[4] = a.m[0] * b.m[4] + a.m[2] * b.m[5] + a.m[4];
    r.m[5] = a.m[1] * b.m[4] + a.m[3] * b.m[5] + a.m[5];
    return r;
}
// Mat2D::invert (src/math/mat2d.cpp:179-200), in float.
bool invert(const Mat& a, Mat* out)
{
    const float aa = a.m[0], ab = a.m[1], ac = a.m[2], ad = a.m[3], atx = a.m[4], aty = a.m[5];
    float det = aa * ad - ab * ac;
    if (det == 0.0f)
        return false;
    det = 1.0f / det;
    out->m[0] = ad * det;
    out->m[1] = -ab * det;
    out->m[2] = -ac * det;
    out->m[3] = aa * det;
    out->m[4] = (ac * aty - ad * atx) * det;
    out->m[5] = (ab * atx - aa * aty) * det;
    return true;
}
V2 map(const Mat& m, V2 p)
{
    return {m.m[0] * p.x + m.m[2] * p.y + m.m[4], m.m[1] * p.x + m.m[3] * p.y + m.m[5]};
}

struct IAABB
{
    int l, t, r, b;
    bool empty() const { return l >= r || t >= b; }
    IAABB intersect(const IAABB& o) const
    {
        return {std::max(l, o.l), std::max(t, o.t), std::min(r, o.r), std::min(b, o.b)};
    }
};
// Mat2D::mapBoundingBox + AABB::roundOut (include/rive/math/aabb.hpp).
IAABB map_bounds_round_out(const Mat& m, const V2* pts, uint32_t n, float outsetX,
                           float outsetY)
{
    float l = INFINITY, t = INFINITY, r = -INFINITY, b = -INFINITY;
    for (uint32_t i = 0; i < n; ++i)
    {
        V2 p = map(m, pts[i]);
        l = std::min(l, p.x);
        t = std::min(t, p.y);
        r = std::max(r, p.x);
        b = std::max(b, p.y);
    }
    l -= outsetX;
    r += outsetX;
    t -= outsetY;
    b += outsetY;
    return {(int)floorf(l), (int)floorf(t), (int)ceilf(r), (int)ceilf(b)};
}

} // namespace

// RiveRenderer::IsAABB (rive_renderer.cpp:16-55): a move and three lines, any
// further points equal to the first, opposite corners sharing x and y.
bool is_aabb(const uint8_t* verbs, uint32_t verbCount, const V2* pts, uint32_t ptCount, float rect[4])
{
    if (verbCount < 4 || verbs[0] != kMove || verbs[1] != kLine || verbs[2] != kLine || verbs[3] != kLine)
        return false;
    for (uint32_t i = 4; i < ptCount; ++i)
        if (!(pts[i].x == pts[0].x && pts[i].y == pts[0].y))
            return false;
    const float c[4] = {pts[0].x, pts[0].y, pts[2].x, pts[2].y};
    const float o[4] = {pts[1].x, pts[1].y, pts[3].x, pts[3].y};
    const bool zyxw = c[0] == o[2] && c[1] == o[1] && c[2] == o[0] && c[3] == o[3];
    const bool xwzy = c[0] == o[0] && c[1] == o[3] && c[2] == o[2] && c[3] == o[1];
    if (!zyxw && !xwzy)
        return false;
    rect[0] = std::min(c[0], c[2]);
    rect[1] = std::min(c[1], c[3]);
    rect[2] = std::max(c[0], c[2]);
    rect[3] = std::max(c[1], c[3]);
    return true;
}

namespace
{
// transform_rect_to_new_space (rive_renderer.cpp:235-265).
bool transform_rect_to_new_space(float rect[4], const Mat& current, const Mat& newMatrix)
{
    if (memcmp(current.m, newMatrix.m, sizeof(current.m)) == 0)
        return true;
    Mat toNew;
    if (!invert(newMatrix, &toNew))
        return false;
    toNew = mul(toNew, current);
    const float maxSkew = fmaxf(fabsf(toNew.m[1]), fabsf(toNew.m[2]));
    const float maxScale = fmaxf(fabsf(toNew.m[0]), fabsf(toNew.m[3]));
    if (maxSkew > 1.0e-5f && maxScale > 1.0e-5f) // math::EPSILON
        return false;
    const V2 a = map(toNew, {rect[0], rect[1]}), b = map(toNew, {rect[2], rect[3]});
    rect[0] = std::min(a.x, b.x);
    rect[1] = std::min(a.y, b.y);
    rect[2] = std::max(a.x, b.x);
    rect[3] = std::max(a.y, b.y);
    return true;
}

// A render state's clip rect (clipRectImpl, rive_renderer.cpp:267-321) and
// its ClipRectInverseMatrix (gpu.cpp:689-706).
struct ClipRectState
{
    bool has = false;
    float rect[4] = {0, 0, 0, 0};
    Mat matrix;
    void add(const float r[4], const Mat& m)
    {
        if (!has)
        {
            memcpy(rect, r, sizeof(rect));
            matrix = m;
            has = true;
            return;
        }
        rect[0] = std::max(rect[0], r[0]);
        rect[1] = std::max(rect[1], r[1]);
        rect[2] = std::min(rect[2], r[2]);
        rect[3] = std::min(rect[3], r[3]);
    }
    IAABB pixelBounds() const
    {
        const V2 corners[4] = {{rect[0], rect[1]}, {rect[2], rect[1]}, {rect[2], rect[3]}, {rect[0], rect[3]}};
        return map_bounds_round_out(matrix, corners, 4, 0, 0);
    }
    bool inverse(Mat* out) const
    {
        if (!has)
            return false;
        const float w = rect[2] - rect[0], h = rect[3] - rect[1];
        Mat toRect;
        toRect.m[0] = w * .5f;
        toRect.m[1] = 0;
        toRect.m[2] = 0;
        toRect.m[3] = h * .5f;
        toRect.m[4] = (rect[0] + rect[2]) * .5f;
        toRect.m[5] = (rect[1] + rect[3]) * .5f;
        if (w <= 0 || h <= 0 || !invert(mul(matrix, toRect), out))
            *out = Mat{{0, 0, 0, 0, 0, 0}};
        return true;
    }
};

// find_clip_rect_coverage_distances (common.glsl:376-400) at the pixel
// center, then min_component of the half4 (draw_raster_order_path.frag:152-159).
float clip_rect_coverage(const float m[6], int x, int y)
{
    const float awx = fabsf(m[0]) + fabsf(m[2]), awy = fabsf(m[1]) + fabsf(m[3]);
    float d;
    if (awx != 0 && awy != 0)
    {
        const float rx = 1.f / awx, ry = 1.f / awy;
        const float fx = float(x) + .5f, fy = float(y) + .5f;
        const float cx = m[0] * fx + m[2] * fy + m[4], cy = m[1] * fx + m[3] * fy + m[5];
        d = std::min(std::min(cx * rx + rx + .5f, cy * ry + ry + .5f),
                     std::min(-cx * rx + rx + .5f, -cy * ry + ry + .5f));
    }
    else
        d = std::min(m[4], m[5]);
    return round_to_f16(d);
}
} // namespace

// ======================= front end ========================================

namespace
{
struct Gradient
{
    uint32_t paintType; // LINEAR/RADIAL
    std::vector<uint32_t> colors;
    std::vector<float> stops;
    float coeffs[3];
};

// Gradient::MakeLinear / MakeRadial (renderer/src/gradient.cpp:38-175).
bool make_gradient(const rive_stream_draw& d, const rive_stream& s, Gradient* g)
{
    size_t count = d.stop_count;
    if (count == 0)
        return false;
    const uint32_t* colors = s.stop_colors + d.first_stop;
    const float* stops = s.stop_offsets + d.first_stop;
    for (size_t i = 0; i < count; ++i)
    {
        if (!(stops[i] >= 0 && stops[i] <= 1))
            return false;
        if (i > 0 && !(stops[i - 1] <= stops[i]))
            return false;
    }
    g->colors.assign(colors, colors + count);
    g->stops.assign(stops, stops + count);
    if (d.paint_type == 1)
    {
        g->paintType = LINEAR_GRADIENT_PAINT_TYPE;
        float sx = d.grad[0], sy = d.grad[1], ex = d.grad[2], ey = d.grad[3];
        float firstStop = stops[0], lastStop = stops[count - 1];
        if ((firstStop != 0 || lastStop != 1) && lastStop - firstStop > (1.f / (1 << 12)))
        {
            // precise_mix(start, end, t) = start*(1-t) + end*t
            float nsx = sx * (1.f - firstStop) + ex * firstStop;
            float nsy = sy * (1.f - firstStop) + ey * firstStop;
            float nex = sx * (1.f - lastStop) + ex * lastStop;
            float ney = sy * (1.f - lastStop) + ey * lastStop;
            sx = nsx, sy = nsy, ex = nex, ey = ney;
            g->stops[0] = 0;
            g->stops[count - 1] = 1;
            if (count > 2)
            {
                float m = 1.f / (lastStop - firstStop);
                float a = -firstStop * m;
                for (size_t i = 1; i < count - 1; ++i)
                    g->stops[i] = stops[i] * m + a;
                for (size_t i = 1; i < count - 1; ++i)
                    g->stops[i] = fmaxf(g->stops[i - 1], g->stops[i]);
                for (size_t i = count - 2; i != 0; --i)
                    g->stops[i] = fminf(g->stops[i], g->stops[i + 1]);
            }
        }
        float vx = ex - sx, vy = ey - sy;
        float inv = 1.f / (vx * vx + vy * vy);
        vx *= inv;
        vy *= inv;
        g->coeffs[0] = vx;
        g->coeffs[1] = vy;
        g->coeffs[2] = -(vx * sx + vy * sy);
    }
    else
    {
        g->paintType = RADIAL_GRADIENT_PAINT_TYPE;
        float cx = d.grad[0], cy = d.grad[1], radius = d.grad[2];
        float lastStop = stops[count - 1];
        if (lastStop != 1 && lastStop > (1.f / (1 << 12)))
        {
            g->stops[count - 1] = 1;
            radius *= lastStop;
            float inverseLastStop = 1.f / lastStop;
            for (size_t i = 0; i < count - 1; ++i)
                g->stops[i] = stops[i] * inverseLastStop;
            if (count > 1)
            {
                g->stops[0] = fmaxf(0, g->stops[0]);
                for (size_t i = 1; i < count - 1; ++i)
                    g->stops[i] = fmaxf(g->stops[i - 1], g->stops[i]);
                for (size_t i = count - 2; i != (size_t)-1; --i)
                    g->stops[i] = fminf(g->stops[i], g->stops[i + 1]);
            }
        }
        g->coeffs[0] = cx;
        g->coeffs[1] = cy;
        g->coeffs[2] = radius;
    }
    return true;
}

struct ClipElement
{
    uint32_t path;
    Mat matrix;
    FillRule fillRule;
    IAABB pixelBounds;
    uint32_t clipID;
    uint32_t flushIndex; // the logical flush clipID belongs to
};

uint32_t swizzle_rgba(uint32_t argb) // SwizzleRiveColorToRGBA
{
    uint32_t a = argb >> 24, r = (argb >> 16) & 0xff, g = (argb >> 8) & 0xff, b = argb & 0xff;
    return r | (g << 8) | (b << 16) | (a << 24);
}
} // namespace

// Resource limits of one logical flush (render_context.cpp:43-53,517-541,
// 484-492, 617-654; gpu.hpp:85-103): path and clip IDs up to
// MaxPathID(1) - 1, contour IDs up to 65535, tessellation vertices up to a
// 2048 x 2048 texture less 25 padding vertices, gradient texture rows up to
// 2048.
constexpr uint64_t kMaxPathID = (0x1fu << 10) - 1 - 1023 - 1;
constexpr uint64_t kMaxContourID = 65535;
constexpr uint64_t kMaxTessVerticesBeforePadding = 2048ull * 2048ull - (8 + 16 + 1);
constexpr uint64_t kMaxGradientRows = 2048;

struct Builder
{
    const rive_stream& s;
    std::vector<Flush>& flushes;
    Flush* fp = nullptr;
    std::shared_ptr<std::vector<OracleImage>> images; // shared by the frame's flushes
    std::vector<ClipElement> clipStack;
    uint32_t clipContentID = 0;
    uint32_t nextClipID = 1;
    uint32_t nextContourID = 1;
    uint32_t tessLocation = kMidpointFanPatchSegmentSpan; // pre-padding
    // LogicalFlush::m_resourceCounts of the current logical flush
    uint64_t resPaths = 0, resContours = 0, resTess = 0;
    // gradient bookkeeping (render_context.cpp:588-674)
    std::vector<std::pair<uint32_t, uint32_t>> simpleRamps;
    std::vector<Gradient> complexRamps;
    struct PendingPaint
    {
        uint32_t pathID;
        bool complex;
        uint32_t row, col;
    };
    std::vector<PendingPaint> pendingGradPaints;
    // Feather-atlas draws of the current logical flush: their tessellation
    // is written after the main draws' (render_context.cpp:2255-2345).
    struct PendingAtlas
    {
        PathDrawDesc d;
        std::vector<uint8_t> verbs;
        std::vector<V2> pts;
        size_t drawIndex;
        bool stroke, scissored;
    };
    std::vector<PendingAtlas> pendingAtlas;

    Builder(const rive_stream& s_, std::vector<Flush>& out) : s(s_), flushes(out) {}
    Flush& cur() { return *fp; }

    const V2* pathPts(uint32_t path) const
    {
        return reinterpret_cast<const V2*>(s.points) + s.paths[path].first_point;
    }

    // Pushes one PathDraw: records + tessellation (LogicalFlush::pushDraws).
    void pushPathDraw(uint32_t pathIdx, const Mat& matrix, float strokeRadius, Join join,
                      Cap cap, FillRule fillRule, uint32_t paintType, uint32_t paintValue,
                      uint32_t clipID, uint32_t outerClipID, const Gradient* grad,
                      uint32_t blendMode = BLEND_SRC_OVER, const Mat* clipRectInverse = nullptr)
    {
        const rive_stream_path& sp = s.paths[pathIdx];
        pushGeometryDraw(s.verbs + sp.first_verb, sp.verb_count, pathPts(pathIdx), sp.point_count, matrix,
                         strokeRadius, join, cap, fillRule, paintType, paintValue, clipID, outerClipID, grad,
                         blendMode, 0, 0, clipRectInverse);
    }

    void pushGeometryDraw(const uint8_t* verbs, uint32_t verbCount, const V2* pts, uint32_t ptCount,
                          const Mat& matrix, float strokeRadius, Join join, Cap cap, FillRule fillRule,
                          uint32_t paintType, uint32_t paintValue, uint32_t clipID, uint32_t outerClipID,
                          const Gradient* grad, uint32_t blendMode, uint32_t image, uint32_t sampler,
                          const Mat* clipRectInverse = nullptr, float featherRadius = 0,
                          const IAABB* atlasBounds = nullptr)
    {
        PathDrawDesc d;
        d.verbs = verbs;
        d.verbCount = verbCount;
        d.pts = pts;
        d.ptCount = ptCount;
        d.matrix = matrix;
        d.strokeRadius = strokeRadius;
        d.join = join;
        d.cap = cap;
        d.fillRule = fillRule;
        d.featherRadius = featherRadius;
        derive_directions(&d);
        const bool atlas = atlasBounds != nullptr;
        if (atlas && strokeRadius == 0)
        {
            // Atlas fills are drawn one way (draw.cpp:659-680): forward, or
            // reverse under a left-handed matrix.
            const float det = matrix.m[0] * matrix.m[3] - matrix.m[1] * matrix.m[2];
            d.dirs = det < 0 ? Dirs::reverse : Dirs::forward;
        }

        uint32_t pathID = (uint32_t)cur().pathData.size(); // record 0 is reserved
        d.pathID = pathID;
        PathTess pt = tessellate_path_records(d, nextContourID, tessLocation);

        // PathData (gpu.cpp:859-877)
        rive_hip_path_data pd{};
        memcpy(pd.matrix, matrix.m, sizeof(pd.matrix));
        pd.stroke_radius = strokeRadius;
        pd.feather_radius = featherRadius;
        float atlasScale = 0, atlasTx = 0, atlasTy = 0;
        uint32_t aw = 0, ah = 0;
        IAABB visible{0, 0, 0, 0};
        if (atlas)
        {
            // The atlas region (draw.cpp:1441-1486): the visible part of the
            // pixel bounds at scale 16 / max(r * maxScale, 16), placed with 2
            // texels of padding on every side (where in the atlas it lands
            // does not change the result: the placement is integral).
            visible = atlasBounds->intersect(IAABB{0, 0, (int)s.width, (int)s.height});
            atlasScale = 16.f / std::max(featherRadius * find_max_scale(matrix), 16.f);
            const uint32_t w = (uint32_t)ceilf((float)(visible.r - visible.l) * atlasScale);
            const uint32_t h = (uint32_t)ceilf((float)(visible.b - visible.t) * atlasScale);
            aw = w + 4;
            ah = h + 4;
            atlasTx = 2.f - (float)visible.l * atlasScale;
            atlasTy = 2.f - (float)visible.t * atlasScale;
            pd.atlas_transform[0] = atlasScale;
            pd.atlas_transform[1] = atlasTx;
            pd.atlas_transform[2] = atlasTy;
        }
        pd.z_index = 0;
        cur().pathData.push_back(pd);

        // PaintData (gpu.cpp:879-939)
        rive_hip_paint_data pp{};
        uint32_t params = paintType;
        if (paintType == CLIP_UPDATE_PAINT_TYPE)
        {
            pp.value = clipID << 16;
            params |= outerClipID << 16;
        }
        else
        {
            params |= clipID << 16 | blendMode << 4;
            pp.value = paintValue;
        }
        if (strokeRadius == 0)
        {
            if (fillRule == FillRule::nonZero)
                params |= PAINT_FLAG_NON_ZERO_FILL;
            else if (fillRule == FillRule::evenOdd)
                params |= PAINT_FLAG_EVEN_ODD_FILL;
        }
        rive_hip_paint_aux aux{};
        cur().pathImage.resize(pathID + 1, 0);
        if (image != 0)
        {
            // PAINT_FLAG_HAS_IMAGE; PaintAuxData image branch (gpu.cpp:1006-1034).
            params |= 0x800u;
            const OracleImage& img = (*images)[image - 1];
            Mat inv;
            if (!invert(matrix, &inv))
                inv = Mat();
            memcpy(aux.image_matrix, inv.m, sizeof(aux.image_matrix));
            float dudx = inv.m[0] * (float)img.width, dudy = inv.m[2] * (float)img.height;
            float dvdx = inv.m[1] * (float)img.width, dvdy = inv.m[3] * (float)img.height;
            float maxScaleFactorPow2 = std::max(dudx * dudx + dvdx * dvdx, dudy * dudy + dvdy * dvdy);
            aux.image_lod = log2f(std::max(maxScaleFactorPow2, 1.f)) * .5f + -.5f; // MIP_MAP_LOD_BIAS
            cur().pathImage[pathID] = image | sampler << 16;
        }
        // PaintAuxData clip rect (gpu.cpp:1044-1062): the inverse matrix and
        // inverse fwidth, or WideOpen {0, 0, 0, 0, 1, 1} without a clip rect.
        if (clipRectInverse != nullptr)
        {
            params |= PAINT_FLAG_HAS_CLIP_RECT;
            memcpy(aux.clip_rect_inverse_matrix, clipRectInverse->m, sizeof(aux.clip_rect_inverse_matrix));
            const float* cm = clipRectInverse->m;
            aux.inverse_fwidth[0] = -1.f / (fabsf(cm[0]) + fabsf(cm[1]));
            aux.inverse_fwidth[1] = -1.f / (fabsf(cm[2]) + fabsf(cm[3]));
        }
        else
        {
            const float wideOpen[6] = {0, 0, 0, 0, 1, 1};
            memcpy(aux.clip_rect_inverse_matrix, wideOpen, sizeof(wideOpen));
        }
        pp.params = params;
        cur().paintData.push_back(pp);

        if (grad != nullptr)
        {
            Mat inv;
            if (!invert(matrix, &inv))
                inv = Mat();
            Mat pm;
            if (grad->paintType == LINEAR_GRADIENT_PAINT_TYPE)
            {
                Mat g;
                g.m[0] = grad->coeffs[0];
                g.m[1] = 0;
                g.m[2] = grad->coeffs[1];
                g.m[3] = 0;
                g.m[4] = grad->coeffs[2];
                g.m[5] = 0;
                pm = mul(g, inv);
            }
            else
            {
                float w = 1 / grad->coeffs[2];
                Mat g;
                g.m[0] = w;
                g.m[1] = 0;
                g.m[2] = 0;
                g.m[3] = w;
                g.m[4] = -grad->coeffs[0] * w;
                g.m[5] = -grad->coeffs[1] * w;
                pm = mul(g, inv);
            }
            memcpy(aux.paint_matrix, pm.m, sizeof(aux.paint_matrix));
            allocateGradient(*grad, pathID);
        }
        cur().paintAux.push_back(aux);

        // Tessellation bookkeeping
        DrawInfo di;
        di.pathID = pathID;
        di.tessLocation = tessLocation;
        di.tessVertexCount = pt.midpointFanTessVertexCount;
        di.dirs = d.dirs;
        di.isStroke = strokeRadius > 0;
        di.featheredFill = strokeRadius == 0 && featherRadius > 0;
        di.outerCubics = uses_interior_triangulation(d);
        if (atlas)
        {
            di.atlas = true;
            di.atlasScale = atlasScale;
            di.atlasTx = atlasTx;
            di.atlasTy = atlasTy;
            di.aw = aw;
            di.ah = ah;
            di.blitL = atlasBounds->l;
            di.blitT = atlasBounds->t;
            di.blitR = atlasBounds->r;
            di.blitB = atlasBounds->b;
            cur().draws.push_back(di);
            cur().pathTessVertexCounts.push_back(pt.midpointFanTessVertexCount);
            if (pt.midpointFanTessVertexCount > 0)
            {
                resPaths += 1;
                resContours += pt.contours.size();
                resTess += pt.midpointFanTessVertexCount;
            }
            // Its tessellation comes after the main draws'.
            PendingAtlas pa;
            pa.d = d;
            pa.verbs.assign(verbs, verbs + verbCount);
            pa.pts.assign(pts, pts + ptCount);
            pa.drawIndex = cur().draws.size() - 1;
            pa.stroke = strokeRadius > 0;
            pa.scissored = !(visible.l == atlasBounds->l && visible.t == atlasBounds->t &&
                             visible.r == atlasBounds->r && visible.b == atlasBounds->b);
            pendingAtlas.push_back(std::move(pa));
            return;
        }
        cur().draws.push_back(di);
        for (size_t c = 0; c < pt.contours.size(); ++c)
            cur().contours.push_back(pt.contours[c]);
        nextContourID += (uint32_t)pt.contours.size();
        for (const CubicRecord& r : pt.records)
        {
            cur().records.push_back(r);
            write_spans(r, d.dirs, &cur().spans);
        }
        cur().pathTessVertexCounts.push_back(pt.midpointFanTessVertexCount);
        tessLocation += pt.midpointFanTessVertexCount;
        if (pt.midpointFanTessVertexCount > 0) // draw.cpp:1378-1391
        {
            resPaths += 1;
            resContours += pt.contours.size();
            resTess += pt.midpointFanTessVertexCount;
        }
    }

    void allocateGradient(const Gradient& g, uint32_t pathID)
    {
        size_t n = g.stops.size();
        PendingPaint pp;
        pp.pathID = pathID;
        if (n == 1 || (n == 2 && g.stops[0] == 0 && g.stops[1] == 1))
        {
            std::pair<uint32_t, uint32_t> key{g.colors[0], g.colors[std::min<size_t>(1, n - 1)]};
            uint32_t idx = ~0u;
            for (size_t i = 0; i < simpleRamps.size(); ++i)
                if (simpleRamps[i] == key)
                    idx = (uint32_t)i * 2;
            if (idx == ~0u)
            {
                idx = (uint32_t)simpleRamps.size() * 2;
                simpleRamps.push_back(key);
            }
            pp.complex = false;
            pp.row = idx / kGradTextureWidth;
            pp.col = idx % kGradTextureWidth;
        }
        else
        {
            uint32_t row = ~0u;
            for (size_t i = 0; i < complexRamps.size(); ++i)
                if (complexRamps[i].colors == g.colors && complexRamps[i].stops == g.stops)
                    row = (uint32_t)i;
            if (row == ~0u)
            {
                row = (uint32_t)complexRamps.size();
                complexRamps.push_back(g);
            }
            pp.complex = true;
            pp.row = row;
            pp.col = 0;
        }
        pendingGradPaints.push_back(pp);
    }

    // Writes gradient spans and patches the gradient paint records
    // (render_context.cpp:1462-1533, gpu.cpp:900-914, 959-985).
    void finishGradients()
    {
        const uint32_t simpleRows =
            (uint32_t)((simpleRamps.size() * 2 + kGradTextureWidth - 1) / kGradTextureWidth);
        const uint32_t complexRows = (uint32_t)complexRamps.size();
        cur().gradSimpleRows = simpleRows;
        cur().gradComplexRows = complexRows;
        const uint32_t height = std::max<uint32_t>(simpleRows + complexRows, 1);
        const uint32_t ONE_TEXEL_FIXED = 65536 / kGradTextureWidth;
        for (size_t i = 0; i < simpleRamps.size(); ++i)
        {
            uint32_t y = (uint32_t)(i / (kGradTextureWidth / 2));
            uint32_t centerX = (uint32_t)((i % (kGradTextureWidth / 2)) * 2 + 1);
            uint32_t cx = centerX * ONE_TEXEL_FIXED;
            rive_hip_grad_span gs;
            gs.horizontal_span = (cx << 16) | cx;
            gs.y_with_flags = GRAD_SPAN_FLAG_LEFT_BORDER | GRAD_SPAN_FLAG_RIGHT_BORDER | y;
            gs.color0 = simpleRamps[i].first;
            gs.color1 = simpleRamps[i].second;
            cur().gradSpans.push_back(gs);
        }
        for (uint32_t i = 0; i < complexRows; ++i)
        {
            const Gradient& g = complexRamps[i];
            uint32_t y = i + simpleRows;
            float m = (kGradTextureWidth - 1.f) * ONE_TEXEL_FIXED;
            float a = .5f * ONE_TEXEL_FIXED;
            uint32_t lastXFixed = (uint32_t)(g.stops[0] * m + a);
            uint32_t lastColor = g.colors[0];
            for (size_t k = 1; k < g.stops.size(); ++k)
            {
                uint32_t xFixed = (uint32_t)(g.stops[k] * m + a);
                uint32_t flags = GRAD_SPAN_FLAG_COMPLEX_BORDER;
                if (k == 1)
                    flags |= GRAD_SPAN_FLAG_LEFT_BORDER;
                if (k == g.stops.size() - 1)
                    flags |= GRAD_SPAN_FLAG_RIGHT_BORDER;
                rive_hip_grad_span gs;
                gs.horizontal_span = (xFixed << 16) | lastXFixed;
                gs.y_with_flags = flags | y;
                gs.color0 = lastColor;
                gs.color1 = g.colors[k];
                cur().gradSpans.push_back(gs);
                lastColor = g.colors[k];
                lastXFixed = xFixed;
            }
        }
        const float inverseHeight = 1.f / height;
        for (const PendingPaint& pp : pendingGradPaints)
        {
            uint32_t row = pp.row + (pp.complex ? simpleRows : 0);
            cur().paintData[pp.pathID].value = fbits((static_cast<float>(row) + .5f) * inverseHeight);
            float left, right;
            if (pp.complex)
            {
                left = 0;
                right = (float)kGradTextureWidth;
            }
            else
            {
                left = (float)pp.col;
                right = left + 2;
            }
            cur().paintAux[pp.pathID].grad_span[0] = (right - left - 1) * (1.f / kGradTextureWidth);
            cur().paintAux[pp.pathID].grad_span[1] = (left + .5f) * (1.f / kGradTextureWidth);
        }
        cur().gradHeight = height;
    }

    // The clip elements a draw under `height` clip paths must (re)issue in
    // the current logical flush: those above the deepest element whose
    // content the clip buffer holds (rive_renderer.cpp:631-824).  A clip ID
    // from an earlier logical flush never matches.
    size_t clipInBuffer(size_t height) const
    {
        if (clipContentID != 0)
            for (size_t i = height - 1; i != (size_t)-1; --i)
                if (clipStack[i].clipID == clipContentID && clipStack[i].flushIndex == flushes.size() - 1)
                    return i;
        return (size_t)-1;
    }

    // applyClip: pushes the clip-update draws the stack still needs and
    // returns the clip ID the draw reads.
    uint32_t applyClip(size_t height, const IAABB& frame)
    {
        if (height == 0)
            return 0;
        const size_t inBuffer = clipInBuffer(height);
        uint32_t parentClipID = inBuffer == (size_t)-1 ? 0 : clipStack[inBuffer].clipID;
        for (size_t i = inBuffer + 1; i < height; ++i)
        {
            ClipElement& ce = clipStack[i];
            ce.clipID = nextClipID++;
            ce.flushIndex = (uint32_t)flushes.size() - 1;
            IAABB cb = ce.pixelBounds;
            if (!cb.intersect(frame).empty())
                pushPathDraw(ce.path, ce.matrix, 0, Join::miter, Cap::butt, ce.fillRule,
                             CLIP_UPDATE_PAINT_TYPE, 0, ce.clipID, parentClipID, nullptr);
            parentClipID = ce.clipID;
        }
        clipContentID = parentClipID;
        return parentClipID;
    }

    // Draw::resourceCounts of one PathDraw (draw.cpp:1378-1391): a path, its
    // contours and its tessellation vertices, or nothing when it tessellates
    // to no vertices.
    struct Needs
    {
        uint64_t paths = 0, contours = 0, tess = 0;
        void add(const Needs& o)
        {
            paths += o.paths;
            contours += o.contours;
            tess += o.tess;
        }
    };
    Needs measure(const uint8_t* verbs, uint32_t verbCount, const V2* pts, uint32_t ptCount, const Mat& matrix,
                  float strokeRadius, Join join, Cap cap, FillRule fillRule, float featherRadius = 0) const
    {
        PathDrawDesc d;
        d.verbs = verbs;
        d.verbCount = verbCount;
        d.pts = pts;
        d.ptCount = ptCount;
        d.matrix = matrix;
        d.strokeRadius = strokeRadius;
        d.join = join;
        d.cap = cap;
        d.fillRule = fillRule;
        d.featherRadius = featherRadius;
        derive_directions(&d);
        d.pathID = 1;
        PathTess pt = tessellate_path_records(d, 1, 0);
        Needs n;
        if (pt.midpointFanTessVertexCount > 0)
        {
            n.paths = 1;
            n.contours = pt.contours.size();
            n.tess = pt.midpointFanTessVertexCount;
        }
        return n;
    }
    Needs measurePath(uint32_t pathIdx, const Mat& m, float strokeRadius, Join join, Cap cap, FillRule rule) const
    {
        const rive_stream_path& sp = s.paths[pathIdx];
        return measure(s.verbs + sp.first_verb, sp.verb_count, pathPts(pathIdx), sp.point_count, m, strokeRadius,
                       join, cap, rule);
    }

    // Whether a draw batch -- the clip draws applyClip would issue now plus
    // the draw itself -- fits the current logical flush: the checks of
    // generateClipID (render_context.cpp:484-492), LogicalFlush::pushDraws
    // (:517-541) and allocateGradient (:617-654).
    bool batchFits(size_t height, const IAABB& frame, const Needs& drawNeeds, const Gradient* grad) const
    {
        Needs n = drawNeeds;
        uint64_t newClips = 0;
        if (height > 0)
            for (size_t i = clipInBuffer(height) + 1; i < height; ++i)
            {
                ++newClips;
                const ClipElement& ce = clipStack[i];
                IAABB cb = ce.pixelBounds;
                if (!cb.intersect(frame).empty())
                    n.add(measurePath(ce.path, ce.matrix, 0, Join::miter, Cap::butt, ce.fillRule));
            }
        if ((uint64_t)(nextClipID - 1) + newClips > kMaxPathID)
            return false;
        if (resPaths + n.paths > kMaxPathID || resContours + n.contours > kMaxContourID ||
            resTess + n.tess > kMaxTessVerticesBeforePadding)
            return false;
        if (grad != nullptr)
        {
            uint64_t simple = simpleRamps.size(), complex = complexRamps.size();
            const size_t k = grad->stops.size();
            if (k == 1 || (k == 2 && grad->stops[0] == 0 && grad->stops[1] == 1))
            {
                std::pair<uint32_t, uint32_t> key{grad->colors[0], grad->colors[std::min<size_t>(1, k - 1)]};
                if (std::find(simpleRamps.begin(), simpleRamps.end(), key) == simpleRamps.end())
                    ++simple;
            }
            else
            {
                bool found = false;
                for (const Gradient& g : complexRamps)
                    found |= g.colors == grad->colors && g.stops == grad->stops;
                complex += found ? 0 : 1;
            }
            if ((simple + kGradTextureWidth / 2 - 1) / (kGradTextureWidth / 2) + complex > kMaxGradientRows)
                return false;
        }
        return true;
    }

    void beginFlush()
    {
        flushes.emplace_back();
        fp = &flushes.back();
        Flush& f = cur();
        f.images = images;
        f.width = s.width;
        f.height = s.height;
        f.clearColor = s.clear_color;
        f.dither = s.dither;
        f.loadPreserve = flushes.size() > 1; // later logical flushes load the target
        // Path record 0 is reserved for the clear color (render_context.cpp:1535-1548).
        f.pathData.push_back(rive_hip_path_data{});
        rive_hip_paint_data clearPaint{};
        clearPaint.params = SOLID_COLOR_PAINT_TYPE;
        clearPaint.value = swizzle_rgba(s.clear_color);
        f.paintData.push_back(clearPaint);
        f.paintAux.push_back(rive_hip_paint_aux{});
        clipContentID = 0;
        nextClipID = 1;
        nextContourID = 1;
        tessLocation = kMidpointFanPatchSegmentSpan;
        resPaths = resContours = resTess = 0;
        simpleRamps.clear();
        complexRamps.clear();
        pendingGradPaints.clear();
    }

    // pushFeatherAtlasTessellation for every atlas draw of the flush, fills
    // then strokes, unscissored then scissored (render_context.cpp:2255-2345).
    void writeAtlasTessellation()
    {
        Flush& f = cur();
        for (bool stroked : {false, true})
            for (bool scissored : {false, true})
                for (PendingAtlas& pa : pendingAtlas)
                {
                    if (pa.stroke != stroked || pa.scissored != scissored)
                        continue;
                    pa.d.verbs = pa.verbs.data();
                    pa.d.pts = pa.pts.data();
                    PathTess pt = tessellate_path_records(pa.d, nextContourID, tessLocation);
                    DrawInfo& di = f.draws[pa.drawIndex];
                    di.tessLocation = tessLocation;
                    di.tessVertexCount = pt.midpointFanTessVertexCount;
                    for (const ContourData& c : pt.contours)
                        f.contours.push_back(c);
                    nextContourID += (uint32_t)pt.contours.size();
                    for (const CubicRecord& r : pt.records)
                    {
                        f.records.push_back(r);
                        write_spans(r, pa.d.dirs, &f.spans);
                    }
                    tessLocation += pt.midpointFanTessVertexCount;
                }
        pendingAtlas.clear();
    }

    void endFlush()
    {
        Flush& f = cur();
        writeAtlasTessellation();
        finishGradients();
        f.pathCount = (uint32_t)f.pathData.size() - 1;
        f.pathImage.resize(f.pathCount + 1, 0);
        f.tessEnd = tessLocation;
        // Padding spans (render_context.cpp:1550-1568): 8 at 0 and 1 at the end.
        std::vector<TessVertexSpan> padding;
        write_padding_span(kMidpointFanPatchSegmentSpan, 0, &padding);
        write_padding_span(1, tessLocation, &padding);
        f.spans.insert(f.spans.begin(), padding.begin(), padding.end());
        f.tessVertexCount = tessLocation + 1;
    }

    // RiveRenderer::clipAndPushDraw (rive_renderer.cpp:496-570): the batch
    // goes into the current logical flush when it fits, else the flush ends
    // and the batch is tried once more in a fresh one (which re-issues every
    // clip draw); a batch that fits no flush is dropped.
    template <typename Push>
    void clipAndPush(size_t height, const IAABB& frame, const Needs& drawNeeds, const Gradient* grad, Push push)
    {
        for (int attempt = 0; attempt < 2; ++attempt)
        {
            if (batchFits(height, frame, drawNeeds, grad))
            {
                const uint32_t clipID = applyClip(height, frame);
                push(clipID);
                return;
            }
            endFlush();
            beginFlush();
        }
    }

    void build()
    {
        images = std::make_shared<std::vector<OracleImage>>(s.image_count);
        for (uint32_t i = 0; i < s.image_count; ++i)
            build_image_mips(s.images[i].width, s.images[i].height, s.images[i].rgba_premul, &(*images)[i]);
        beginFlush();

        const IAABB frame{0, 0, (int)s.width, (int)s.height};
        for (uint32_t di = 0; di < s.draw_count; ++di)
        {
            const rive_stream_draw& d = s.draws[di];
            Mat m;
            memcpy(m.m, d.matrix, sizeof(m.m));
            // save(); transform(matrix); clipPath()... (rive_renderer.cpp:93-112,
            // 199-228, 323-380)
            size_t height = 0;
            IAABB overall = frame;
            bool clippedOut = false;
            ClipRectState cr;
            for (uint32_t c = 0; c < d.clip_count && !clippedOut; ++c)
            {
                uint32_t cp = d.clip_paths[c];
                const rive_stream_path& sp = s.paths[cp];
                if (sp.verb_count == 0)
                {
                    clippedOut = true;
                    break;
                }
                // clipPath (rive_renderer.cpp:199-228): axis-aligned rectangles
                // become clip rects (clipRectImpl, :267-321).
                float rect[4];
                if (is_aabb(s.verbs + sp.first_verb, sp.verb_count, pathPts(cp), sp.point_count, rect))
                {
                    if (!(rect[2] - rect[0] > 0 && rect[3] - rect[1] > 0))
                    {
                        clippedOut = true;
                        break;
                    }
                    if (!cr.has || transform_rect_to_new_space(rect, m, cr.matrix))
                    {
                        cr.add(rect, m);
                        overall = overall.intersect(cr.pixelBounds());
                        if (overall.empty())
                            clippedOut = true;
                        continue;
                    }
                }
                bool equivalent = clipStack.size() > height && clipStack[height].path == cp &&
                                  memcmp(clipStack[height].matrix.m, m.m, sizeof(m.m)) == 0;
                if (!equivalent)
                {
                    IAABB pb = map_bounds_round_out(m, pathPts(cp), sp.point_count, 0, 0);
                    overall = overall.intersect(pb);
                    if (overall.empty())
                    {
                        clippedOut = true;
                        break;
                    }
                    clipStack.resize(height);
                    clipStack.push_back({cp, m, (FillRule)sp.fill_rule, pb, 0, 0});
                }
                else
                {
                    overall = overall.intersect(clipStack[height].pixelBounds);
                }
                ++height;
            }
            if (clippedOut || overall.empty())
                continue;

            // drawImage (rive_renderer.cpp:383-446): the unit rect under
            // scale(width, height) with an image paint.
            static const uint8_t kUnitVerbs[4] = {kMove, kLine, kLine, kLine};
            static const V2 kUnitPts[4] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            if (d.image != 0)
            {
                if (d.image > images->size())
                    continue;
                const OracleImage& img = (*images)[d.image - 1];
                Mat scale;
                scale.m[0] = (float)img.width;
                scale.m[3] = (float)img.height;
                Mat im = mul(m, scale);
                IAABB pixelBounds = map_bounds_round_out(im, kUnitPts, 4, 0, 0);
                if (pixelBounds.intersect(frame).empty() || pixelBounds.intersect(overall).empty())
                    continue;
                float op = std::max(0.f, std::min(1.f, std::max(0.f, d.opacity)));
                uint32_t color = ((uint32_t)std::lround(255.f * op) << 24) | 0x00ffffffu;
                Mat crInv;
                const bool hasCR = cr.inverse(&crInv);
                const Needs needs = measure(kUnitVerbs, 4, kUnitPts, 4, im, 0, Join::miter, Cap::butt,
                                            FillRule::nonZero);
                clipAndPush(height, frame, needs, nullptr, [&](uint32_t drawClipID) {
                    pushGeometryDraw(kUnitVerbs, 4, kUnitPts, 4, im, 0, Join::miter, Cap::butt, FillRule::nonZero,
                                     SOLID_COLOR_PAINT_TYPE, swizzle_rgba(color), drawClipID, 0, nullptr,
                                     shader_blend_mode(d.blend_mode), d.image, d.image_sampler,
                                     hasCR ? &crInv : nullptr);
                });
                continue;
            }
            // drawPath (rive_renderer.cpp:121-197)
            const rive_stream_path& sp = s.paths[d.path];
            if (sp.verb_count == 0)
                continue;
            bool isStroke = d.style == 1;
            if (isStroke && !(d.thickness > 0))
                continue;
            if (!(d.feather >= 0))
                continue;
            // Feathers (rive_renderer.cpp:164-184): only clockwise fills are
            // feathered, and a fill whose feather spans more than a pixel on
            // screen is drawn from a copy of its path softened for it.
            const uint8_t* dverbs = s.verbs + sp.first_verb;
            uint32_t dverbCount = sp.verb_count;
            const V2* dpts = pathPts(d.path);
            uint32_t dptCount = sp.point_count;
            std::vector<uint8_t> softVerbs;
            std::vector<V2> softPts;
            if (d.feather != 0 && !isStroke)
            {
                if ((FillRule)sp.fill_rule != FillRule::clockwise)
                    continue;
                const float matrixMaxScale = find_max_scale(m);
                if (d.feather * matrixMaxScale > 1)
                {
                    soften_path_for_feathering(dverbs, dverbCount, dpts, dptCount, d.feather, matrixMaxScale,
                                               &softVerbs, &softPts);
                    dverbs = softVerbs.data();
                    dverbCount = (uint32_t)softVerbs.size();
                    dpts = softPts.data();
                    dptCount = (uint32_t)softPts.size();
                }
            }
            const float featherRadius = d.feather != 0 ? find_feather_radius(d.feather) : 0.f;
            // PathDraw::Make bounds (draw.cpp:452-494)
            float ox = 0, oy = 0;
            if (isStroke || featherRadius != 0)
            {
                float outset = 0;
                if (isStroke)
                {
                    outset = d.thickness * .5f;
                    if ((Join)d.join == Join::miter)
                        outset *= 4.f;
                    else if ((Cap)d.cap == Cap::square)
                        outset *= 1.41421356f;
                }
                outset += featherRadius;
                // strokePixelOutset = paintMatrix.mapBoundingBox({0,0,outset,outset})
                V2 corners[4] = {{0, 0}, {outset, 0}, {0, outset}, {outset, outset}};
                Mat lin = m;
                lin.m[4] = lin.m[5] = 0;
                float l = INFINITY, t = INFINITY, r = -INFINITY, b = -INFINITY;
                for (V2 c : corners)
                {
                    V2 p = map(lin, c);
                    l = std::min(l, p.x), t = std::min(t, p.y);
                    r = std::max(r, p.x), b = std::max(b, p.y);
                }
                ox = (r - l) + 1;
                oy = (b - t) + 1;
            }
            IAABB pixelBounds = map_bounds_round_out(m, dpts, dptCount, ox, oy);
            if (pixelBounds.intersect(frame).empty())
                continue;
            // applyClip (rive_renderer.cpp:631-824)
            if (pixelBounds.intersect(overall).empty())
                continue;
            Gradient grad;
            const Gradient* gp = nullptr;
            uint32_t paintType = SOLID_COLOR_PAINT_TYPE;
            uint32_t paintValue = swizzle_rgba(d.color);
            if (d.paint_type != 0)
            {
                if (make_gradient(d, s, &grad))
                {
                    gp = &grad;
                    paintType = grad.paintType;
                }
                else
                {
                    paintValue = swizzle_rgba(0); // invalid gradients draw nothing visible
                }
            }
            float strokeRadius = 0;
            if (isStroke)
                strokeRadius = fmaxf(d.thickness * .5f, 1.17549435e-38f);
            Mat crInv;
            const bool hasCR = cr.inverse(&crInv);
            // SelectCoverageType (draw.cpp:387-418): a feather is drawn
            // through the atlas once it can render at half resolution or less.
            const bool atlas =
                featherRadius != 0 && 16.f / std::max(featherRadius * find_max_scale(m), 16.f) <= .5f;
            Needs needs = measure(dverbs, dverbCount, dpts, dptCount, m, strokeRadius, (Join)d.join, (Cap)d.cap,
                                  (FillRule)sp.fill_rule, featherRadius);
            if (atlas && !isStroke)
            {
                // counted one way (forward or reverse) rather than both
                PathDrawDesc one;
                one.verbs = dverbs;
                one.verbCount = dverbCount;
                one.pts = dpts;
                one.ptCount = dptCount;
                one.matrix = m;
                one.fillRule = (FillRule)sp.fill_rule;
                one.featherRadius = featherRadius;
                derive_directions(&one);
                one.dirs = Dirs::forward;
                one.pathID = 1;
                PathTess pt = tessellate_path_records(one, 1, 0);
                needs.tess = pt.midpointFanTessVertexCount;
            }
            clipAndPush(height, frame, needs, gp, [&](uint32_t drawClipID) {
                pushGeometryDraw(dverbs, dverbCount, dpts, dptCount, m, strokeRadius, (Join)d.join, (Cap)d.cap,
                                 (FillRule)sp.fill_rule, paintType, paintValue, drawClipID, 0, gp,
                                 shader_blend_mode(d.blend_mode), 0, 0, hasCR ? &crInv : nullptr, featherRadius,
                                 atlas ? &pixelBounds : nullptr);
            });
        }
        endFlush();
    }
};

void build_flushes(const rive_stream& s, std::vector<Flush>* out)
{
    out->clear();
    Builder(s, *out).build();
}

// ======================= color ramps ======================================

void render_ramps(const Flush& f, std::vector<uint32_t>* texture)
{
    // color_ramp.glsl:40-106: each GradientSpan is a 1px-tall strip; color0 at
    // x0, color1 at x1, linearly interpolated, with solid borders.
    const uint32_t W = kGradTextureWidth;
    texture->assign((size_t)W * f.gradHeight, 0);
    for (const rive_hip_grad_span& gs : f.gradSpans)
    {
        float x0 = float(gs.horizontal_span & 0xffffu) / 65536.f;
        float x1 = float(gs.horizontal_span >> 16) / 65536.f;
        uint32_t yflags = gs.y_with_flags;
        uint32_t y = yflags & ~GRAD_SPAN_FLAGS_MASK;
        float left = x0, right = x1;
        if (yflags & GRAD_SPAN_FLAG_LEFT_BORDER)
            left = (yflags & GRAD_SPAN_FLAG_COMPLEX_BORDER) ? 0.f : x0 - 1.f / W;
        if (yflags & GRAD_SPAN_FLAG_RIGHT_BORDER)
            right = (yflags & GRAD_SPAN_FLAG_COMPLEX_BORDER) ? 1.f : x1 + 1.f / W;
        float c0[4], c1[4];
        for (int k = 0; k < 4; ++k)
        {
            int shift = k < 3 ? 16 - 8 * k : 24; // r, g, b, a
            c0[k] = float((gs.color0 >> shift) & 0xff) / 255.f;
            c1[k] = float((gs.color1 >> shift) & 0xff) / 255.f;
        }
        for (uint32_t x = 0; x < W; ++x)
        {
            float u = (float(x) + .5f) / W;
            if (!(u >= left && u < right))
                continue;
            float c[4];
            if (u < x0)
                memcpy(c, c0, sizeof(c));
            else if (u >= x1)
                memcpy(c, c1, sizeof(c));
            else
            {
                float t = (u - x0) / (x1 - x0);
                for (int k = 0; k < 4; ++k)
                    c[k] = (c1[k] - c0[k]) * t + c0[k];
            }
            uint32_t px = 0;
            for (int k = 0; k < 4; ++k)
            {
                float v = std::min(std::max(c[k], 0.f), 1.f);
                px |= (uint32_t)floorf(v * 255.f + .5f) << (8 * k);
            }
            (*texture)[(size_t)y * W + x] = px; // RGBA8 little endian
        }
    }
}

// ======================= raster ===========================================

namespace
{
struct FxP
{
    int64_t x, y;
};
// Vertex positions on the 1/256-px grid.  A non-finite position has no
// place on it: the triangles using it are dropped, as a GPU rasteriser drops
// primitives with NaN or infinite positions.  Finite positions are kept
// exactly up to 2^30 subpixels (4 194 304 px) from the origin and clamped
// beyond, so that the int64 orientation products below stay exact.
constexpr float kSnapLimit = 1073741824.f; // 2^30
inline bool snappable(float v) { return std::isfinite(v * 256.f); }
inline int64_t snap(float v)
{
    const float c = std::min(std::max(v * 256.f, -kSnapLimit), kSnapLimit);
    return (int64_t)rintf(c);
}
// Orientation with symbolic perturbation of q by (+e, +e^2).
inline int sos_sign(FxP a, FxP b, FxP q)
{
    int64_t dx = b.x - a.x, dy = b.y - a.y;
    int64_t D = dx * (q.y - a.y) - dy * (q.x - a.x);
    if (D != 0)
        return D > 0 ? 1 : -1;
    if (dy != 0)
        return dy > 0 ? -1 : 1;
    return dx > 0 ? 1 : (dx < 0 ? -1 : 0);
}
inline int64_t edgeD(FxP a, FxP b, FxP q)
{
    return (b.x - a.x) * (q.y - a.y) - (b.y - a.y) * (q.x - a.x);
}

struct PatchVertexDef
{
    float id, outset, cov;
    float mid, mout, mcov;
    int type;
};

std::array<PatchVertexDef, 42> make_midpoint_fan_patch(std::array<uint16_t, 72>* idx)
{
    // gpu.cpp:446-687 (PatchType::midpointFan)
    std::array<PatchVertexDef, 42> v{};
    int n = 0;
    for (int i = 0; i < 8; ++i)
    {
        float l = (float)i, r = l + 1;
        v[n + 0] = {l, -1, 1, r - 1, -1, 1, 0};
        v[n + 1] = {l, +1, 0, l - 1, -1, 1, 0};
        v[n + 2] = {r, -1, 1, r - 1, +1, 0, 0};
        v[n + 3] = {r, +1, 0, l - 1, +1, 0, 0};
        n += 4;
    }
    for (int i = 0; i <= 8; ++i)
        v[n++] = {(float)i, -1, 1, (float)i - 1, -1, 1, 1};
    v[n++] = {0, 0, 1, 0, 0, 1, 2};
    int k = 0;
    const uint16_t pat[6] = {0, 1, 2, 2, 1, 3};
    for (int s = 0; s < 8; ++s)
        for (int j = 0; j < 6; ++j)
            (*idx)[k++] = (uint16_t)(s * 4 + pat[j]);
    const int fan = 32;
    for (int step = 1; step < 8; step <<= 1)
        for (int i = 0; i < 8; i += step * 2)
        {
            (*idx)[k++] = (uint16_t)(fan + i);
            (*idx)[k++] = (uint16_t)(fan + i + step);
            (*idx)[k++] = (uint16_t)(fan + i + step * 2);
        }
    (*idx)[k++] = (uint16_t)(fan);
    (*idx)[k++] = (uint16_t)(fan + 8);
    (*idx)[k++] = (uint16_t)(41);
    return v;
}

// PatchType::midpointFanCenterAA (gpu.cpp:446-687), the patch of feathered
// fills: an AA border on both sides of the edge (coverage 1/2 on the edge, 0
// at outset +-1), the curve fan and the triangle to the contour midpoint.
std::array<PatchVertexDef, 74> make_midpoint_fan_center_aa_patch(std::array<uint16_t, 120>* idx)
{
    std::array<PatchVertexDef, 74> v{};
    int n = 0;
    for (int i = 0; i < 8; ++i)
    {
        float l = (float)i, r = l + 1;
        v[n + 0] = {l, 0, .5f, r - 1, 0, .5f, 0};
        v[n + 1] = {l, 1, 0, l - 1, 0, .5f, 0};
        v[n + 2] = {r, 0, .5f, r - 1, 1, 0, 0};
        v[n + 3] = {r, 1, 0, l - 1, 1, 0, 0};
        n += 4;
    }
    for (int i = 0; i < 8; ++i)
    {
        float l = (float)i, r = l + 1;
        v[n + 0] = {l, -0.f, .5f, r - 1, -0.f, .5f, 0};
        v[n + 1] = {r, -0.f, .5f, r - 1, -1, 0, 0};
        v[n + 2] = {l, -1, 0, l - 1, -0.f, .5f, 0};
        v[n + 3] = {r, -1, 0, l - 1, -1, 0, 0};
        n += 4;
    }
    for (int i = 0; i <= 8; ++i)
        v[n++] = {(float)i, 0, 1, (float)i - 1, 0, 1, 1};
    v[n++] = {0, 0, 1, 0, 0, 1, 2};
    int k = 0;
    const uint16_t pat[6] = {0, 1, 2, 2, 1, 3};
    const uint16_t neg[6] = {0, 2, 1, 1, 2, 3};
    for (int s = 0; s < 8; ++s)
        for (int j = 0; j < 6; ++j)
            (*idx)[k++] = (uint16_t)(s * 4 + pat[j]);
    for (int s = 0; s < 8; ++s)
        for (int j = 0; j < 6; ++j)
            (*idx)[k++] = (uint16_t)(32 + s * 4 + neg[j]);
    const int fan = 64;
    for (int step = 1; step < 8; step <<= 1)
        for (int i = 0; i < 8; i += step * 2)
        {
            (*idx)[k++] = (uint16_t)(fan + i);
            (*idx)[k++] = (uint16_t)(fan + i + step);
            (*idx)[k++] = (uint16_t)(fan + i + step * 2);
        }
    (*idx)[k++] = (uint16_t)(fan);
    (*idx)[k++] = (uint16_t)(fan + 8);
    (*idx)[k++] = (uint16_t)(73);
    return v;
}

struct OutVertex
{
    bool discard;
    float x, y;   // pixel coordinates
    float c[4];   // coverages
    uint32_t pathID;
};

inline float signf_(float x) { return x > 0 ? 1.f : (x < 0 ? -1.f : 0.f); }
inline float manhattan_pixel_width(const float M[4], float nx, float ny)
{
    // draw_path_common.glsl:268-273
    float vx = M[0] * nx + M[2] * ny;
    float vy = M[1] * nx + M[3] * ny;
    return (fabsf(vx) + fabsf(vy)) * (1.f / (vx * vx + vy * vy));
}

// unpack_tessellated_path_vertex (draw_path_common.glsl:275-789), midpoint-fan
// and midpoint-fan-center-AA patches, no MSAA.
OutVertex draw_vertex(const Flush& f, const std::vector<TessVertex>& tex,
                      const PatchVertexDef& pv, int instanceID)
{
    OutVertex o{};
    int localVertexID = (int)pv.id;
    float outset = pv.outset;
    float fillCoverage = pv.cov;
    const int patchSegmentSpan = 8;
    int vertexIDOnContour = std::min(localVertexID, patchSegmentSpan - 1);
    int tessVertexIdx = instanceID * patchSegmentSpan + vertexIDOnContour;
    TessVertex tv = tex[(size_t)tessVertexIdx];
    uint32_t contourIDWithFlags = tv.contourIDWithFlags;
    uint32_t contourID = std::max(contourIDWithFlags & CONTOUR_ID_MASK, 1u);
    const ContourData& cd = f.contours[contourID - 1];
    V2 midpoint = cd.midpoint;
    uint32_t pathID = cd.pathID & 0xffffu;
    uint32_t vertexIndex0 = cd.vertexIndex0;
    const rive_hip_path_data& pd = f.pathData[pathID];
    const float M[4] = {pd.matrix[0], pd.matrix[1], pd.matrix[2], pd.matrix[3]};
    float tx = pd.matrix[4], ty = pd.matrix[5];
    float strokeRadius = pd.stroke_radius;
    float featherRadius = pd.feather_radius;
    uint32_t mirroredContourFlag = contourIDWithFlags & MIRRORED_CONTOUR_CONTOUR_FLAG;
    if (mirroredContourFlag != 0u)
    {
        localVertexID = (int)pv.mid;
        outset = pv.mout;
        fillCoverage = pv.mcov;
    }
    if (localVertexID != vertexIDOnContour)
    {
        int replacement = tessVertexIdx + localVertexID - vertexIDOnContour;
        TessVertex rv = tex[(size_t)replacement];
        if ((rv.contourIDWithFlags & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu)) !=
            (contourIDWithFlags & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu)))
        {
            bool isClosed = strokeRadius == 0.f || midpoint.x != 0.f;
            if (isClosed)
            {
                tessVertexIdx = (int)vertexIndex0;
                tv = tex[(size_t)tessVertexIdx];
            }
        }
        else
        {
            tessVertexIdx = replacement;
            tv = rv;
        }
        contourIDWithFlags = (tv.contourIDWithFlags & ~MIRRORED_CONTOUR_CONTOUR_FLAG) |
                             mirroredContourFlag;
    }
    float theta;
    float featherJoinEdge0Theta = 0, featherJoinCornerTheta = 0;
    const bool featherJoin =
        (contourIDWithFlags & JOIN_TYPE_MASK) == FEATHER_JOIN_CONTOUR_FLAG && pv.type == 0;
    if (featherJoin)
    {
        // draw_path_common.glsl:373-468: the join's stepping is worked out
        // here from the tessellation vertices on either side of it.
        uint32_t joinDataPacked = tv.z;
        float joinVertexID = float(joinDataPacked & 0xffffu);
        float joinSegmentCount = float(joinDataPacked >> 16);
        int ex = (int)(-joinVertexID - 1.f), ey = (int)(joinSegmentCount - joinVertexID + 1.f);
        if ((contourIDWithFlags & MIRRORED_CONTOUR_CONTOUR_FLAG) != 0u)
        {
            ex = -ex;
            ey = -ey;
        }
        TessVertex before = tex[(size_t)(tessVertexIdx + ex)];
        TessVertex after = tex[(size_t)(tessVertexIdx + ey)];
        if ((after.contourIDWithFlags & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu)) !=
            (before.contourIDWithFlags & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu)))
            after = tex[(size_t)vertexIndex0];
        featherJoinEdge0Theta = ubits(before.z);
        float featherJoinEdge1Theta = ubits(after.z);
        featherJoinCornerTheta = featherJoinEdge1Theta - featherJoinEdge0Theta;
        if (fabsf(featherJoinCornerTheta) > PI_)
            featherJoinCornerTheta -= _2PI_ * signf_(featherJoinCornerTheta);
        float nonHelperSegmentCount = joinSegmentCount + 1.f - float(FEATHER_JOIN_HELPER_VERTEX_COUNT);
        float forwardSegmentCount = std::min(
            std::max(rintf(fabsf(featherJoinCornerTheta) / PI_ * nonHelperSegmentCount), 1.f),
            nonHelperSegmentCount - 1.f);
        float backwardSegmentCount = nonHelperSegmentCount - forwardSegmentCount;
        if (joinVertexID <= backwardSegmentCount)
        {
            featherJoinCornerTheta = -(PI_ * signf_(featherJoinCornerTheta) - featherJoinCornerTheta);
            joinSegmentCount = backwardSegmentCount;
            if (joinVertexID == backwardSegmentCount)
                outset = -outset;
        }
        else if (joinVertexID == backwardSegmentCount + 1.f)
        {
            joinVertexID = 0;
            joinSegmentCount = 0;
            outset = 0;
        }
        else
        {
            joinVertexID -= backwardSegmentCount + 2.f;
            joinSegmentCount = forwardSegmentCount;
        }
        if (joinVertexID == joinSegmentCount)
            theta = featherJoinEdge1Theta;
        else
            theta = featherJoinEdge0Theta + featherJoinCornerTheta * (joinVertexID / joinSegmentCount);
    }
    else
    {
        theta = ubits(tv.z);
    }
    float nx = sinf(theta), ny = -cosf(theta);
    float ox = tv.x, oy = tv.y;
    float offx = 0, offy = 0;
    float covx, covy, covz = 0, covw = 0;
    if (featherRadius != 0.f)
    {
        // Never softer than 1.5 standard deviations across 1/2 px (:478-485).
        float mx = M[0] * nx + M[2] * ny, my = M[1] * nx + M[3] * ny;
        featherRadius = std::max(featherRadius, 1.f / sqrtf(mx * mx + my * my));
    }
    if (strokeRadius != 0.f)
    {
        float det = M[0] * M[3] - M[2] * M[1];
        outset *= signf_(det);
        if (contourIDWithFlags & LEFT_JOIN_CONTOUR_FLAG)
            outset = std::min(outset, 0.f);
        if (contourIDWithFlags & RIGHT_JOIN_CONTOUR_FLAG)
            outset = std::max(outset, 0.f);
        float aaRadius = featherRadius != 0.f ? featherRadius : manhattan_pixel_width(M, nx, ny) * AA_RADIUS;
        float globalCoverage = 1.f;
        if (aaRadius > strokeRadius && featherRadius == 0.f)
        {
            globalCoverage = strokeRadius / aaRadius;
            strokeRadius = aaRadius;
        }
        float vox = nx * (strokeRadius + aaRadius), voy = ny * (strokeRadius + aaRadius);
        float x = outset * (strokeRadius + aaRadius);
        covx = (1.f / (aaRadius * 2.f)) * (x + strokeRadius) + .5f;
        covy = (1.f / (aaRadius * 2.f)) * (-x + strokeRadius) + .5f;
        uint32_t joinType = contourIDWithFlags & JOIN_TYPE_MASK;
        if (joinType > ROUND_JOIN_CONTOUR_FLAG)
        {
            int peekDir = 2;
            if ((contourIDWithFlags & JOIN_TANGENT_0_CONTOUR_FLAG) == 0u)
                peekDir = -peekDir;
            if ((contourIDWithFlags & MIRRORED_CONTOUR_CONTOUR_FLAG) != 0u)
                peekDir = -peekDir;
            TessVertex other = tex[(size_t)(tessVertexIdx + peekDir)];
            float otherJoinTheta = ubits(other.z);
            float joinAngle = fabsf(otherJoinTheta - theta);
            if (joinAngle > PI_)
                joinAngle = _2PI_ - joinAngle;
            bool isTan0 = (contourIDWithFlags & JOIN_TANGENT_0_CONTOUR_FLAG) != 0u;
            bool isLeftJoin = (contourIDWithFlags & LEFT_JOIN_CONTOUR_FLAG) != 0u;
            float bisectTheta = joinAngle * (isTan0 == isLeftJoin ? -.5f : .5f) + theta;
            float bx = sinf(bisectTheta), by = -cosf(bisectTheta);
            float bisectPixelWidth = manhattan_pixel_width(M, bx, by);
            float miterRatio = cosf(joinAngle * .5f);
            float clipRadius;
            if ((joinType == MITER_CLIP_JOIN_CONTOUR_FLAG) ||
                (joinType == MITER_REVERT_JOIN_CONTOUR_FLAG && miterRatio >= .25f))
            {
                float miterInverseLimit =
                    (contourIDWithFlags & EMULATED_STROKE_CAP_CONTOUR_FLAG) != 0u ? 1.f : .25f;
                clipRadius = strokeRadius * (1.f / std::max(miterRatio, miterInverseLimit));
            }
            else
            {
                clipRadius = strokeRadius * miterRatio + bisectPixelWidth * .5f;
            }
            float clipAARadius = clipRadius + bisectPixelWidth * AA_RADIUS;
            if ((contourIDWithFlags & JOIN_TANGENT_INNER_CONTOUR_FLAG) != 0u)
            {
                float strokeAARaidus = strokeRadius + aaRadius;
                float slop = aaRadius * .125f;
                if (strokeAARaidus <= clipAARadius * miterRatio + slop)
                {
                    float miterAARadius = strokeAARaidus * (1.f / miterRatio);
                    vox = bx * miterAARadius;
                    voy = by * miterAARadius;
                }
                else
                {
                    float bax = bx * clipAARadius, bay = by * clipAARadius;
                    float k0 = vox * vox + voy * voy, k1 = bax * bax + bay * bay;
                    // k * inverse(float2x2(vertexOffset, bisectAAOffset))
                    float a = vox, b = voy, c = bax, d = bay;
                    float idet = 1.f / (a * d - c * b);
                    float nvx = (k0 * d - k1 * b) * idet;
                    float nvy = (-k0 * c + k1 * a) * idet;
                    vox = nvx;
                    voy = nvy;
                }
            }
            float ptx = fabsf(outset) * vox, pty = fabsf(outset) * voy;
            float clipDistance =
                (clipAARadius - (ptx * bx + pty * by)) / (bisectPixelWidth * (AA_RADIUS * 2.f));
            if ((contourIDWithFlags & LEFT_JOIN_CONTOUR_FLAG) != 0u)
                covy = clipDistance;
            else
                covx = clipDistance;
        }
        covx *= globalCoverage;
        covy *= globalCoverage;
        covy = std::max(covy, 1e-4f);
        if (featherRadius != 0.f)
            covx = FEATHER_COVERAGE_BIAS - covx; // a feathered stroke
        float px_ = outset * vox, py_ = outset * voy;
        offx = M[0] * px_ + M[2] * py_;
        offy = M[1] * px_ + M[3] * py_;
        if (pv.type != 0)
        {
            o.discard = true;
            return o;
        }
    }
    else
    {
        covx = fillCoverage;
        covy = -1.f;
        if (featherRadius != 0.f)
        {
            // draw_path_common.glsl:661-744: a feathered edge starts out as a
            // flat edge feathering y0 = fillCoverage; the spokes of a feather
            // join carry the corner's coordinates instead.
            covy = FEATHER_COVERAGE_BIAS;
            covz = HORIZONTAL_COTANGENT_VALUE;
            covw = fillCoverage;
            if (featherJoin)
            {
                if (featherJoinCornerTheta < 0.f)
                {
                    featherJoinEdge0Theta += featherJoinCornerTheta;
                    featherJoinCornerTheta = -featherJoinCornerTheta;
                }
                float spokeTheta = theta - featherJoinEdge0Theta;
                float sm = spokeTheta + kPI_OVER_2;
                spokeTheta = (sm - _2PI_ * floorf(sm / _2PI_)) - kPI_OVER_2; // GLSL mod
                spokeTheta = std::min(std::max(spokeTheta, 0.f), featherJoinCornerTheta);
                if (spokeTheta > featherJoinCornerTheta * .5f)
                    spokeTheta = featherJoinCornerTheta - spokeTheta;
                float c4[4];
                pack_feathered_fill_coverages(featherJoinCornerTheta, V2{sinf(spokeTheta), cosf(spokeTheta)},
                                              outset, c4);
                covx = c4[0];
                covy = c4[1];
                covz = c4[2];
                covw = c4[3];
            }
            float px_ = (outset * featherRadius) * nx, py_ = (outset * featherRadius) * ny;
            offx = M[0] * px_ + M[2] * py_;
            offy = M[1] * px_ + M[3] * py_;
        }
        else
        {
            // sign(MUL(outset * norm, inverse(M))) * AA_RADIUS
            float det = M[0] * M[3] - M[2] * M[1];
            float idet = 1.f / det;
            float vx = outset * nx, vy = outset * ny;
            // row vector * inverse(M); inverse(M) columns: (M3, -M1)/det, (-M2, M0)/det
            float rx = (vx * M[3] + vy * -M[1]) * idet;
            float ry = (vx * -M[2] + vy * M[0]) * idet;
            offx = signf_(rx) * AA_RADIUS;
            offy = signf_(ry) * AA_RADIUS;
        }
        if (bool(contourIDWithFlags & MIRRORED_CONTOUR_CONTOUR_FLAG) !=
            bool(contourIDWithFlags & NEGATE_PATH_FILL_COVERAGE_FLAG))
            covx = -covx;
        if (pv.type == 2)
        {
            ox = midpoint.x;
            oy = midpoint.y;
        }
    }
    o.x = (M[0] * ox + M[2] * oy) + offx + tx;
    o.y = (M[1] * ox + M[3] * oy) + offy + ty;
    o.c[0] = covx;
    o.c[1] = covy;
    o.c[2] = covz;
    o.c[3] = covw;
    o.pathID = pathID;
    o.discard = false;
    return o;
}

float dither_noise(int x, int y)
{
    // common.glsl:269-274 with ditherScale 1/256, ditherBias -1/512
    // (gpu.cpp:822-824); fragCoord is the pixel center.
    float fx = float(x) + .5f, fy = float(y) + .5f;
    float v = 0.06711056f * fx + 0.00583715f * fy;
    float v1 = v - floorf(v);
    float w = 52.9829189f * v1;
    float v2 = w - floorf(w);
    return v2 * (1.f / 256.f) + (-.5f / 256.f);
}
inline float unorm8(uint32_t packed, int k) { return float((packed >> (8 * k)) & 0xff) / 255.f; }
inline uint32_t to_unorm8(float v)
{
    v = std::min(std::max(v, 0.f), 1.f);
    return (uint32_t)floorf(v * 255.f + .5f);
}
} // namespace

void build_image_mips(uint32_t width, uint32_t height, const uint8_t* rgba, OracleImage* out)
{
    // Full mip chain (the reference builds it on the GPU, render_context_impl.hpp:43-70),
    // each texel the rounded mean of its 2x2 footprint, clamped at odd edges.
    uint32_t levels = 1;
    while ((std::max(width, height) >> levels) != 0)
        ++levels;
    levels = std::min<uint32_t>(levels, 15);
    out->width = width;
    out->height = height;
    out->levels = levels;
    out->levelOffset.assign(levels + 1, 0);
    uint32_t total = 0;
    for (uint32_t l = 0; l < levels; ++l)
    {
        out->levelOffset[l] = total;
        total += std::max(width >> l, 1u) * std::max(height >> l, 1u);
    }
    out->levelOffset[levels] = total;
    out->texels.assign(total, 0);
    memcpy(out->texels.data(), rgba, (size_t)width * height * 4);
    for (uint32_t l = 0; l + 1 < levels; ++l)
    {
        const int sw = (int)std::max(width >> l, 1u), sh = (int)std::max(height >> l, 1u);
        const int dw = (int)std::max(width >> (l + 1), 1u), dh = (int)std::max(height >> (l + 1), 1u);
        const uint32_t* src = out->texels.data() + out->levelOffset[l];
        uint32_t* dst = out->texels.data() + out->levelOffset[l + 1];
        for (int y = 0; y < dh; ++y)
            for (int x = 0; x < dw; ++x)
            {
                const int x0 = std::min(2 * x, sw - 1), x1 = std::min(2 * x + 1, sw - 1);
                const int y0 = std::min(2 * y, sh - 1), y1 = std::min(2 * y + 1, sh - 1);
                const uint32_t t[4] = {src[y0 * sw + x0], src[y0 * sw + x1], src[y1 * sw + x0], src[y1 * sw + x1]};
                uint32_t px = 0;
                for (int k = 0; k < 4; ++k)
                {
                    uint32_t sum = 0;
                    for (uint32_t v : t)
                        sum += (v >> (8 * k)) & 0xffu;
                    px |= ((sum + 2u) >> 2) << (8 * k);
                }
                dst[y * dw + x] = px;
            }
    }
}

namespace
{
int wrap_coord(int i, int n, uint32_t mode)
{
    if (mode == 1) // repeat
    {
        int m = i % n;
        return m < 0 ? m + n : m;
    }
    if (mode == 2) // mirrored repeat
    {
        int m = i % (2 * n);
        if (m < 0)
            m += 2 * n;
        return m < n ? m : 2 * n - 1 - m;
    }
    return std::min(std::max(i, 0), n - 1);
}
} // namespace

void sample_image(const OracleImage& img, float u, float v, float lod, uint32_t sampler, float out[4])
{
    // VkSamplerCreateInfo of pipeline_manager_vulkan.cpp:95-110: min/mag filter
    // per ImageFilter, mipmapMode NEAREST (level = ceil(lod + .5) - 1), address
    // mode per ImageWrap.
    const uint32_t wrapX = sampler % 3, wrapY = (sampler / 3) % 3, filter = sampler / 9;
    int level = 0;
    if (lod > 0.f)
        level = std::min((int)ceilf(lod + .5f) - 1, (int)img.levels - 1);
    level = std::max(level, 0);
    const int w = std::max((int)img.width >> level, 1), h = std::max((int)img.height >> level, 1);
    const uint32_t* tex = img.texels.data() + img.levelOffset[level];
    auto unorm = [](uint32_t t, int k) { return float((t >> (8 * k)) & 0xffu) / 255.f; };
    if (filter == 1)
    {
        const int i = wrap_coord((int)floorf(u * (float)w), w, wrapX);
        const int j = wrap_coord((int)floorf(v * (float)h), h, wrapY);
        for (int k = 0; k < 4; ++k)
            out[k] = unorm(tex[j * w + i], k);
        return;
    }
    const float x = u * (float)w - .5f, y = v * (float)h - .5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int i0 = wrap_coord((int)fx0, w, wrapX), i1 = wrap_coord((int)fx0 + 1, w, wrapX);
    const int j0 = wrap_coord((int)fy0, h, wrapY), j1 = wrap_coord((int)fy0 + 1, h, wrapY);
    for (int k = 0; k < 4; ++k)
    {
        const float c00 = unorm(tex[j0 * w + i0], k), c10 = unorm(tex[j0 * w + i1], k);
        const float c01 = unorm(tex[j1 * w + i0], k), c11 = unorm(tex[j1 * w + i1], k);
        const float top = (c10 - c00) * ax + c00;
        const float bot = (c11 - c01) * ax + c01;
        out[k] = (bot - top) * ay + top;
    }
}

// Gradient paint color, unpremultiplied (draw_path.vert:285-331,450-476): the
// paint matrix maps the pixel to t, the ramp row is sampled with linear
// filtering at texel centers.
void sample_gradient(const Flush& f, const rive_hip_paint_aux& aux, const rive_hip_paint_data& paint,
                     const std::vector<uint32_t>& ramps, int x, int y, float color[4])
{
    const uint32_t paintType = paint.params & 0xfu;
    float fx = float(x) + .5f, fy = float(y) + .5f;
    float pcx = aux.paint_matrix[0] * fx + aux.paint_matrix[2] * fy + aux.paint_matrix[4];
    float pcy = aux.paint_matrix[1] * fx + aux.paint_matrix[3] * fy + aux.paint_matrix[5];
    float t = paintType == LINEAR_GRADIENT_PAINT_TYPE ? pcx : sqrtf(pcx * pcx + pcy * pcy);
    t = std::min(std::max(t, 0.f), 1.f);
    float span = aux.grad_span[0] > .9f ? 2.f : aux.grad_span[1];
    float u = span > 1.f ? (1.f - 1.f / 512.f) * t + (.5f / 512.f) : (1.f / 512.f) * t + span;
    float rowf = ubits(paint.value);
    uint32_t row = (uint32_t)(rowf * f.gradHeight);
    float tc = u * 512.f - .5f;
    int i0 = (int)floorf(tc);
    float fr = tc - (float)i0;
    int ia = std::min(std::max(i0, 0), 511), ib = std::min(std::max(i0 + 1, 0), 511);
    uint32_t ta = ramps[(size_t)row * 512 + ia], tb = ramps[(size_t)row * 512 + ib];
    for (int k = 0; k < 4; ++k)
        color[k] = (unorm8(tb, k) - unorm8(ta, k)) * fr + unorm8(ta, k);
}

// Resolves every pixel a path touched, in one band of rows
// (draw_raster_order_path.frag:14-238).
void resolve_touched(const Flush& f, const std::vector<uint32_t>& ramps, uint32_t pathID, bool isStroke,
                     const std::vector<uint32_t>& touched, const std::vector<float>& cov,
                     std::vector<float>& clipCov, std::vector<uint32_t>& clipIDs, uint32_t* rgba,
                     uint8_t* edgeMask)
{
    const uint32_t W = f.width;
    const rive_hip_paint_data& paint = f.paintData[pathID];
    const rive_hip_paint_aux& aux = f.paintAux[pathID];
    const uint32_t paintType = paint.params & 0xfu;
    const bool evenOdd = (paint.params & PAINT_FLAG_EVEN_ODD_FILL) != 0u;
    const bool clockwise = !isStroke && !evenOdd && !(paint.params & PAINT_FLAG_NON_ZERO_FILL);
    for (uint32_t p : touched)
    {
        int x = (int)(p % W), y = (int)(p / W);
        float count = cov[p];
        float coverage;
        if (clockwise)
            coverage = std::min(std::max(count, 0.f), 1.f);
        else
        {
            coverage = fabsf(count);
            if (evenOdd)
            {
                float h = coverage * .5f;
                coverage = 1.f - fabsf((h - floorf(h)) * 2.f + -1.f);
            }
            coverage = std::min(coverage, 1.f);
        }
        if (paintType == CLIP_UPDATE_PAINT_TYPE)
        {
            uint32_t clipID = paint.value >> 16;
            uint32_t outerClipID = paint.params >> 16;
            if (outerClipID != 0)
            {
                float outer = clipIDs[p] == outerClipID ? clipCov[p] : 0.f;
                coverage = std::min(coverage, outer);
            }
            clipCov[p] = round_to_f16(coverage);
            clipIDs[p] = clipID;
            continue;
        }
        uint32_t clipID = paint.params >> 16;
        if (clipID != 0)
            coverage = clipIDs[p] == clipID ? std::min(clipCov[p], coverage) : 0.f;
        if (paint.params & PAINT_FLAG_HAS_CLIP_RECT)
            coverage = std::min(std::max(clip_rect_coverage(aux.clip_rect_inverse_matrix, x, y), 0.f), coverage);
        if (edgeMask != nullptr && coverage > 0.f && coverage < 1.f)
            edgeMask[p] = 1;
        const uint32_t blendMode = (paint.params >> 4) & 0xfu;
        uint32_t dst = rgba[p];
        float color[4];
        // Image paints (draw_path.vert:479-493): the paint color times the
        // image sample; unmultiplied under advanced blend.
        float imageColor[4] = {1.f, 1.f, 1.f, 1.f};
        const uint32_t pimg = f.pathImage[pathID];
        if ((pimg & 0xffffu) != 0)
        {
            const float fx = float(x) + .5f, fy = float(y) + .5f;
            const float u = aux.image_matrix[0] * fx + aux.image_matrix[2] * fy + aux.image_matrix[4];
            const float v = aux.image_matrix[1] * fx + aux.image_matrix[3] * fy + aux.image_matrix[5];
            sample_image((*f.images)[(pimg & 0xffffu) - 1], u, v, aux.image_lod, pimg >> 16, imageColor);
            if (blendMode != BLEND_SRC_OVER)
            {
                const float inv = imageColor[3] != 0.f ? 1.f / imageColor[3] : 0.f;
                for (int k = 0; k < 3; ++k)
                    imageColor[k] *= inv;
            }
        }
        if (blendMode != BLEND_SRC_OVER)
        {
            // draw_raster_order_path.frag:191-204: the paint stays
            // unpremultiplied, the advanced blend mixes it with the blend
            // coefficients by the destination alpha, then premultiplies.
            float src[4];
            if (paintType == SOLID_COLOR_PAINT_TYPE)
            {
                for (int k = 0; k < 4; ++k)
                    src[k] = unorm8(paint.value, k);
            }
            else
            {
                sample_gradient(f, aux, paint, ramps, x, y, src);
            }
            src[3] *= coverage;
            for (int k = 0; k < 4; ++k)
                src[k] *= imageColor[k];
            const float dp[4] = {unorm8(dst, 0), unorm8(dst, 1), unorm8(dst, 2), unorm8(dst, 3)};
            float coeffs[3];
            advanced_blend_coeffs(src, dp, blendMode, coeffs);
            for (int k = 0; k < 3; ++k)
                color[k] = (coeffs[k] - src[k]) * dp[3] + src[k];
            color[3] = src[3];
            for (int k = 0; k < 3; ++k)
                color[k] *= color[3];
        }
        else if (paintType == SOLID_COLOR_PAINT_TYPE)
        {
            float a = unorm8(paint.value, 3);
            color[0] = unorm8(paint.value, 0) * a;
            color[1] = unorm8(paint.value, 1) * a;
            color[2] = unorm8(paint.value, 2) * a;
            color[3] = a;
            for (int k = 0; k < 4; ++k)
                color[k] *= coverage;
            for (int k = 0; k < 4; ++k)
                color[k] *= imageColor[k];
        }
        else
        {
            sample_gradient(f, aux, paint, ramps, x, y, color);
            color[3] *= coverage;
            color[0] *= color[3];
            color[1] *= color[3];
            color[2] *= color[3];
        }
        float paintAlpha = color[3];
        float out[4];
        for (int k = 0; k < 4; ++k)
            out[k] = color[k] + unorm8(dst, k) * (1.f - paintAlpha);
        if (f.dither && paintAlpha != 0.f)
        {
            float n = dither_noise(x, y);
            out[0] += n;
            out[1] += n;
            out[2] += n;
        }
        rgba[p] = to_unorm8(out[0]) | (to_unorm8(out[1]) << 8) | (to_unorm8(out[2]) << 16) |
                  (to_unorm8(out[3]) << 24);
    }
}

void render_flush(const Flush& f, const std::vector<TessVertex>& tex, uint32_t* rgba, uint8_t* edgeMask)
{
    const uint32_t W = f.width, H = f.height;
    std::vector<uint32_t> ramps;
    render_ramps(f, &ramps);

    // Initial color: clear (premultiplied, gpu.cpp SwizzleRiveColorToRGBAPremul),
    // unless this logical flush loads the target.
    if (!f.loadPreserve)
    {
        uint32_t c = f.clearColor;
        float a = float(c >> 24) / 255.f;
        float r = float((c >> 16) & 0xff) / 255.f * a;
        float g = float((c >> 8) & 0xff) / 255.f * a;
        float b = float(c & 0xff) / 255.f * a;
        uint32_t px = to_unorm8(r) | (to_unorm8(g) << 8) | (to_unorm8(b) << 16) | (to_unorm8(a) << 24);
        std::fill(rgba, rgba + (size_t)W * H, px);
    }
    std::vector<float> clipCov((size_t)W * H, 0.f);
    std::vector<uint32_t> clipIDs((size_t)W * H, 0u);
    std::vector<float> cov((size_t)W * H, 0.f);
    std::vector<uint32_t> stamp((size_t)W * H, 0u);

    std::array<uint16_t, 72> fanIdx;
    const auto fanPatch = make_midpoint_fan_patch(&fanIdx);
    std::array<uint16_t, 120> caaIdx;
    const auto caaPatch = make_midpoint_fan_center_aa_patch(&caaIdx);

    // The feather atlas (render_atlas.glsl): every atlas draw's triangles,
    // placed by its atlas transform, rendered into its own region of an R16F
    // texture -- fills add signed coverage with no culling (additive blend,
    // the sign from the triangle's facing), strokes keep the maximum of
    // clockwise triangles (max blend) -- each blend rounding to fp16
    // (gpu.hpp:2076-2080, pipeline_manager_vulkan.cpp:69).
    std::vector<std::vector<float>> atlases(f.draws.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < (int)f.draws.size(); ++k)
    {
        const DrawInfo& di = f.draws[(size_t)k];
        if (!di.atlas || di.tessVertexCount == 0)
            continue;
        std::vector<float>& A = atlases[(size_t)k];
        A.assign((size_t)di.aw * di.ah, 0.f);
        const PatchVertexDef* patch = di.featheredFill ? caaPatch.data() : fanPatch.data();
        const uint16_t* idx = di.featheredFill ? caaIdx.data() : fanIdx.data();
        const int NV = di.featheredFill ? 74 : 42, NT = di.featheredFill ? 40 : 24;
        const int firstInstance = (int)(di.tessLocation / 8);
        const int nInst = (int)((di.tessLocation + di.tessVertexCount) / 8) - firstInstance;
        std::vector<OutVertex> iv((size_t)NV);
        for (int inst = 0; inst < nInst; ++inst)
        {
            for (int v = 0; v < NV; ++v)
            {
                OutVertex o = draw_vertex(f, tex, patch[v], firstInstance + inst);
                o.x = o.x * di.atlasScale + di.atlasTx; // render_atlas.glsl:35-37
                o.y = o.y * di.atlasScale + di.atlasTy;
                iv[(size_t)v] = o;
            }
            for (int t = 0; t < NT; ++t)
            {
                const OutVertex* V[3] = {&iv[idx[t * 3 + 0]], &iv[idx[t * 3 + 1]], &iv[idx[t * 3 + 2]]};
                if (V[0]->discard || V[1]->discard || V[2]->discard)
                    continue;
                bool ok = true;
                for (int q = 0; q < 3; ++q)
                    ok = ok && snappable(V[q]->x) && snappable(V[q]->y);
                if (!ok)
                    continue;
                FxP a{snap(V[0]->x), snap(V[0]->y)}, b{snap(V[1]->x), snap(V[1]->y)}, c{snap(V[2]->x), snap(V[2]->y)};
                int64_t area = edgeD(a, b, c);
                if (area == 0 || (di.isStroke && area < 0))
                    continue;
                float sign = 1.f;
                if (area < 0)
                {
                    // a back-facing fill triangle subtracts; test it clockwise
                    std::swap(b, c);
                    std::swap(V[1], V[2]);
                    area = -area;
                    sign = -1.f;
                }
                int64_t minx = std::min({a.x, b.x, c.x}), maxx = std::max({a.x, b.x, c.x});
                int64_t miny = std::min({a.y, b.y, c.y}), maxy = std::max({a.y, b.y, c.y});
                int x0 = (int)std::max<int64_t>(0, (minx - 128) >> 8);
                int x1 = (int)std::min<int64_t>((int64_t)di.aw - 1, (maxx - 128 + 255) >> 8);
                int y0 = (int)std::max<int64_t>(0, (miny - 128) >> 8);
                int y1 = (int)std::min<int64_t>((int64_t)di.ah - 1, (maxy - 128 + 255) >> 8);
                const double inv = 1.0 / (double)area;
                for (int y = y0; y <= y1; ++y)
                    for (int x = x0; x <= x1; ++x)
                    {
                        FxP q{(int64_t)x * 256 + 128, (int64_t)y * 256 + 128};
                        if (sos_sign(a, b, q) <= 0 || sos_sign(b, c, q) <= 0 || sos_sign(c, a, q) <= 0)
                            continue;
                        double la = (double)edgeD(b, c, q) * inv;
                        double lb = (double)edgeD(c, a, q) * inv;
                        double lc = (double)edgeD(a, b, q) * inv;
                        float c4[4];
                        for (int r = 0; r < 4; ++r)
                            c4[r] = (float)(la * V[0]->c[r] + lb * V[1]->c[r] + lc * V[2]->c[r]);
                        float& texel = A[(size_t)y * di.aw + x];
                        if (di.isStroke)
                        {
                            float fc = c4[0] < FEATHER_COVERAGE_THRESHOLD ? eval_feathered_stroke(c4)
                                                                          : std::min(c4[0], c4[1]);
                            texel = std::max(texel, round_to_f16(fc));
                        }
                        else
                        {
                            float fc = c4[1] < FEATHER_COVERAGE_THRESHOLD ? eval_feathered_fill(c4) : c4[0];
                            texel = round_to_f16(texel + sign * fc);
                        }
                    }
            }
        }
    }

    // Rows are split into bands, one thread per band at a time: a pixel's
    // triangles still accumulate in submission order, so the result does not
    // depend on the thread count.
    const int bands = std::max(1, std::min((int)H, 2 * omp_get_max_threads()));
    std::vector<std::vector<uint32_t>> touchedBands(bands);
    std::vector<OutVertex> ov;
    std::vector<int> instRows;

    for (size_t drawIdx = 0; drawIdx < f.draws.size(); ++drawIdx)
    {
        const DrawInfo& di = f.draws[drawIdx];
        if (di.tessVertexCount == 0)
            continue;
        if (di.atlas)
        {
            // The blit (draw_mesh.frag:88-99): over the pixel bounds, the
            // atlas sampled with linear filtering at pixel * scale +
            // translate, clamped to [0, 1].
            const std::vector<float>& A = atlases[drawIdx];
            const uint32_t pathID = di.pathID;
#pragma omp parallel for schedule(dynamic, 1)
            for (int band = 0; band < bands; ++band)
            {
                const int yb0 = std::max((int)((int64_t)band * H / bands), di.blitT);
                const int yb1 = std::min((int)((int64_t)(band + 1) * H / bands), di.blitB) - 1;
                std::vector<uint32_t>& touched = touchedBands[band];
                touched.clear();
                for (int y = yb0; y <= yb1; ++y)
                    for (int x = std::max(di.blitL, 0); x < std::min(di.blitR, (int)W); ++x)
                    {
                        const float u = ((float)x + .5f) * di.atlasScale + di.atlasTx - .5f;
                        const float v = ((float)y + .5f) * di.atlasScale + di.atlasTy - .5f;
                        const float fu = floorf(u), fv = floorf(v);
                        const float wu = u - fu, wv = v - fv;
                        auto texel = [&](int i, int j) {
                            i = std::min(std::max(i, 0), (int)di.aw - 1);
                            j = std::min(std::max(j, 0), (int)di.ah - 1);
                            return A[(size_t)j * di.aw + i];
                        };
                        const int i0 = (int)fu, j0 = (int)fv;
                        const float top = texel(i0, j0) * (1.f - wu) + texel(i0 + 1, j0) * wu;
                        const float bot = texel(i0, j0 + 1) * (1.f - wu) + texel(i0 + 1, j0 + 1) * wu;
                        const float sample = top * (1.f - wv) + bot * wv;
                        const size_t p = (size_t)y * W + x;
                        stamp[p] = pathID;
                        touched.push_back((uint32_t)p);
                        cov[p] = std::min(std::max(sample, 0.f), 1.f);
                    }
                resolve_touched(f, ramps, pathID, di.isStroke, touched, cov, clipCov, clipIDs, rgba, edgeMask);
            }
            continue;
        }
        const uint32_t pathID = di.pathID;
        const bool isStroke = di.isStroke;
        const int firstInstance = (int)(di.tessLocation / 8);
        const int endInstance = (int)((di.tessLocation + di.tessVertexCount) / 8);
        const int nInst = endInstance - firstInstance;
        // Feathered fills draw midpoint-fan-center-AA patches (draw.cpp:1866-1870).
        const PatchVertexDef* patch = di.featheredFill ? caaPatch.data() : fanPatch.data();
        const uint16_t* idx = di.featheredFill ? caaIdx.data() : fanIdx.data();
        const int NV = di.featheredFill ? 74 : 42, NT = di.featheredFill ? 40 : 24;
        ov.resize((size_t)nInst * NV);
        instRows.resize((size_t)nInst * 2);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nInst; ++k)
        {
            int64_t lo = INT64_MAX, hi = INT64_MIN;
            for (int v = 0; v < NV; ++v)
            {
                const OutVertex o = draw_vertex(f, tex, patch[v], firstInstance + k);
                ov[(size_t)k * NV + v] = o;
                if (!o.discard && snappable(o.x) && snappable(o.y))
                {
                    const int64_t y = snap(o.y);
                    lo = std::min(lo, y);
                    hi = std::max(hi, y);
                }
            }
            instRows[2 * k] = lo <= hi ? (int)std::max<int64_t>(0, (lo - 128) >> 8) : 1;
            instRows[2 * k + 1] = lo <= hi ? (int)std::min<int64_t>((int64_t)H - 1, (hi - 128 + 255) >> 8) : 0;
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int band = 0; band < bands; ++band)
        {
            const int yb0 = (int)((int64_t)band * H / bands), yb1 = (int)((int64_t)(band + 1) * H / bands) - 1;
            std::vector<uint32_t>& touched = touchedBands[band];
            touched.clear();
            for (int k = 0; k < nInst; ++k)
            {
                if (instRows[2 * k] > yb1 || instRows[2 * k + 1] < yb0)
                    continue;
                const OutVertex* iv = &ov[(size_t)k * NV];
                for (int t = 0; t < NT; ++t)
                {
                    const OutVertex& A = iv[idx[t * 3 + 0]];
                    const OutVertex& B = iv[idx[t * 3 + 1]];
                    const OutVertex& C = iv[idx[t * 3 + 2]];
                    if (A.discard || B.discard || C.discard)
                        continue;
                    if (!snappable(A.x) || !snappable(A.y) || !snappable(B.x) || !snappable(B.y) ||
                        !snappable(C.x) || !snappable(C.y))
                        continue;
                    FxP a{snap(A.x), snap(A.y)}, b{snap(B.x), snap(B.y)}, c{snap(C.x), snap(C.y)};
                    int64_t area = edgeD(a, b, c);
                    if (area <= 0)
                        continue; // counterclockwise or degenerate: culled
                    int64_t minx = std::min({a.x, b.x, c.x}), maxx = std::max({a.x, b.x, c.x});
                    int64_t miny = std::min({a.y, b.y, c.y}), maxy = std::max({a.y, b.y, c.y});
                    int x0 = (int)std::max<int64_t>(0, (minx - 128) >> 8);
                    int x1 = (int)std::min<int64_t>((int64_t)W - 1, (maxx - 128 + 255) >> 8);
                    int y0 = (int)std::max<int64_t>(yb0, (miny - 128) >> 8);
                    int y1 = (int)std::min<int64_t>(yb1, (maxy - 128 + 255) >> 8);
                    const double inv = 1.0 / (double)area;
                    for (int y = y0; y <= y1; ++y)
                    {
                        // Candidate columns: the row's crossings of the
                        // triangle's edges, widened by 2 px (the exact
                        // tests below decide; this only skips pixels far
                        // outside a thin triangle's bounding box).
                        int cx0 = x0, cx1 = x1;
                        {
                            const double yc = (double)y * 256 + 128;
                            double lo = INFINITY, hi = -INFINITY;
                            int crossings = 0;
                            const FxP* vs[3] = {&a, &b, &c};
                            for (int e = 0; e < 3; ++e)
                            {
                                const FxP& p0 = *vs[e];
                                const FxP& p1 = *vs[(e + 1) % 3];
                                if (p0.y == p1.y || yc < (double)std::min(p0.y, p1.y) ||
                                    yc > (double)std::max(p0.y, p1.y))
                                    continue;
                                const double xe = (double)p0.x + (yc - (double)p0.y) *
                                                                      (double)(p1.x - p0.x) / (double)(p1.y - p0.y);
                                lo = std::min(lo, xe);
                                hi = std::max(hi, xe);
                                ++crossings;
                            }
                            if (crossings >= 2)
                            {
                                cx0 = std::max(x0, (int)std::floor((lo - 128) / 256) - 2);
                                cx1 = std::min(x1, (int)std::ceil((hi - 128) / 256) + 2);
                            }
                        }
                        for (int x = cx0; x <= cx1; ++x)
                        {
                            FxP q{(int64_t)x * 256 + 128, (int64_t)y * 256 + 128};
                            if (sos_sign(a, b, q) <= 0 || sos_sign(b, c, q) <= 0 || sos_sign(c, a, q) <= 0)
                                continue;
                            double la = (double)edgeD(b, c, q) * inv;
                            double lb = (double)edgeD(c, a, q) * inv;
                            double lc = (double)edgeD(a, b, q) * inv;
                            size_t p = (size_t)y * W + x;
                            float c4[4];
                            for (int q = 0; q < 4; ++q)
                                c4[q] = (float)(la * A.c[q] + lb * B.c[q] + lc * C.c[q]);
                            float count = stamp[p] == pathID ? cov[p] : 0.f;
                            if (stamp[p] != pathID)
                            {
                                stamp[p] = pathID;
                                touched.push_back((uint32_t)p);
                            }
                            // apply_frag_coverage (draw_path.vert:500-545)
                            if (isStroke)
                            {
                                float fc = c4[0] < FEATHER_COVERAGE_THRESHOLD ? eval_feathered_stroke(c4)
                                                                              : std::min(c4[0], c4[1]);
                                count = std::max(fc, count);
                            }
                            else
                            {
                                float fc = c4[1] < FEATHER_COVERAGE_THRESHOLD ? eval_feathered_fill(c4) : c4[0];
                                count = count + fc;
                            }
                            cov[p] = round_to_f16(count);
                        }
                    }
                }
            }
            resolve_touched(f, ramps, pathID, isStroke, touched, cov, clipCov, clipIDs, rgba, edgeMask);
        }
    }
}

void tessellate_flush(const Flush& f, std::vector<TessVertex>* tex)
{
    uint32_t rows = (f.tessVertexCount + kTessTextureWidth - 1) / kTessTextureWidth;
    rows = std::max(rows, 1u);
    tex->assign((size_t)rows * kTessTextureWidth, TessVertex{0, 0, 0, 0});
    for (const TessVertexSpan& s : f.spans)
        tessellate_span(s, f.contours.data(), reinterpret_cast<const float*>(f.pathData.data()), tex->data(),
                        rows);
}
} // namespace oracle
