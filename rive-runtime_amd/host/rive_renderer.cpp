// Clean-room front end: RiveRenderer (clip stack, clip IDs), RenderContext
// (per-flush records, gradients) and RenderContextHIPImpl.  See
// rive_renderer.hpp for the reference interfaces these mirror.
#include "rive_renderer.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <utility>

namespace rive
{
// ---------------------------------------------------------------- Mat2D
Mat2D Mat2D::operator*(const Mat2D& b) const
{
    const float* a = m;
    const float* c = b.m;
    return Mat2D(a[0] * c[0] + a[2] * c[1], a[1] * c[0] + a[3] * c[1], a[0] * c[2] + a[2] * c[3],
                 a[1] * c[2] + a[3] * c[3], a[0] * c[4] + a[2] * c[5] + a[4],
                 a[1] * c[4] + a[3] * c[5] + a[5]);
}

bool Mat2D::invert(Mat2D* out) const
{
    // src/math/mat2d.cpp:179-200, in float.
    const float aa = m[0], ab = m[1], ac = m[2], ad = m[3], atx = m[4], aty = m[5];
    float det = aa * ad - ab * ac;
    if (det == 0.f)
        return false;
    det = 1.f / det;
    *out = Mat2D(ad * det, -ab * det, -ac * det, aa * det, (ac * aty - ad * atx) * det,
                 (ab * atx - aa * aty) * det);
    return true;
}

float Mat2D::findMaxScale() const
{
    // src/math/mat2d_find_max_scale.cpp:24-67
    const float xx = m[0], xy = m[1], yx = m[2], yy = m[3];
    if (xy == 0 && yx == 0)
        return std::max(fabsf(xx), fabsf(yy));
    float a = xx * xx + xy * xy;
    float b = xx * yx + yy * xy;
    float c = yx * yx + yy * yy;
    float bSqd = b * b;
    const float eps = 1.f / (1 << 12);
    float result;
    if (bSqd <= eps * eps)
        result = std::max(a, c);
    else
    {
        float aminusc = a - c;
        float apluscdiv2 = (a + c) * .5f;
        float x = sqrtf(aminusc * aminusc + 4 * bSqd) * .5f;
        result = apluscdiv2 + x;
    }
    if (!std::isfinite(result))
        result = 0;
    result = std::max(result, 0.f);
    return sqrtf(result);
}

// ---------------------------------------------------------------- RawPath
void RawPath::injectImplicitMoveIfNeeded()
{
    if (!m_contourIsOpen)
        move(m_points.empty() ? Vec2D{0, 0} : m_points[m_lastMoveIdx]);
}
void RawPath::move(Vec2D p)
{
    m_contourIsOpen = true;
    m_lastMoveIdx = m_points.size();
    m_points.push_back(p);
    m_verbs.push_back(PathVerb::move);
}
void RawPath::line(Vec2D p)
{
    injectImplicitMoveIfNeeded();
    m_points.push_back(p);
    m_verbs.push_back(PathVerb::line);
}
void RawPath::cubic(Vec2D a, Vec2D b, Vec2D c)
{
    injectImplicitMoveIfNeeded();
    m_points.push_back(a);
    m_points.push_back(b);
    m_points.push_back(c);
    m_verbs.push_back(PathVerb::cubic);
}
void RawPath::quad(Vec2D a, Vec2D b)
{
    // RawPath::quadToCubic (src/math/raw_path.cpp:484-495)
    injectImplicitMoveIfNeeded();
    Vec2D p0 = m_points.back();
    auto lerp = [](Vec2D u, Vec2D v, float t) { return Vec2D{u.x + (v.x - u.x) * t, u.y + (v.y - u.y) * t}; };
    cubic(lerp(p0, a, 2 / 3.f), lerp(b, a, 2 / 3.f), b);
}
void RawPath::close()
{
    if (m_contourIsOpen)
    {
        m_verbs.push_back(PathVerb::close);
        m_contourIsOpen = false;
    }
}

// ---------------------------------------------------------------- Gradient
namespace
{
bool valid_stops(const float stops[], size_t count)
{
    if (count == 0)
        return false;
    for (size_t i = 0; i < count; ++i)
    {
        if (!(stops[i] >= 0 && stops[i] <= 1))
            return false;
        if (i > 0 && !(stops[i - 1] <= stops[i]))
            return false;
    }
    return true;
}
} // namespace

std::shared_ptr<Gradient> Gradient::MakeLinear(float sx, float sy, float ex, float ey, const ColorInt colors[],
                                               const float stops[], size_t count)
{
    // renderer/src/gradient.cpp:38-108
    if (!valid_stops(stops, count))
        return nullptr;
    auto g = std::make_shared<Gradient>();
    g->paintType = 2;
    g->colors.assign(colors, colors + count);
    g->stops.assign(stops, stops + count);
    float firstStop = stops[0], lastStop = stops[count - 1];
    if ((firstStop != 0 || lastStop != 1) && lastStop - firstStop > (1.f / (1 << 12)))
    {
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
    return g;
}

std::shared_ptr<Gradient> Gradient::MakeRadial(float cx, float cy, float radius, const ColorInt colors[],
                                               const float stops[], size_t count)
{
    // renderer/src/gradient.cpp:110-175
    if (!valid_stops(stops, count))
        return nullptr;
    auto g = std::make_shared<Gradient>();
    g->paintType = 3;
    g->colors.assign(colors, colors + count);
    g->stops.assign(stops, stops + count);
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
    return g;
}

// ---------------------------------------------------------------- RiveRenderer
ColorInt colorModulateOpacity(ColorInt c, float opacity)
{
    const float o = (float)(c >> 24) / 255.f * opacity;
    const uint32_t a = (uint32_t)std::lround(255.f * std::max(0.f, std::min(1.f, o)));
    return (a << 24) | (c & 0x00ffffffu);
}

std::shared_ptr<Gradient> Gradient::getModulated(float opacity) const
{
    if (opacity == 1.f)
        return std::const_pointer_cast<Gradient>(shared_from_this());
    if (m_lastModulated && m_lastModulatedOpacity == opacity)
        return m_lastModulated;
    auto g = std::make_shared<Gradient>();
    g->paintType = paintType;
    g->stops = stops;
    memcpy(g->coeffs, coeffs, sizeof(coeffs));
    for (ColorInt c : colors)
        g->colors.push_back(colorModulateOpacity(c, opacity));
    m_lastModulated = g;
    m_lastModulatedOpacity = opacity;
    return g;
}

std::shared_ptr<RenderBuffer> Factory::makeRenderBuffer(RenderBufferType type, uint32_t, size_t bytes)
{
    return std::make_shared<HostRenderBuffer>(type, bytes);
}
std::shared_ptr<RenderShader> Factory::makeLinearGradient(float sx, float sy, float ex, float ey,
                                                          const ColorInt colors[], const float stops[], size_t count)
{
    return Gradient::MakeLinear(sx, sy, ex, ey, colors, stops, count);
}
std::shared_ptr<RenderShader> Factory::makeRadialGradient(float cx, float cy, float radius, const ColorInt colors[],
                                                          const float stops[], size_t count)
{
    return Gradient::MakeRadial(cx, cy, radius, colors, stops, count);
}
std::shared_ptr<RiveRenderPath> Factory::makeRenderPath(RawPath& path, FillRule rule)
{
    auto p = std::make_shared<RiveRenderPath>(std::move(path), rule);
    path = RawPath();
    return p;
}
std::shared_ptr<RiveRenderPath> Factory::makeEmptyRenderPath() { return std::make_shared<RiveRenderPath>(); }
std::shared_ptr<RiveRenderPaint> Factory::makeRenderPaint() { return std::make_shared<RiveRenderPaint>(); }
std::shared_ptr<RenderImage> Factory::decodeImage(const uint8_t*, size_t) { return nullptr; }

RiveRenderer::IAABB RiveRenderer::IAABB::intersect(const IAABB& o) const
{
    return {std::max(l, o.l), std::max(t, o.t), std::min(r, o.r), std::min(b, o.b)};
}

RiveRenderer::IAABB RiveRenderer::MapBounds(const Mat2D& m, const RawPath& raw, float ox, float oy)
{
    float l = INFINITY, t = INFINITY, r = -INFINITY, b = -INFINITY;
    for (Vec2D p : raw.points())
    {
        Vec2D q = m.map(p);
        l = std::min(l, q.x), t = std::min(t, q.y);
        r = std::max(r, q.x), b = std::max(b, q.y);
    }
    l -= ox, r += ox, t -= oy, b += oy;
    return {(int)floorf(l), (int)floorf(t), (int)ceilf(r), (int)ceilf(b)};
}

RiveRenderer::RiveRenderer(gpu::RenderContext* ctx) : m_context(ctx)
{
    // The unit rect drawImage draws through (rive_renderer.cpp:428-434).
    m_unitRectPath.lineTo(1, 0);
    m_unitRectPath.lineTo(1, 1);
    m_unitRectPath.lineTo(0, 1);
    RenderState s;
    s.overallClipPixelBounds = {0, 0, (int)ctx->frameDescriptor().renderTargetWidth,
                                (int)ctx->frameDescriptor().renderTargetHeight};
    m_stack.push_back(s);
}

void RiveRenderer::save() { m_stack.push_back(m_stack.back()); }
void RiveRenderer::restore()
{
    if (m_stack.size() > 1)
        m_stack.pop_back();
}
void RiveRenderer::transform(const Mat2D& m) { m_stack.back().matrix = m_stack.back().matrix * m; }

void RiveRenderer::modulateOpacity(float opacity)
{
    // rive_renderer.cpp:115-119
    m_stack.back().modulatedOpacity = std::max(0.f, m_stack.back().modulatedOpacity * opacity);
}

void RiveRenderer::drawImageMesh(const RenderImage*, ImageSampler, std::shared_ptr<RenderBuffer>,
                                 std::shared_ptr<RenderBuffer>, std::shared_ptr<RenderBuffer>, uint32_t, uint32_t,
                                 BlendMode, float)
{
    ++m_unsupportedDraws;
    fprintf(stderr, "RiveRenderer::drawImageMesh: image meshes are not supported by the HIP backend\n");
}

void RiveRenderer::drawImage(const RenderImage* renderImage, ImageSampler sampler, BlendMode blendMode,
                             float opacity)
{
    // rive_renderer.cpp:383-446: an image is the unit rect under
    // scale(width, height) with an image paint.
    auto* image = static_cast<const RiveRenderImage*>(renderImage);
    if (image == nullptr || image->handle() == nullptr)
        return;
    save();
    transform(Mat2D((float)image->width(), 0, 0, (float)image->height(), 0, 0));
    RiveRenderPaint paint;
    // rive_renderer.cpp:400-402: the modulated opacity scales the image's;
    // the unit rect below is then drawn at opacity 1.
    paint.image(image, std::max(0.f, opacity * m_stack.back().modulatedOpacity));
    m_stack.back().modulatedOpacity = 1.f;
    paint.blendMode(blendMode);
    paint.imageSampler(sampler);
    drawPath(&m_unitRectPath, &paint);
    restore();
}

bool RiveRenderer::IsAABB(const RawPath& raw, float rect[4])
{
    // rive_renderer.cpp:16-55: a move and three lines, any further points
    // equal to the first, with opposite corners sharing x and y.
    const auto& verbs = raw.verbs();
    const auto& pts = raw.points();
    if (verbs.size() < 4 || verbs[0] != PathVerb::move || verbs[1] != PathVerb::line ||
        verbs[2] != PathVerb::line || verbs[3] != PathVerb::line)
        return false;
    for (size_t i = 4; i < pts.size(); ++i)
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
// rive_renderer.cpp:235-265: re-expresses `rect` under `newMatrix` so that
// currentMatrix * rect == newMatrix * newRect, if it stays a rect.
bool transform_rect_to_new_space(float rect[4], const Mat2D& currentMatrix, const Mat2D& newMatrix)
{
    if (currentMatrix == newMatrix)
        return true;
    Mat2D currentToNew;
    if (!newMatrix.invert(&currentToNew))
        return false;
    currentToNew = currentToNew * currentMatrix;
    const float maxSkew = fmaxf(fabsf(currentToNew[1]), fabsf(currentToNew[2]));
    const float maxScale = fmaxf(fabsf(currentToNew[0]), fabsf(currentToNew[3]));
    if (maxSkew > 1.0e-5f && maxScale > 1.0e-5f) // math::EPSILON
        return false;
    const Vec2D a = currentToNew.map({rect[0], rect[1]}), b = currentToNew.map({rect[2], rect[3]});
    rect[0] = std::min(a.x, b.x);
    rect[1] = std::min(a.y, b.y);
    rect[2] = std::max(a.x, b.x);
    rect[3] = std::max(a.y, b.y);
    return true;
}
} // namespace

void RiveRenderer::clipPath(RenderPath* renderPath)
{
    // rive_renderer.cpp:199-228: axis-aligned rectangles become clip rects
    // (raster ordering supports them, render_context.cpp:456-461).
    auto* path = static_cast<RiveRenderPath*>(renderPath);
    RenderState& st = m_stack.back();
    if (st.overallClipPixelBounds.empty())
        return;
    if (path->getRawPath().empty())
    {
        st.overallClipPixelBounds = {};
        return;
    }
    float rect[4];
    if (IsAABB(path->getRawPath(), rect))
        clipRectImpl(rect, path);
    else
        clipPathImpl(path);
}

void RiveRenderer::clipRectImpl(float rect[4], RiveRenderPath* path)
{
    // rive_renderer.cpp:267-321
    RenderState& st = m_stack.back();
    if (!(rect[2] - rect[0] > 0 && rect[3] - rect[1] > 0)) // isEmptyOrNaN
    {
        st.overallClipPixelBounds = {};
        return;
    }
    if (st.hasClipRect && !transform_rect_to_new_space(rect, st.matrix, st.clipRectMatrix))
    {
        clipPathImpl(path);
        return;
    }
    if (!st.hasClipRect)
    {
        memcpy(st.clipRect, rect, sizeof(st.clipRect));
        st.clipRectMatrix = st.matrix;
    }
    else
    {
        st.clipRect[0] = fmaxf(st.clipRect[0], rect[0]);
        st.clipRect[1] = fmaxf(st.clipRect[1], rect[1]);
        st.clipRect[2] = fminf(st.clipRect[2], rect[2]);
        st.clipRect[3] = fminf(st.clipRect[3], rect[3]);
    }
    st.hasClipRect = true;
    RawPath corners;
    corners.move({st.clipRect[0], st.clipRect[1]});
    corners.line({st.clipRect[2], st.clipRect[1]});
    corners.line({st.clipRect[2], st.clipRect[3]});
    corners.line({st.clipRect[0], st.clipRect[3]});
    st.overallClipPixelBounds = st.overallClipPixelBounds.intersect(MapBounds(st.clipRectMatrix, corners, 0, 0));
    // ClipRectInverseMatrix::reset (gpu.cpp:689-706): pixel space -> the
    // normalized rect [-1, -1, +1, +1]; an empty rect or singular matrix clips
    // everything away.
    const float w = st.clipRect[2] - st.clipRect[0], h = st.clipRect[3] - st.clipRect[1];
    const Mat2D toRect = st.clipRectMatrix * Mat2D(w * .5f, 0, 0, h * .5f, (st.clipRect[0] + st.clipRect[2]) * .5f,
                                                   (st.clipRect[1] + st.clipRect[3]) * .5f);
    if (w <= 0 || h <= 0 || !toRect.invert(&st.clipRectInverse))
        st.clipRectInverse = Mat2D(0, 0, 0, 0, 0, 0);
}

void RiveRenderer::clipPathImpl(RiveRenderPath* path)
{
    // rive_renderer.cpp:323-380
    RenderState& st = m_stack.back();
    const size_t h = st.clipStackHeight;
    bool equivalent = m_clipStack.size() > h && m_clipStack[h].path == path && m_clipStack[h].matrix == st.matrix;
    if (!equivalent)
    {
        IAABB pb = MapBounds(st.matrix, path->getRawPath(), 0, 0);
        st.overallClipPixelBounds = st.overallClipPixelBounds.intersect(pb);
        if (st.overallClipPixelBounds.empty())
            return;
        m_clipStack.resize(h);
        m_clipStack.push_back({st.matrix, path, path->getFillRule(), pb, m_context->newClipSerial()});
    }
    else
    {
        st.overallClipPixelBounds = st.overallClipPixelBounds.intersect(m_clipStack[h].pixelBounds);
    }
    st.clipStackHeight = h + 1;
}

void RiveRenderer::drawPath(RenderPath* renderPath, RenderPaint* renderPaint)
{
    // rive_renderer.cpp:121-197 and applyClip (:631-824), rasterOrdering.
    auto* path = static_cast<RiveRenderPath*>(renderPath);
    auto* paint = static_cast<RiveRenderPaint*>(renderPaint);
    const RenderState& st = m_stack.back();
    if (path->getRawPath().empty())
        return;
    if (paint->getIsStroked() && !(paint->getThickness() > 0))
        return;
    if (!(paint->getFeather() >= 0))
        return;
    if (st.overallClipPixelBounds.empty())
        return;
    // Feathers (rive_renderer.cpp:164-184): only clockwise fills are
    // feathered; a fill whose feather spans more than a pixel on screen draws
    // a copy of its path softened for the feather.
    const float feather = paint->getFeather();
    if (feather != 0 && !paint->getIsStroked())
    {
        if (path->getFillRule() != FillRule::clockwise)
            return;
        const float matrixMaxScale = st.matrix.findMaxScale();
        if (feather * matrixMaxScale > 1)
            path = m_context->keepForFrame(path->makeSoftenedCopyForFeathering(feather, matrixMaxScale));
    }
    float ox = 0, oy = 0;
    if (paint->getIsStroked() || feather != 0)
    {
        // PathDraw::Make stroke and feather outset (draw.cpp:452-494)
        float outset = 0;
        if (paint->getIsStroked())
        {
            outset = paint->getThickness() * .5f;
            if (paint->getJoin() == StrokeJoin::miter)
                outset *= 4.f;
            else if (paint->getCap() == StrokeCap::square)
                outset *= 1.41421356f;
        }
        if (feather != 0)
            outset += feather * (3.f / 2); // find_feather_radius
        Mat2D lin(st.matrix[0], st.matrix[1], st.matrix[2], st.matrix[3], 0, 0);
        Vec2D c[4] = {{0, 0}, {outset, 0}, {0, outset}, {outset, outset}};
        float l = INFINITY, t = INFINITY, r = -INFINITY, b = -INFINITY;
        for (Vec2D p : c)
        {
            Vec2D q = lin.map(p);
            l = std::min(l, q.x), t = std::min(t, q.y);
            r = std::max(r, q.x), b = std::max(b, q.y);
        }
        ox = (r - l) + 1;
        oy = (b - t) + 1;
    }
    const IAABB frame{0, 0, (int)m_context->frameDescriptor().renderTargetWidth,
                      (int)m_context->frameDescriptor().renderTargetHeight};
    IAABB pixelBounds = MapBounds(st.matrix, path->getRawPath(), ox, oy);
    if (pixelBounds.intersect(frame).empty())
        return;
    if (pixelBounds.intersect(st.overallClipPixelBounds).empty())
        return;
    // The clip stack the draw sits under; RenderContext issues its clip
    // draws and IDs per logical flush (applyClip).
    std::vector<gpu::ClipRef> clips;
    clips.reserve(st.clipStackHeight);
    for (size_t i = 0; i < st.clipStackHeight; ++i)
    {
        const ClipElement& ce = m_clipStack[i];
        clips.push_back({ce.serial, ce.path, ce.matrix, ce.fillRule, !ce.pixelBounds.intersect(frame).empty()});
    }
    gpu::DrawRecord rec = m_context->makeRecord(path, st.matrix, paint, 0, false, 0, path->getFillRule(),
                                           st.hasClipRect ? &st.clipRectInverse : nullptr, st.modulatedOpacity);
    const float featherRadius = feather * (3.f / 2);
    if (feather != 0 && 16.f / std::max(featherRadius * st.matrix.findMaxScale(), 16.f) <= .5f)
    {
        // SelectCoverageType (draw.cpp:387-418): once the feather can render
        // at half resolution or less the reference draws it through an
        // offscreen atlas.  Its PathData carries the atlas transform (the
        // visible pixel bounds at scale 16 / (r * maxScale), 2 texels of
        // padding) and its fill is tessellated one way (draw.cpp:659-680).
        const IAABB visible = pixelBounds.intersect(frame);
        const float scale = 16.f / std::max(featherRadius * st.matrix.findMaxScale(), 16.f);
        rec.path.atlas_transform[0] = scale;
        rec.path.atlas_transform[1] = 2.f - (float)visible.l * scale;
        rec.path.atlas_transform[2] = 2.f - (float)visible.t * scale;
        rec.atlas = true;
        rec.atlasScissored = !(visible.l == pixelBounds.l && visible.t == pixelBounds.t &&
                               visible.r == pixelBounds.r && visible.b == pixelBounds.b);
        if (!paint->getIsStroked())
        {
            const Mat2D& m = st.matrix;
            rec.draw.contour_directions = m[0] * m[3] - m[2] * m[1] < 0 ? 1u : 0u; // reverse : forward
        }
    }
    m_context->pushDrawBatch(std::move(clips), rec);
}

namespace gpu
{
namespace
{
// ConvertBlendModeToPLSBlendMode (gpu.cpp:717-753): the shader's 4-bit mode.
uint32_t shader_blend_mode(BlendMode m)
{
    switch (m)
    {
        case BlendMode::screen: return 1;
        case BlendMode::overlay: return 2;
        case BlendMode::darken: return 3;
        case BlendMode::lighten: return 4;
        case BlendMode::colorDodge: return 5;
        case BlendMode::colorBurn: return 6;
        case BlendMode::hardLight: return 7;
        case BlendMode::softLight: return 8;
        case BlendMode::difference: return 9;
        case BlendMode::exclusion: return 10;
        case BlendMode::multiply: return 11;
        case BlendMode::hue: return 12;
        case BlendMode::saturation: return 13;
        case BlendMode::color: return 14;
        case BlendMode::luminosity: return 15;
        default: return 0;
    }
}
uint32_t swizzle_rgba(uint32_t argb)
{
    uint32_t a = argb >> 24, r = (argb >> 16) & 0xff, g = (argb >> 8) & 0xff, b = argb & 0xff;
    return r | (g << 8) | (b << 16) | (a << 24);
}
// calc_polar_segments_per_radian<kPolarPrecision> (bezier_utils.hpp:108-113)
float polar_segments_per_radian(float r)
{
    float cosTheta = 1.f - (1.f / 8.f) / r;
    return .5f / acosf(std::max(cosTheta, -1.f));
}
} // namespace

void RenderContext::beginFrame(const FrameDescriptor& fd)
{
    m_frame = fd;
    m_batches.clear();
    m_modulatedGradients.clear();
    m_framePaths.clear();
    m_canvases.clear();
    m_provLastClipID = 0;
    m_provContentSerial = 0;
    m_provClipIDs.clear();
}

void RenderContext::beginCanvas(const FrameDescriptor& fd, rive_hip_target* target)
{
    m_frame = fd;
    m_canvases.push_back({target, m_batches.size(), fd});
    m_provLastClipID = 0;
    m_provContentSerial = 0;
    m_provClipIDs.clear();
}

void RenderContext::pushDrawBatch(std::vector<ClipRef> clips, const DrawRecord& draw)
{
    size_t first = 0;
    if (m_provContentSerial != 0)
        for (size_t i = clips.size(); i-- > 0;)
            if (clips[i].serial == m_provContentSerial)
            {
                first = i + 1;
                break;
            }
    for (size_t i = first; i < clips.size(); ++i)
        m_provClipIDs.push_back({clips[i].serial, ++m_provLastClipID});
    if (!clips.empty())
        m_provContentSerial = clips.back().serial;
    m_batches.push_back({std::move(clips), draw});
}

uint32_t RenderContext::getClipContentID() const
{
    for (auto it = m_provClipIDs.rbegin(); it != m_provClipIDs.rend(); ++it)
        if (it->first == m_provContentSerial)
            return it->second;
    return 0;
}


DrawRecord RenderContext::makeRecord(const RiveRenderPath* path, const Mat2D& m, const RiveRenderPaint* paint,
                                     uint32_t clipID, bool clipUpdate, uint32_t outerClipID, FillRule fillRule,
                                     const Mat2D* clipRectInverse, float opacity)
{
    DrawRecord r{};
    r.raw = &path->getRawPath();
    r.residentVerb = path->residentVerb();
    r.residentPoint = path->residentPoint();
    const bool stroke = paint != nullptr && paint->getIsStroked();
    float strokeRadius = stroke ? fmaxf(paint->getThickness() * .5f, 1.17549435e-38f) : 0.f;
    // find_feather_radius (draw.cpp:313-318): two standard deviations wide,
    // the radius covers three.
    const float featherRadius =
        !clipUpdate && paint != nullptr && paint->getFeather() != 0 ? paint->getFeather() * (3.f / 2) : 0.f;
    memcpy(r.path.matrix, m.values(), sizeof(r.path.matrix));
    r.path.stroke_radius = strokeRadius;
    r.path.feather_radius = featherRadius;
    // PaintData (gpu.cpp:879-939)
    uint32_t params;
    if (clipUpdate)
    {
        params = 0; // clipUpdate paint type
        r.paint.value = clipID << 16;
        params |= outerClipID << 16;
    }
    else
    {
        const Gradient* g = paint->getGradient();
        if (g != nullptr)
        {
            params = g->paintType;
            if (opacity != 1.f)
            {
                m_modulatedGradients.push_back(g->getModulated(opacity));
                g = m_modulatedGradients.back().get();
            }
            r.gradient = g;
        }
        else
        {
            params = 1; // solid color
            ColorInt color = paint->getColor();
            if (opacity != 1.f)
                color = colorModulateOpacity(color, opacity);
            r.paint.value = swizzle_rgba(color);
        }
        params |= clipID << 16 | shader_blend_mode(paint->getBlendMode()) << 4;
        if (paint->getImage() != nullptr)
        {
            // PAINT_FLAG_HAS_IMAGE; PaintAuxData image branch (gpu.cpp:1006-1034):
            // the inverse image matrix and a constant mip LOD.
            params |= 0x800u;
            r.image = paint->getImage();
            r.draw.image_sampler = paint->getImageSampler().asKey();
            Mat2D inv;
            if (!m.invert(&inv))
                inv = Mat2D();
            memcpy(r.aux.image_matrix, inv.values(), sizeof(r.aux.image_matrix));
            const float w = (float)r.image->width(), h = (float)r.image->height();
            const float dudx = inv[0] * w, dudy = inv[2] * h;
            const float dvdx = inv[1] * w, dvdy = inv[3] * h;
            const float maxScaleFactorPow2 = std::max(dudx * dudx + dvdx * dvdx, dudy * dudy + dvdy * dvdy);
            r.aux.image_lod = log2f(std::max(maxScaleFactorPow2, 1.f)) * .5f + -.5f; // MIP_MAP_LOD_BIAS
        }
    }
    if (!stroke)
    {
        if (fillRule == FillRule::nonZero)
            params |= 0x100u;
        else if (fillRule == FillRule::evenOdd)
            params |= 0x200u;
    }
    // PaintAuxData clip rect (gpu.cpp:1044-1062): the inverse matrix and
    // inverse fwidth, or WideOpen (scale 0, translate 1: uniform coverage 1).
    if (clipRectInverse != nullptr)
    {
        params |= 0x400u; // PAINT_FLAG_HAS_CLIP_RECT
        const Mat2D& cm = *clipRectInverse;
        memcpy(r.aux.clip_rect_inverse_matrix, cm.values(), sizeof(r.aux.clip_rect_inverse_matrix));
        r.aux.inverse_fwidth[0] = -1.f / (fabsf(cm[0]) + fabsf(cm[1]));
        r.aux.inverse_fwidth[1] = -1.f / (fabsf(cm[2]) + fabsf(cm[3]));
    }
    else
    {
        const float wideOpen[6] = {0, 0, 0, 0, 1, 1};
        memcpy(r.aux.clip_rect_inverse_matrix, wideOpen, sizeof(wideOpen));
    }
    r.paint.params = params;
    // PathDraw's per-draw constants (draw.cpp:561-825)
    rive_hip_draw& d = r.draw;
    d.is_stroke = stroke ? 1 : (featherRadius != 0 ? 2 : 0);
    // PathDraw::Make (draw.cpp:510-538): with interior triangulation on, a
    // plain fill of at least 512 x 512 px on screen (find_transformed_area,
    // gpu.cpp:1160-1173) and at most 256 verbs is laid out as outer-cubic
    // patches (the triangulation controller's guards alone decide).
    if (d.is_stroke == 0 && m_impl->interiorTriangulation() && path->getRawPath().verbs().size() <= 256 &&
        !path->getRawPath().points().empty())
    {
        const auto& pts = path->getRawPath().points();
        float l = pts[0].x, t = pts[0].y, rr = l, b = t;
        for (const Vec2D& q : pts)
        {
            l = std::min(l, q.x), rr = std::max(rr, q.x);
            t = std::min(t, q.y), b = std::max(b, q.y);
        }
        const float* mv = m.values();
        auto map = [mv](float x, float y) { return Vec2D{mv[0] * x + mv[2] * y + mv[4], mv[1] * x + mv[3] * y + mv[5]}; };
        const Vec2D q0 = map(l, t), q1 = map(rr, t), q2 = map(rr, b), q3 = map(l, b);
        const Vec2D v0{q1.x - q0.x, q1.y - q0.y}, v1{q2.x - q0.x, q2.y - q0.y}, v2{q3.x - q0.x, q3.y - q0.y};
        const float area = (fabsf(v0.x * v1.y - v0.y * v1.x) + fabsf(v1.x * v2.y - v1.y * v2.x)) * .5f;
        if (area >= 512.f * 512.f)
            d.is_stroke = 3;
    }
    d.join = stroke ? (uint32_t)paint->getJoin() : 0;
    d.cap = stroke ? (uint32_t)paint->getCap() : 0;
    memcpy(d.matrix, m.values(), sizeof(d.matrix));
    d.stroke_radius = strokeRadius;
    d.parametric_precision = 4.f;
    if (stroke || featherRadius != 0)
    {
        // initForMidpointFan (draw.cpp:776-889): strokes and feathers take
        // polar segments for their radius on screen, a feather's capped where
        // its segments reach pi/16; wide feathers tessellate coarser.
        d.matrix_max_scale = m.findMaxScale();
        float r_ = 0;
        if (featherRadius != 0)
            r_ = std::min(featherRadius * d.matrix_max_scale, 1 / (8 * (1 - 0.99518472667f)));
        if (stroke)
            r_ += strokeRadius * d.matrix_max_scale;
        d.polar_segments_per_radian = polar_segments_per_radian(r_);
        if (featherRadius > 1)
            d.parametric_precision = std::min(4.f * 100.f / (featherRadius * d.matrix_max_scale), 4.f);
    }
    if (stroke)
    {
        d.contour_directions = 0; // forward
    }
    else if (fillRule == FillRule::clockwise)
    {
        float det = m[0] * m[3] - m[2] * m[1];
        d.contour_directions = det < 0 ? 2 : 3;
        if (det < 0)
            d.contour_flags |= 1u << 24; // NEGATE_PATH_FILL_COVERAGE_FLAG
    }
    else
    {
        d.contour_directions = 3; // reverseThenForward
    }
    return r;
}

namespace
{
// Resource limits of one logical flush (render_context.cpp:43-53,484-492,
// 517-541,617-654; gpu.hpp:85-103): path and clip IDs up to MaxPathID(1) - 1,
// contour IDs up to 65535, tessellation vertices up to a 2048 x 2048 texture
// less 25 padding vertices, gradient texture rows up to 2048.
constexpr uint64_t kMaxPathID = (0x1fu << 10) - 1 - 1023 - 1;
constexpr uint64_t kMaxContourID = 65535;
constexpr uint64_t kMaxTessVerticesBeforePadding = 2048ull * 2048ull - (8 + 16 + 1);
constexpr uint64_t kMaxGradientRows = 2048;

bool is_simple_ramp(const Gradient& g)
{
    const size_t n = g.stops.size();
    return n == 1 || (n == 2 && g.stops[0] == 0 && g.stops[1] == 1);
}
std::pair<uint32_t, uint32_t> simple_key(const Gradient& g)
{
    return {g.colors[0], g.colors[std::min<size_t>(1, g.stops.size() - 1)]};
}

// One logical flush being filled: LogicalFlush's resource counters, its
// clip IDs and clip content, and its gradient ramps.
struct LogicalFlushState
{
    std::vector<DrawRecord> draws;
    uint64_t paths = 0, contours = 0, tess = 0;
    uint32_t lastClipID = 0;
    uint64_t contentSerial = 0; // clip element whose content the clip buffer holds
    std::vector<std::pair<uint64_t, uint32_t>> clipIDs; // serial -> clip ID in this flush
    std::vector<std::pair<uint32_t, uint32_t>> simple;
    std::vector<const Gradient*> complex;

    uint32_t clipIDOf(uint64_t serial) const
    {
        for (auto it = clipIDs.rbegin(); it != clipIDs.rend(); ++it)
            if (it->first == serial)
                return it->second;
        return 0;
    }
    // The first clip element (index) the clip buffer does not already hold.
    size_t firstClipToIssue(const std::vector<ClipRef>& clips) const
    {
        if (contentSerial != 0)
            for (size_t i = clips.size(); i-- > 0;)
                if (clips[i].serial == contentSerial)
                    return i + 1;
        return 0;
    }
    bool newGradient(const Gradient* g, bool* simpleRamp) const
    {
        *simpleRamp = is_simple_ramp(*g);
        if (*simpleRamp)
            return std::find(simple.begin(), simple.end(), simple_key(*g)) == simple.end();
        for (const Gradient* c : complex)
            if (c->colors == g->colors && c->stops == g->stops)
                return false;
        return true;
    }
};
} // namespace

bool RenderContext::prepareCanvases(rive_hip_target* target, FlushDescriptor* out)
{
    if (m_canvases.empty())
        return false;
    std::vector<DrawRecord> draws;
    std::vector<rive_hip_canvas> canvases;
    for (size_t c = 0; c < m_canvases.size(); ++c)
    {
        const size_t end = c + 1 < m_canvases.size() ? m_canvases[c + 1].firstBatch : m_batches.size();
        LogicalFlushState st;
        const size_t first = draws.size();
        for (size_t b = m_canvases[c].firstBatch; b < end; ++b)
        {
            Batch& batch = m_batches[b];
            const size_t firstClip = st.firstClipToIssue(batch.clips);
            uint32_t parent = firstClip > 0 ? st.clipIDOf(batch.clips[firstClip - 1].serial) : 0;
            for (size_t i = firstClip; i < batch.clips.size(); ++i)
            {
                const ClipRef& ce = batch.clips[i];
                const uint32_t id = ++st.lastClipID;
                st.clipIDs.push_back({ce.serial, id});
                if (ce.onFrame)
                    draws.push_back(makeRecord(ce.path, ce.matrix, nullptr, id, true, parent, ce.fillRule));
                parent = id;
            }
            if (!batch.clips.empty())
                st.contentSerial = batch.clips.back().serial;
            draws.push_back(batch.draw);
            draws.back().paint.params = (draws.back().paint.params & 0xffffu) | (parent << 16);
        }
        canvases.push_back({m_canvases[c].target, (uint32_t)first + 1, (uint32_t)(draws.size() - first),
                            m_canvases[c].frame.clearColor, m_canvases[c].frame.loadAction});
    }
    *out = layoutFlush(draws, target, m_frame.loadAction, &canvases);
    m_batches.clear();
    m_canvases.clear();
    return true;
}

bool RenderContext::flush(rive_hip_target* target, uint32_t maxFlushes, uint32_t* flushCount)
{
    if (flushCount != nullptr)
        *flushCount = 0;
    if (!m_canvases.empty())
    {
        // Batched: every canvas in one flush, each with its own clip IDs; no
        // logical-flush split (no counterpart in the reference).
        FlushDescriptor fd;
        prepareCanvases(target, &fd);
        bool ok = m_impl->flushFrame({fd});
        if (flushCount != nullptr)
            *flushCount = 1;
        m_modulatedGradients.clear();
    m_framePaths.clear();
        return ok;
    }

    // 1. Resource counts of every draw and every clip element, on the
    //    device: they decide where the frame splits.
    std::vector<DrawRecord> items;
    std::vector<size_t> drawItem(m_batches.size());
    std::vector<std::pair<uint64_t, size_t>> clipItem; // serial -> item
    auto clipItemOf = [&](uint64_t serial) -> size_t {
        for (auto it = clipItem.rbegin(); it != clipItem.rend(); ++it)
            if (it->first == serial)
                return it->second;
        return (size_t)-1;
    };
    for (size_t b = 0; b < m_batches.size(); ++b)
    {
        for (const ClipRef& ce : m_batches[b].clips)
            if (ce.onFrame && clipItemOf(ce.serial) == (size_t)-1)
            {
                clipItem.push_back({ce.serial, items.size()});
                items.push_back(makeRecord(ce.path, ce.matrix, nullptr, 0, true, 0, ce.fillRule));
            }
        drawItem[b] = items.size();
        items.push_back(m_batches[b].draw);
    }
    std::vector<uint32_t> tess, contours;
    if (!items.empty() && !m_impl->countDraws(items, &tess, &contours))
        return false;

    // 2. RiveRenderer::clipAndPushDraw over the batches (rive_renderer.cpp:
    //    496-570): a batch -- the clip draws applyClip issues plus the draw --
    //    joins the current logical flush when it fits, else the flush ends
    //    and the batch is retried once in a fresh flush, which re-issues its
    //    whole clip stack; a batch that fits no flush is dropped.
    std::vector<LogicalFlushState> flushes(1);
    for (size_t b = 0; b < m_batches.size(); ++b)
    {
        Batch& batch = m_batches[b];
        bool pushed = false;
        for (int attempt = 0; attempt < 2 && !pushed; ++attempt)
        {
            LogicalFlushState& st = flushes.back();
            const size_t firstClip = st.firstClipToIssue(batch.clips);
            uint64_t paths = 0, ctrs = 0, verts = 0;
            auto add = [&](size_t item) {
                if (tess[item] > 0) // draw.cpp:1378-1391
                {
                    paths += 1;
                    ctrs += contours[item];
                    verts += tess[item];
                }
            };
            for (size_t i = firstClip; i < batch.clips.size(); ++i)
                if (batch.clips[i].onFrame)
                    add(clipItemOf(batch.clips[i].serial));
            add(drawItem[b]);
            bool fits = (uint64_t)st.lastClipID + (batch.clips.size() - firstClip) <= kMaxPathID &&
                        st.paths + paths <= kMaxPathID && st.contours + ctrs <= kMaxContourID &&
                        st.tess + verts <= kMaxTessVerticesBeforePadding;
            bool simpleRamp = false, newRamp = false;
            if (fits && batch.draw.gradient != nullptr)
            {
                newRamp = st.newGradient(batch.draw.gradient, &simpleRamp);
                const uint64_t simple = st.simple.size() + (newRamp && simpleRamp ? 1 : 0);
                const uint64_t complex = st.complex.size() + (newRamp && !simpleRamp ? 1 : 0);
                fits = (simple + 255) / 256 + complex <= kMaxGradientRows;
            }
            if (!fits)
            {
                if (attempt == 0)
                    flushes.emplace_back();
                continue;
            }
            // The draw reads the innermost clip of its stack; elements below
            // firstClip are already in the clip buffer.
            uint32_t parent = firstClip > 0 ? st.clipIDOf(batch.clips[firstClip - 1].serial) : 0;
            for (size_t i = firstClip; i < batch.clips.size(); ++i)
            {
                const ClipRef& ce = batch.clips[i];
                const uint32_t id = ++st.lastClipID;
                st.clipIDs.push_back({ce.serial, id});
                if (ce.onFrame)
                    st.draws.push_back(makeRecord(ce.path, ce.matrix, nullptr, id, true, parent, ce.fillRule));
                parent = id;
            }
            if (!batch.clips.empty())
                st.contentSerial = batch.clips.back().serial;
            st.draws.push_back(batch.draw);
            st.draws.back().paint.params = (st.draws.back().paint.params & 0xffffu) | (parent << 16);
            st.paths += paths;
            st.contours += ctrs;
            st.tess += verts;
            if (newRamp)
            {
                if (simpleRamp)
                    st.simple.push_back(simple_key(*batch.draw.gradient));
                else
                    st.complex.push_back(batch.draw.gradient);
            }
            pushed = true;
        }
        if (!pushed)
            fprintf(stderr, "RiveRenderer: a draw and its clip stack exceed one logical flush; dropped\n");
    }

    // 3. Each logical flush to the backend; later ones load the target
    //    (render_context.cpp:1218-1222).
    if (flushCount != nullptr)
        *flushCount = (uint32_t)flushes.size();
    std::vector<FlushDescriptor> descs;
    for (size_t k = 0; k < flushes.size(); ++k)
    {
        if (maxFlushes != 0 && k >= maxFlushes)
            break;
        descs.push_back(layoutFlush(flushes[k].draws, target, k == 0 ? m_frame.loadAction : 1u, nullptr));
    }
    bool ok = m_impl->flushFrame(descs);
    m_batches.clear();
    m_modulatedGradients.clear();
    m_framePaths.clear();
    m_provLastClipID = 0;
    m_provContentSerial = 0;
    m_provClipIDs.clear();
    return ok;
}

FlushDescriptor RenderContext::layoutFlush(std::vector<DrawRecord>& draws, rive_hip_target* target,
                                           uint32_t loadAction, const std::vector<rive_hip_canvas>* canvases)
{
    // LogicalFlush::layoutResources + writeResources (render_context.cpp:1116-2408)
    FlushDescriptor fd;
    fd.renderTarget = target;
    fd.loadAction = loadAction;
    fd.clearColor = m_frame.clearColor;
    fd.dither = m_frame.ditherMode;
    fd.paths.push_back(rive_hip_path_data{});
    rive_hip_paint_data clearPaint{};
    clearPaint.params = 1;
    clearPaint.value = swizzle_rgba(m_frame.clearColor);
    fd.paints.push_back(clearPaint);
    fd.paintAux.push_back(rive_hip_paint_aux{});

    struct PendingGrad
    {
        uint32_t pathID;
        bool complex;
        uint32_t row, col;
    };
    std::vector<std::pair<uint32_t, uint32_t>> simple;
    std::vector<const Gradient*> complex;
    std::vector<PendingGrad> pending;
    // Draws whose paths all have resident geometry read it in place; any
    // other flush copies every path's verbs and points.
    bool resident = !draws.empty();
    for (const DrawRecord& r : draws)
        resident = resident && r.residentVerb >= 0;
    fd.residentGeometry = resident;
    for (DrawRecord& r : draws)
    {
        const uint32_t pathID = (uint32_t)fd.paths.size();
        r.draw.path_id = pathID;
        r.draw.verb_count = (uint32_t)r.raw->verbs().size();
        r.draw.point_count = (uint32_t)r.raw->points().size();
        if (resident)
        {
            r.draw.first_verb = (uint32_t)r.residentVerb;
            r.draw.first_point = r.residentPoint;
        }
        else
        {
            r.draw.first_verb = (uint32_t)fd.verbs.size();
            r.draw.first_point = (uint32_t)(fd.points.size() / 2);
            for (PathVerb v : r.raw->verbs())
                fd.verbs.push_back((uint8_t)v);
            for (Vec2D p : r.raw->points())
            {
                fd.points.push_back(p.x);
                fd.points.push_back(p.y);
            }
        }
        if (r.gradient != nullptr)
        {
            // PaintAux paint matrix (gpu.cpp:950-985) and ramp allocation
            // (render_context.cpp:588-674).
            const Gradient* g = r.gradient;
            Mat2D view(r.path.matrix[0], r.path.matrix[1], r.path.matrix[2], r.path.matrix[3], r.path.matrix[4],
                       r.path.matrix[5]);
            Mat2D inv;
            if (!view.invert(&inv))
                inv = Mat2D();
            Mat2D pm;
            if (g->paintType == 2)
                pm = Mat2D(g->coeffs[0], 0, g->coeffs[1], 0, g->coeffs[2], 0) * inv;
            else
            {
                float w = 1 / g->coeffs[2];
                pm = Mat2D(w, 0, 0, w, -g->coeffs[0] * w, -g->coeffs[1] * w) * inv;
            }
            memcpy(r.aux.paint_matrix, pm.values(), sizeof(r.aux.paint_matrix));
            size_t n = g->stops.size();
            PendingGrad pg{pathID, false, 0, 0};
            if (n == 1 || (n == 2 && g->stops[0] == 0 && g->stops[1] == 1))
            {
                std::pair<uint32_t, uint32_t> key{g->colors[0], g->colors[std::min<size_t>(1, n - 1)]};
                auto it = std::find(simple.begin(), simple.end(), key);
                uint32_t idx;
                if (it != simple.end())
                    idx = (uint32_t)(it - simple.begin()) * 2;
                else
                {
                    idx = (uint32_t)simple.size() * 2;
                    simple.push_back(key);
                }
                pg.row = idx / 512;
                pg.col = idx % 512;
            }
            else
            {
                uint32_t row = ~0u;
                for (size_t i = 0; i < complex.size(); ++i)
                    if (complex[i]->colors == g->colors && complex[i]->stops == g->stops)
                        row = (uint32_t)i;
                if (row == ~0u)
                {
                    row = (uint32_t)complex.size();
                    complex.push_back(g);
                }
                pg.complex = true;
                pg.row = row;
            }
            pending.push_back(pg);
        }
        if (r.image != nullptr)
        {
            auto it = std::find(fd.images.begin(), fd.images.end(), r.image->handle());
            if (it == fd.images.end())
            {
                fd.images.push_back(r.image->handle());
                it = fd.images.end() - 1;
            }
            r.draw.image_index = (uint32_t)(it - fd.images.begin()) + 1;
        }
        fd.paths.push_back(r.path);
        fd.paints.push_back(r.paint);
        fd.paintAux.push_back(r.aux);
        fd.draws.push_back(r.draw);
    }
    if (canvases == nullptr)
    {
        // Atlas feathers are tessellated after the flush's other draws:
        // fills then strokes, unscissored then scissored
        // (render_context.cpp:2255-2345).  Their path IDs stay in submission
        // order; only the tessellation order moves.
        std::vector<rive_hip_draw> ordered;
        ordered.reserve(fd.draws.size());
        for (size_t i = 0; i < draws.size(); ++i)
            if (!draws[i].atlas)
                ordered.push_back(fd.draws[i]);
        for (int stroked = 0; stroked < 2; ++stroked)
            for (int scissored = 0; scissored < 2; ++scissored)
                for (size_t i = 0; i < draws.size(); ++i)
                    if (draws[i].atlas && (int)(draws[i].draw.is_stroke == 1) == stroked &&
                        (int)draws[i].atlasScissored == scissored)
                        ordered.push_back(fd.draws[i]);
        fd.draws.swap(ordered);
    }
    // Gradient spans (render_context.cpp:1462-1533).
    const uint32_t simpleRows = (uint32_t)((simple.size() * 2 + 511) / 512);
    const uint32_t complexRows = (uint32_t)complex.size();
    const uint32_t height = std::max<uint32_t>(simpleRows + complexRows, 1);
    const uint32_t ONE_TEXEL_FIXED = 65536 / 512;
    for (size_t i = 0; i < simple.size(); ++i)
    {
        uint32_t y = (uint32_t)(i / 256);
        uint32_t cx = (uint32_t)((i % 256) * 2 + 1) * ONE_TEXEL_FIXED;
        fd.gradSpans.push_back({(cx << 16) | cx, 0x80000000u | 0x40000000u | y, simple[i].first, simple[i].second});
    }
    for (uint32_t i = 0; i < complexRows; ++i)
    {
        const Gradient* g = complex[i];
        uint32_t y = i + simpleRows;
        float m = (512 - 1.f) * ONE_TEXEL_FIXED;
        float a = .5f * ONE_TEXEL_FIXED;
        uint32_t lastX = (uint32_t)(g->stops[0] * m + a);
        uint32_t lastColor = g->colors[0];
        for (size_t k = 1; k < g->stops.size(); ++k)
        {
            uint32_t x = (uint32_t)(g->stops[k] * m + a);
            uint32_t flags = 0x20000000u;
            if (k == 1)
                flags |= 0x80000000u;
            if (k == g->stops.size() - 1)
                flags |= 0x40000000u;
            fd.gradSpans.push_back({(x << 16) | lastX, flags | y, lastColor, g->colors[k]});
            lastColor = g->colors[k];
            lastX = x;
        }
    }
    const float invH = 1.f / height;
    for (const PendingGrad& pg : pending)
    {
        uint32_t row = pg.row + (pg.complex ? simpleRows : 0);
        float v = (static_cast<float>(row) + .5f) * invH;
        memcpy(&fd.paints[pg.pathID].value, &v, 4);
        float left = pg.complex ? 0.f : (float)pg.col;
        float right = pg.complex ? 512.f : left + 2;
        fd.paintAux[pg.pathID].grad_span[0] = (right - left - 1) * (1.f / 512);
        fd.paintAux[pg.pathID].grad_span[1] = (left + .5f) * (1.f / 512);
    }
    fd.gradSimpleRows = simpleRows;
    fd.gradComplexRows = complexRows;
    // Batched: each canvas owns the paths of its draws (path ID = draw index + 1).
    if (canvases != nullptr)
        fd.canvases = *canvases;
    return fd;
}
} // namespace gpu

// ---------------------------------------------------------------- HIP impl
bool RenderContextHIPImpl::upload(int kind, const void* data, size_t bytes)
{
    size_t need = std::max<size_t>(bytes, 16);
    if (m_sizes[kind] < need)
    {
        // Grow like the reference's resource allocator: round up by 1.25x.
        size_t grown = need + need / 4;
        if (rive_hip_buffer_resize(m_ctx, kind, grown) != RIVE_HIP_OK)
            return false;
        m_sizes[kind] = grown;
    }
    void* p = rive_hip_buffer_map(m_ctx, kind, bytes);
    if (p == nullptr && bytes > 0)
        return false;
    if (bytes > 0)
        memcpy(p, data, bytes);
    return rive_hip_buffer_unmap(m_ctx, kind, bytes) == RIVE_HIP_OK;
}

bool RenderContextHIPImpl::countDraws(const std::vector<gpu::DrawRecord>& draws, std::vector<uint32_t>* tess,
                                      std::vector<uint32_t>* contours)
{
    std::vector<rive_hip_draw> recs(draws.size());
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    for (size_t i = 0; i < draws.size(); ++i)
    {
        const gpu::DrawRecord& r = draws[i];
        recs[i] = r.draw;
        recs[i].path_id = (uint32_t)i + 1;
        recs[i].first_verb = (uint32_t)verbs.size();
        recs[i].verb_count = (uint32_t)r.raw->verbs().size();
        recs[i].first_point = (uint32_t)(points.size() / 2);
        recs[i].point_count = (uint32_t)r.raw->points().size();
        for (PathVerb v : r.raw->verbs())
            verbs.push_back((uint8_t)v);
        for (Vec2D p : r.raw->points())
        {
            points.push_back(p.x);
            points.push_back(p.y);
        }
    }
    if (!upload(RIVE_HIP_BUFFER_DRAW, recs.data(), recs.size() * sizeof(rive_hip_draw)) ||
        !upload(RIVE_HIP_BUFFER_VERB, verbs.data(), verbs.size()) ||
        !upload(RIVE_HIP_BUFFER_POINT, points.data(), points.size() * sizeof(float)))
        return false;
    tess->assign(draws.size(), 0);
    contours->assign(draws.size(), 0);
    return rive_hip_count_draws(m_ctx, (uint32_t)draws.size(), tess->data(), contours->data()) == RIVE_HIP_OK;
}

bool RenderContextHIPImpl::flushFrame(const std::vector<gpu::FlushDescriptor>& flushes)
{
    // One upload per buffer for the whole frame (the reference maps each
    // buffer once per frame); each logical flush then reads its own run.
    std::vector<rive_hip_path_data> paths;
    std::vector<rive_hip_paint_data> paints;
    std::vector<rive_hip_paint_aux> aux;
    std::vector<rive_hip_draw> draws;
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    std::vector<rive_hip_grad_span> grads;
    std::vector<rive_hip_flush_desc> descs;
    for (const gpu::FlushDescriptor& fd : flushes)
    {
        rive_hip_flush_desc d{};
        d.target = fd.renderTarget;
        d.canvases = fd.canvases.empty() ? nullptr : fd.canvases.data();
        d.canvas_count = (uint32_t)fd.canvases.size();
        d.mode = RIVE_HIP_FLUSH_GEOMETRY;
        d.load_action = fd.loadAction;
        d.clear_color = fd.clearColor;
        d.dither = fd.dither;
        d.path_count = (uint32_t)fd.paths.size() - 1;
        d.draw_count = (uint32_t)fd.draws.size();
        d.grad_span_count = (uint32_t)fd.gradSpans.size();
        d.grad_simple_rows = fd.gradSimpleRows;
        d.grad_complex_rows = fd.gradComplexRows;
        d.first_path_tess_location = 8;
        d.images = fd.images.empty() ? nullptr : fd.images.data();
        d.image_count = (uint32_t)fd.images.size();
        d.first_path = paths.size();
        d.first_paint = paints.size();
        d.first_paint_aux = aux.size();
        d.first_draw = draws.size();
        d.first_grad_span = grads.size();
        const uint32_t verbBase = (uint32_t)verbs.size(), pointBase = (uint32_t)(points.size() / 2);
        paths.insert(paths.end(), fd.paths.begin(), fd.paths.end());
        paints.insert(paints.end(), fd.paints.begin(), fd.paints.end());
        aux.insert(aux.end(), fd.paintAux.begin(), fd.paintAux.end());
        for (rive_hip_draw r : fd.draws)
        {
            r.first_verb += verbBase;
            r.first_point += pointBase;
            draws.push_back(r);
        }
        verbs.insert(verbs.end(), fd.verbs.begin(), fd.verbs.end());
        points.insert(points.end(), fd.points.begin(), fd.points.end());
        grads.insert(grads.end(), fd.gradSpans.begin(), fd.gradSpans.end());
        descs.push_back(d);
    }
    bool resident = !flushes.empty();
    for (const gpu::FlushDescriptor& fd : flushes)
        resident = resident && fd.residentGeometry;
    if (!upload(RIVE_HIP_BUFFER_PATH, paths.data(), paths.size() * sizeof(rive_hip_path_data)) ||
        !upload(RIVE_HIP_BUFFER_PAINT, paints.data(), paints.size() * sizeof(rive_hip_paint_data)) ||
        !upload(RIVE_HIP_BUFFER_PAINT_AUX, aux.data(), aux.size() * sizeof(rive_hip_paint_aux)) ||
        !upload(RIVE_HIP_BUFFER_DRAW, draws.data(), draws.size() * sizeof(rive_hip_draw)) ||
        (!resident && !upload(RIVE_HIP_BUFFER_VERB, verbs.data(), verbs.size())) ||
        (!resident && !upload(RIVE_HIP_BUFFER_POINT, points.data(), points.size() * sizeof(float))) ||
        !upload(RIVE_HIP_BUFFER_GRAD_SPAN, grads.data(), grads.size() * sizeof(rive_hip_grad_span)))
        return false;
    return descs.empty() || rive_hip_flush_frame(m_ctx, descs.data(), (uint32_t)descs.size()) == RIVE_HIP_OK;
}
} // namespace rive

// Replays a rive_stream through RiveRenderer (the draw-stream contract in
// include/rive_hip.h).
constexpr uint64_t kMaxContoursPerFlush = 65535; // contour IDs are 16 bits
namespace
{
struct StreamReplay
{
    std::vector<rive::RiveRenderPath> paths;
    std::vector<rive::RiveRenderPaint> paints;
    std::vector<rive_hip_image*> handles;
    std::vector<rive::RiveRenderImage> images;
    ~StreamReplay()
    {
        for (rive_hip_image* h : handles)
            rive_hip_image_destroy(h);
    }
};
} // namespace

namespace
{
// The stream's paths and images as front-end objects.
int load_stream(rive_hip_context* ctx, const rive_stream* s, StreamReplay& r)
{
    r.paths.reserve(s->path_count);
    for (uint32_t i = 0; i < s->path_count; ++i)
    {
        const rive_stream_path& sp = s->paths[i];
        rive::RiveRenderPath p((rive::FillRule)sp.fill_rule);
        const float* pts = s->points + 2 * (size_t)sp.first_point;
        uint32_t pc = 0;
        for (uint32_t k = 0; k < sp.verb_count; ++k)
        {
            switch (s->verbs[sp.first_verb + k])
            {
                case 0:
                    p.moveTo(pts[2 * pc], pts[2 * pc + 1]);
                    pc += 1;
                    break;
                case 1:
                    p.lineTo(pts[2 * pc], pts[2 * pc + 1]);
                    pc += 1;
                    break;
                case 2:
                    p.mutableRawPath().quad({pts[2 * pc], pts[2 * pc + 1]}, {pts[2 * pc + 2], pts[2 * pc + 3]});
                    pc += 2;
                    break;
                case 4:
                    p.cubicTo(pts[2 * pc], pts[2 * pc + 1], pts[2 * pc + 2], pts[2 * pc + 3], pts[2 * pc + 4],
                              pts[2 * pc + 5]);
                    pc += 3;
                    break;
                case 5:
                    p.close();
                    break;
                default:
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
            }
        }
        r.paths.push_back(std::move(p));
    }
    r.handles.reserve(s->image_count);
    r.images.reserve(s->image_count);
    for (uint32_t i = 0; i < s->image_count; ++i)
    {
        const rive_stream_image& si = s->images[i];
        rive_hip_image* h = nullptr;
        int rc = rive_hip_image_create(ctx, si.width, si.height, 0, si.rgba_premul, &h);
        if (rc != RIVE_HIP_OK)
            return rc;
        r.handles.push_back(h);
        r.images.emplace_back(h, si.width, si.height);
    }
    return RIVE_HIP_OK;
}

// Replays the stream's draws through `renderer` (whose frame or canvas is
// already begun).  The renderer, like the paths and paints, must outlive the
// flush: image draws reference its unit rect.
int draw_stream(rive::RiveRenderer& renderer, const rive_stream* s, StreamReplay& r)
{
    r.paints.resize(s->draw_count);
    for (uint32_t i = 0; i < s->draw_count; ++i)
    {
        const rive_stream_draw& d = s->draws[i];
        rive::RiveRenderPaint& paint = r.paints[i];
        paint.style(d.style == 1 ? rive::RenderPaintStyle::stroke : rive::RenderPaintStyle::fill);
        paint.thickness(d.thickness);
        paint.join((rive::StrokeJoin)d.join);
        paint.cap((rive::StrokeCap)d.cap);
        paint.feather(d.feather);
        paint.blendMode((rive::BlendMode)d.blend_mode);
        if (d.paint_type == 0)
            paint.color(d.color);
        else
        {
            const uint32_t* colors = s->stop_colors + d.first_stop;
            const float* stops = s->stop_offsets + d.first_stop;
            auto g = d.paint_type == 1
                         ? rive::Gradient::MakeLinear(d.grad[0], d.grad[1], d.grad[2], d.grad[3], colors, stops,
                                                      d.stop_count)
                         : rive::Gradient::MakeRadial(d.grad[0], d.grad[1], d.grad[2], colors, stops, d.stop_count);
            if (g)
                paint.shader(g);
            else
                paint.color(0);
        }
        renderer.save();
        renderer.transform(rive::Mat2D(d.matrix[0], d.matrix[1], d.matrix[2], d.matrix[3], d.matrix[4], d.matrix[5]));
        for (uint32_t c = 0; c < d.clip_count && c < 2; ++c)
            renderer.clipPath(&r.paths[d.clip_paths[c]]);
        if (d.image != 0)
        {
            if (d.image > r.images.size())
                return RIVE_HIP_ERROR_INVALID_ARGUMENT;
            renderer.drawImage(&r.images[d.image - 1], rive::ImageSampler::FromKey((uint8_t)d.image_sampler),
                               (rive::BlendMode)d.blend_mode, d.opacity);
        }
        else
            renderer.drawPath(&r.paths[d.path], &paint);
        renderer.restore();
    }
    return RIVE_HIP_OK;
}

rive::gpu::RenderContext::FrameDescriptor stream_frame(const rive_stream* s)
{
    rive::gpu::RenderContext::FrameDescriptor fd;
    fd.renderTargetWidth = s->width;
    fd.renderTargetHeight = s->height;
    fd.clearColor = s->clear_color;
    fd.ditherMode = s->dither;
    return fd;
}
} // namespace

extern "C" int rive_hip_render_stream_ex(rive_hip_context* ctx, const rive_stream* s, rive_hip_target* target,
                                         uint32_t maxFlushes, uint32_t* flushCount)
{
    if (ctx == nullptr || s == nullptr || target == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    StreamReplay r;
    int st = load_stream(ctx, s, r);
    if (st != RIVE_HIP_OK)
        return st;
    rive::gpu::RenderContext rc(std::make_unique<rive::RenderContextHIPImpl>(ctx));
    rc.beginFrame(stream_frame(s));
    rive::RiveRenderer renderer(&rc);
    st = draw_stream(renderer, s, r);
    if (st != RIVE_HIP_OK)
        return st;
    return rc.flush(target, maxFlushes, flushCount) ? RIVE_HIP_OK : RIVE_HIP_ERROR_DEVICE;
}

extern "C" int rive_hip_render_stream(rive_hip_context* ctx, const rive_stream* s, rive_hip_target* target)
{
    return rive_hip_render_stream_ex(ctx, s, target, 0, nullptr);
}

extern "C" uint32_t rive_stream_contour_bound(const rive_stream* s)
{
    if (s == nullptr)
        return 0;
    auto moves = [s](uint32_t path) {
        if (path >= s->path_count)
            return 0u;
        const rive_stream_path& sp = s->paths[path];
        uint32_t m = 0;
        for (uint32_t k = 0; k < sp.verb_count; ++k)
            m += s->verbs[sp.first_verb + k] == 0 ? 1u : 0u;
        return m;
    };
    uint64_t total = 0;
    for (uint32_t i = 0; i < s->draw_count; ++i)
    {
        const rive_stream_draw& d = s->draws[i];
        total += d.image != 0 ? 1u : moves(d.path);
        for (uint32_t c = 0; c < d.clip_count && c < 2; ++c)
            total += moves(d.clip_paths[c]);
    }
    return (uint32_t)std::min<uint64_t>(total, 0xffffffffu);
}

extern "C" int rive_hip_render_streams(rive_hip_context* ctx, const rive_stream* const* streams,
                                       rive_hip_target* const* targets, uint32_t n)
{
    if (ctx == nullptr || streams == nullptr || targets == nullptr || n == 0)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i)
        if (streams[i] == nullptr || targets[i] == nullptr)
            return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    // Consecutive canvases share a flush while their contours fit the 16-bit
    // contour IDs (as the reference starts a new logical flush when a
    // resource limit is reached, render_context.cpp:1012-1016).
    uint32_t begin = 0;
    while (begin < n)
    {
        uint64_t contours = rive_stream_contour_bound(streams[begin]);
        uint32_t end = begin + 1;
        while (end < n && contours + rive_stream_contour_bound(streams[end]) <= kMaxContoursPerFlush)
            contours += rive_stream_contour_bound(streams[end++]);
        std::vector<StreamReplay> replays(end - begin);
        std::vector<std::unique_ptr<rive::RiveRenderer>> renderers;
        rive::gpu::RenderContext rc(std::make_unique<rive::RenderContextHIPImpl>(ctx));
        rc.beginFrame(stream_frame(streams[begin]));
        for (uint32_t i = begin; i < end; ++i)
        {
            int st = load_stream(ctx, streams[i], replays[i - begin]);
            if (st != RIVE_HIP_OK)
                return st;
            rc.beginCanvas(stream_frame(streams[i]), targets[i]);
            renderers.push_back(std::make_unique<rive::RiveRenderer>(&rc));
            st = draw_stream(*renderers.back(), streams[i], replays[i - begin]);
            if (st != RIVE_HIP_OK)
                return st;
        }
        if (!rc.flush(targets[begin]))
            return RIVE_HIP_ERROR_DEVICE;
        begin = end;
    }
    return RIVE_HIP_OK;
}

namespace rive
{
int ResidentSession::init(rive_hip_context* ctx, const std::vector<ResidentScene>& scenes)
{
    m_ctx = ctx;
    m_paths.assign(scenes.size(), {});
    m_paints.assign(scenes.size(), {});
    m_contours.assign(scenes.size(), 0);
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    for (size_t s = 0; s < scenes.size(); ++s)
    {
        const ResidentScene& sc = scenes[s];
        m_paths[s].reserve(sc.pathCount);
        m_paints[s].resize(sc.pathCount);
        for (uint32_t i = 0; i < sc.pathCount; ++i)
        {
            const rive_stream_path& sp = sc.paths[i];
            RiveRenderPath p((FillRule)sp.fill_rule);
            const float* pts = sc.points + 2 * (size_t)sp.first_point;
            uint32_t pc = 0;
            for (uint32_t k = 0; k < sp.verb_count; ++k)
            {
                switch (sc.verbs[sp.first_verb + k])
                {
                    case 0:
                        p.moveTo(pts[2 * pc], pts[2 * pc + 1]);
                        pc += 1;
                        ++m_contours[s];
                        break;
                    case 1:
                        p.lineTo(pts[2 * pc], pts[2 * pc + 1]);
                        pc += 1;
                        break;
                    case 4:
                        p.cubicTo(pts[2 * pc], pts[2 * pc + 1], pts[2 * pc + 2], pts[2 * pc + 3], pts[2 * pc + 4],
                                  pts[2 * pc + 5]);
                        pc += 3;
                        break;
                    case 5:
                        p.close();
                        break;
                    default:
                        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                }
            }
            p.setResident((uint32_t)verbs.size(), (uint32_t)(points.size() / 2));
            for (PathVerb v : p.getRawPath().verbs())
                verbs.push_back((uint8_t)v);
            for (Vec2D q : p.getRawPath().points())
            {
                points.push_back(q.x);
                points.push_back(q.y);
            }
            m_paths[s].push_back(std::move(p));
            m_paints[s][i].color(sc.colors[i]);
        }
    }
    // the resident geometry: one upload each for the whole session
    auto put = [&](int kind, const void* data, size_t bytes) -> int {
        const size_t need = std::max<size_t>(bytes, 16);
        int st = rive_hip_buffer_resize(ctx, kind, need);
        if (st != RIVE_HIP_OK)
            return st;
        void* p = rive_hip_buffer_map(ctx, kind, bytes);
        if (p == nullptr && bytes > 0)
            return RIVE_HIP_ERROR_DEVICE;
        if (bytes > 0)
            memcpy(p, data, bytes);
        return rive_hip_buffer_unmap(ctx, kind, bytes);
    };
    int st = put(RIVE_HIP_BUFFER_VERB, verbs.data(), verbs.size());
    if (st == RIVE_HIP_OK)
        st = put(RIVE_HIP_BUFFER_POINT, points.data(), points.size() * sizeof(float));
    return st;
}

int ResidentSession::render(const std::vector<SceneCanvas>& canvases)
{
    std::vector<gpu::FlushDescriptor> flushes;
    int st = prepare(canvases, &flushes);
    return st != RIVE_HIP_OK ? st : submit(flushes);
}

int ResidentSession::submit(const std::vector<gpu::FlushDescriptor>& flushes)
{
    RenderContextHIPImpl impl(m_ctx);
    for (const gpu::FlushDescriptor& fd : flushes)
        if (!impl.flushFrame({fd}))
            return RIVE_HIP_ERROR_DEVICE;
    return RIVE_HIP_OK;
}

int ResidentSession::prepare(const std::vector<SceneCanvas>& canvases, std::vector<gpu::FlushDescriptor>* out)
{
    out->clear();
    size_t begin = 0;
    while (begin < canvases.size())
    {
        uint64_t contours = m_contours[canvases[begin].scene];
        size_t end = begin + 1;
        while (end < canvases.size() && contours + m_contours[canvases[end].scene] <= kMaxContoursPerFlush)
            contours += m_contours[canvases[end++].scene];
        gpu::RenderContext rc(std::make_unique<RenderContextHIPImpl>(m_ctx));
        std::vector<std::unique_ptr<RiveRenderer>> renderers;
        for (size_t c = begin; c < end; ++c)
        {
            const SceneCanvas& cv = canvases[c];
            gpu::RenderContext::FrameDescriptor fd;
            fd.renderTargetWidth = cv.width;
            fd.renderTargetHeight = cv.height;
            fd.clearColor = 0;
            fd.ditherMode = 1;
            if (c == begin)
                rc.beginFrame(fd);
            rc.beginCanvas(fd, cv.target);
            renderers.push_back(std::make_unique<RiveRenderer>(&rc));
            RiveRenderer& r = *renderers.back();
            std::vector<RiveRenderPath>& paths = m_paths[cv.scene];
            for (size_t i = 0; i < paths.size(); ++i)
            {
                const float* m = cv.matrices + 6 * i;
                r.save();
                r.transform(Mat2D(m[0], m[1], m[2], m[3], m[4], m[5]));
                r.drawPath(&paths[i], &m_paints[cv.scene][i]);
                r.restore();
            }
        }
        out->emplace_back();
        if (!rc.prepareCanvases(canvases[begin].target, &out->back()))
            return RIVE_HIP_ERROR_INVALID_ARGUMENT;
        begin = end;
    }
    return RIVE_HIP_OK;
}
} // namespace rive
