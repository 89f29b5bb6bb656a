// Clean-room C++ front end with the reference's API names:
//   rive::Renderer / RiveRenderer          include/rive/renderer.hpp:209-236,
//                                          renderer/include/rive/renderer/rive_renderer.hpp:26-144
//   rive::RiveRenderPath / RiveRenderPaint  renderer/src/rive_render_path.hpp:18-120,
//                                          renderer/src/rive_render_paint.hpp:19-86
//   rive::gpu::RenderContext               renderer/include/rive/renderer/render_context.hpp:87-1057
//   rive::gpu::RenderContextImpl           renderer/include/rive/renderer/render_context_impl.hpp:26-243
// Draw streams written against these names run unchanged; RenderContext hands
// each flush to a RenderContextImpl, and RenderContextHIPImpl sends it to the
// gfx950 backend through include/rive_hip.h.
#pragma once

#include "../../include/rive_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

namespace rive
{
struct Vec2D
{
    float x = 0, y = 0;
};
inline Vec2D operator-(Vec2D a, Vec2D b) { return {a.x - b.x, a.y - b.y}; }

// [xx, xy, yx, yy, tx, ty]: x' = xx*x + yx*y + tx, y' = xy*x + yy*y + ty.
class Mat2D
{
public:
    Mat2D() = default;
    Mat2D(float xx, float xy, float yx, float yy, float tx, float ty) : m{xx, xy, yx, yy, tx, ty} {}
    const float* values() const { return m; }
    float operator[](int i) const { return m[i]; }
    Vec2D map(Vec2D p) const { return {m[0] * p.x + m[2] * p.y + m[4], m[1] * p.x + m[3] * p.y + m[5]}; }
    Mat2D operator*(const Mat2D& b) const;
    bool invert(Mat2D* out) const;
    float findMaxScale() const; // src/math/mat2d_find_max_scale.cpp:24-67
    static Mat2D fromRotation(float radians)
    {
        const float s = sinf(radians), c = cosf(radians);
        return Mat2D(c, s, -s, c, 0, 0);
    }
    bool operator==(const Mat2D& o) const
    {
        for (int i = 0; i < 6; ++i)
            if (m[i] != o.m[i])
                return false;
        return true;
    }

private:
    float m[6] = {1, 0, 0, 1, 0, 0};
};

enum class FillRule : uint32_t { nonZero, evenOdd, clockwise };
enum class StrokeJoin : uint32_t { miter = 0, round = 1, bevel = 2 };
enum class StrokeCap : uint32_t { butt = 0, round = 1, square = 2 };
enum class RenderPaintStyle : uint32_t { stroke, fill };
// include/rive/shapes/paint/blend_mode.hpp:5-23
enum class BlendMode : uint8_t
{
    srcOver = 3,
    screen = 14,
    overlay = 15,
    darken = 16,
    lighten = 17,
    colorDodge = 18,
    colorBurn = 19,
    hardLight = 20,
    softLight = 21,
    difference = 22,
    exclusion = 23,
    multiply = 24,
    hue = 25,
    saturation = 26,
    color = 27,
    luminosity = 28
};
using ColorInt = uint32_t;

enum class PathVerb : uint8_t { move = 0, line = 1, quad = 2, cubic = 4, close = 5 };

// Verb + point storage (include/rive/math/raw_path.hpp:25-297).
class RawPath
{
public:
    void move(Vec2D p);
    void line(Vec2D p);
    void cubic(Vec2D a, Vec2D b, Vec2D c);
    void quad(Vec2D a, Vec2D b); // stored as the equivalent cubic
    void close();
    bool empty() const { return m_verbs.empty(); }
    const std::vector<PathVerb>& verbs() const { return m_verbs; }
    const std::vector<Vec2D>& points() const { return m_points; }

private:
    void injectImplicitMoveIfNeeded();
    std::vector<PathVerb> m_verbs;
    std::vector<Vec2D> m_points;
    bool m_contourIsOpen = false;
    size_t m_lastMoveIdx = 0;
};

class RenderShader
{
public:
    virtual ~RenderShader() = default;
};

// Linear and radial gradients (renderer/src/gradient.cpp:38-175).
class Gradient : public RenderShader, public std::enable_shared_from_this<Gradient>
{
public:
    static std::shared_ptr<Gradient> MakeLinear(float sx, float sy, float ex, float ey, const ColorInt colors[],
                                                const float stops[], size_t count);
    static std::shared_ptr<Gradient> MakeRadial(float cx, float cy, float radius, const ColorInt colors[],
                                                const float stops[], size_t count);
    // gradient.cpp:191-220: this gradient with every stop's alpha scaled by
    // opacity (itself when opacity is 1); the last modulated copy is cached.
    std::shared_ptr<Gradient> getModulated(float opacity) const;
    size_t count() const { return colors.size(); }
    uint32_t paintType = 0; // 2 linear, 3 radial
    std::vector<ColorInt> colors;
    std::vector<float> stops;
    float coeffs[3] = {0, 0, 0};

private:
    mutable float m_lastModulatedOpacity = -1.f;
    mutable std::shared_ptr<Gradient> m_lastModulated;
};

// colorModulateOpacity (src/shapes/paint/color.cpp:46-67): the color with
// its alpha scaled by opacity, rounded back to 8 bits.
ColorInt colorModulateOpacity(ColorInt color, float opacity);
inline uint32_t colorAlpha(ColorInt c) { return c >> 24; }

class RenderPath
{
public:
    virtual ~RenderPath() = default;
};

// include/rive/shapes/paint/image_sampler.hpp:7-61
enum class ImageFilter : uint8_t { bilinear = 0, nearest = 1 };
enum class ImageWrap : uint8_t { clamp = 0, repeat = 1, mirror = 2 };
struct ImageSampler
{
    ImageWrap wrapX = ImageWrap::clamp;
    ImageWrap wrapY = ImageWrap::clamp;
    ImageFilter filter = ImageFilter::bilinear;
    uint8_t asKey() const { return (uint8_t)((int)wrapX + (int)wrapY * 3 + (int)filter * 9); }
    static ImageSampler FromKey(uint8_t key)
    {
        ImageSampler s;
        s.wrapX = (ImageWrap)(key % 3);
        s.wrapY = (ImageWrap)((key / 3) % 3);
        s.filter = (ImageFilter)((key / 9) % 2);
        return s;
    }
};

class RenderImage
{
public:
    virtual ~RenderImage() = default;
};

// A decoded, premultiplied RGBA8 image resident in HBM
// (renderer/src/rive_render_image.hpp, gpu::Texture).
class RiveRenderImage : public RenderImage
{
public:
    // `owned`: the image is destroyed with this object (a decoded image);
    // otherwise the caller keeps it.
    RiveRenderImage(rive_hip_image* image, uint32_t w, uint32_t h, bool owned = false) :
        m_image(image), m_w(w), m_h(h), m_owned(owned)
    {}
    ~RiveRenderImage() override
    {
        if (m_owned)
            rive_hip_image_destroy(m_image);
    }
    RiveRenderImage(RiveRenderImage&& o) noexcept : m_image(o.m_image), m_w(o.m_w), m_h(o.m_h), m_owned(o.m_owned)
    {
        o.m_owned = false;
    }
    RiveRenderImage(const RiveRenderImage&) = delete;
    RiveRenderImage& operator=(const RiveRenderImage&) = delete;
    rive_hip_image* handle() const { return m_image; }
    uint32_t width() const { return m_w; }
    uint32_t height() const { return m_h; }

private:
    rive_hip_image* m_image;
    uint32_t m_w, m_h;
    bool m_owned;
};
class RenderPaint
{
public:
    virtual ~RenderPaint() = default;
};

class RiveRenderPath : public RenderPath
{
public:
    explicit RiveRenderPath(FillRule rule = FillRule::nonZero) : m_fillRule(rule) {}
    RiveRenderPath(RawPath&& raw, FillRule rule) : m_raw(std::move(raw)), m_fillRule(rule) {}
    void moveTo(float x, float y) { m_raw.move({x, y}); }
    void lineTo(float x, float y) { m_raw.line({x, y}); }
    void cubicTo(float ox, float oy, float ix, float iy, float x, float y)
    {
        m_raw.cubic({ox, oy}, {ix, iy}, {x, y});
    }
    void close() { m_raw.close(); }
    void fillRule(FillRule r) { m_fillRule = r; }
    FillRule getFillRule() const { return m_fillRule; }
    const RawPath& getRawPath() const { return m_raw; }
    RawPath& mutableRawPath() { return m_raw; }
    // Where this path's verbs and points already sit in the context's VERB
    // and POINT buffers (uploaded once and kept resident); -1 when they are
    // copied into every flush.
    void setResident(uint32_t verbBase, uint32_t pointBase)
    {
        m_residentVerb = verbBase;
        m_residentPoint = pointBase;
    }
    int64_t residentVerb() const { return m_residentVerb; }
    uint32_t residentPoint() const { return m_residentPoint; }
    // A copy chopped wherever its curves turn by an angle set by the feather's
    // size on screen (rive_render_path.cpp:323-449).
    std::shared_ptr<RiveRenderPath> makeSoftenedCopyForFeathering(float feather, float matrixMaxScale) const;

private:
    RawPath m_raw;
    FillRule m_fillRule;
    int64_t m_residentVerb = -1;
    uint32_t m_residentPoint = 0;
};

class RiveRenderPaint : public RenderPaint
{
public:
    void style(RenderPaintStyle s) { m_stroked = s == RenderPaintStyle::stroke; }
    void color(ColorInt c)
    {
        m_color = c;
        m_shader.reset();
    }
    void thickness(float t) { m_thickness = t; }
    void join(StrokeJoin j) { m_join = j; }
    void cap(StrokeCap c) { m_cap = c; }
    void feather(float f) { m_feather = f; }
    void blendMode(BlendMode b) { m_blend = b; }
    void shader(std::shared_ptr<Gradient> g) { m_shader = std::move(g); }
    void shader(std::shared_ptr<RenderShader> s) { m_shader = std::dynamic_pointer_cast<Gradient>(s); }
    // RiveRenderPaint::image (rive_render_paint.cpp:58-64): white at `opacity`
    // modulated by the image.
    void image(const RiveRenderImage* img, float opacity)
    {
        m_shader.reset();
        float a = std::max(0.f, std::min(1.f, opacity));
        m_color = ((uint32_t)std::lround(255.f * a) << 24) | 0x00ffffffu;
        m_image = img;
    }
    void imageSampler(ImageSampler s) { m_sampler = s; }
    const RiveRenderImage* getImage() const { return m_image; }
    ImageSampler getImageSampler() const { return m_sampler; }
    bool getIsStroked() const { return m_stroked; }
    ColorInt getColor() const { return m_color; }
    float getThickness() const { return m_thickness; }
    StrokeJoin getJoin() const { return m_join; }
    StrokeCap getCap() const { return m_cap; }
    float getFeather() const { return m_feather; }
    BlendMode getBlendMode() const { return m_blend; }
    const Gradient* getGradient() const { return m_shader.get(); }

private:
    bool m_stroked = false;
    ColorInt m_color = 0xff000000;
    float m_thickness = 1;
    StrokeJoin m_join = StrokeJoin::miter;
    StrokeCap m_cap = StrokeCap::butt;
    float m_feather = 0;
    BlendMode m_blend = BlendMode::srcOver;
    std::shared_ptr<Gradient> m_shader;
    const RiveRenderImage* m_image = nullptr;
    ImageSampler m_sampler;
};

// A GPU buffer handle (renderer.hpp RenderBuffer), for drawImageMesh.
enum class RenderBufferType : uint8_t { index, vertex };
class RenderBuffer
{
public:
    virtual ~RenderBuffer() = default;
};

class Renderer
{
public:
    virtual ~Renderer() = default;
    virtual void save() = 0;
    virtual void restore() = 0;
    virtual void transform(const Mat2D& m) = 0;
    virtual void drawPath(RenderPath* path, RenderPaint* paint) = 0;
    virtual void clipPath(RenderPath* path) = 0;
    virtual void drawImage(const RenderImage* image, ImageSampler sampler, BlendMode blendMode, float opacity) = 0;
    virtual void drawImageMesh(const RenderImage* image, ImageSampler sampler,
                               std::shared_ptr<RenderBuffer> vertices_f32, std::shared_ptr<RenderBuffer> uvCoords_f32,
                               std::shared_ptr<RenderBuffer> indices_u16, uint32_t vertexCount, uint32_t indexCount,
                               BlendMode blendMode, float opacity) = 0;
    // renderer.hpp:231-235: scales the opacity of later draws; stacked
    // multiplicatively, captured by save() and restored by restore().
    virtual void modulateOpacity(float opacity) = 0;
    // renderer.cpp: conveniences over transform().
    void translate(float x, float y) { transform(Mat2D(1, 0, 0, 1, x, y)); }
    void scale(float sx, float sy) { transform(Mat2D(sx, 0, 0, sy, 0, 0)); }
    void rotate(float radians) { transform(Mat2D::fromRotation(radians)); }
};

// A render buffer in host memory (the front end reads image-mesh buffers on
// the host; nothing here draws from them).
class HostRenderBuffer : public RenderBuffer
{
public:
    HostRenderBuffer(RenderBufferType type, size_t bytes) : m_type(type), m_data(bytes) {}
    RenderBufferType type() const { return m_type; }
    uint8_t* data() { return m_data.data(); }
    size_t size() const { return m_data.size(); }

private:
    RenderBufferType m_type;
    std::vector<uint8_t> m_data;
};

// The factory draw streams create their objects with
// (include/rive/factory.hpp:31-69; the reference's RenderContext is one).
class Factory
{
public:
    virtual ~Factory() = default;
    virtual std::shared_ptr<RenderBuffer> makeRenderBuffer(RenderBufferType type, uint32_t flags, size_t bytes);
    virtual std::shared_ptr<RenderShader> makeLinearGradient(float sx, float sy, float ex, float ey,
                                                             const ColorInt colors[], const float stops[],
                                                             size_t count);
    virtual std::shared_ptr<RenderShader> makeRadialGradient(float cx, float cy, float radius,
                                                             const ColorInt colors[], const float stops[],
                                                             size_t count);
    // Takes the path's verbs and points (the RawPath is left empty).
    virtual std::shared_ptr<RiveRenderPath> makeRenderPath(RawPath& path, FillRule rule);
    virtual std::shared_ptr<RiveRenderPath> makeEmptyRenderPath();
    virtual std::shared_ptr<RiveRenderPaint> makeRenderPaint();
    // An encoded image (PNG or JPEG) as a RenderImage; null when it cannot be
    // decoded.  The base factory has no device: always null.
    virtual std::shared_ptr<RenderImage> decodeImage(const uint8_t* bytes, size_t size);
};

namespace gpu
{
// One PathDraw's records (PathData/PaintData/PaintAux + counting inputs).
// The clip ID a draw reads is patched into paint.params when its logical
// flush is laid out.
struct DrawRecord
{
    rive_hip_path_data path;
    rive_hip_paint_data paint;
    rive_hip_paint_aux aux;
    rive_hip_draw draw;
    const RawPath* raw;
    const Gradient* gradient; // null for solid/clip paints
    const RiveRenderImage* image;
    int64_t residentVerb = -1; // the path's resident geometry, if any
    uint32_t residentPoint = 0;
    // A feather the reference draws through its offscreen atlas
    // (draw.cpp:387-418): tessellated one way, after the flush's other draws;
    // scissored when its pixel bounds leave the render target.
    bool atlas = false, atlasScissored = false;
};

// One element of the clip stack a draw sits under (rive_renderer.hpp
// ClipElement): the serial names the element; equal serials are the same
// clip content.
struct ClipRef
{
    uint64_t serial;
    const RiveRenderPath* path;
    Mat2D matrix;
    FillRule fillRule;
    bool onFrame; // its pixel bounds meet the frame (else no draw, only an ID)
};

// The flush descriptor handed to the backend (gpu.hpp:1320-1438, reduced).
struct FlushDescriptor
{
    rive_hip_target* renderTarget = nullptr;
    uint32_t loadAction = 0;
    ColorInt clearColor = 0;
    uint32_t dither = 1;
    std::vector<rive_hip_path_data> paths; // [0] reserved
    std::vector<rive_hip_paint_data> paints;
    std::vector<rive_hip_paint_aux> paintAux;
    std::vector<rive_hip_draw> draws;
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    std::vector<rive_hip_grad_span> gradSpans;
    uint32_t gradSimpleRows = 0, gradComplexRows = 0;
    std::vector<rive_hip_image*> images;
    std::vector<rive_hip_canvas> canvases; // batched flush; empty: renderTarget
    bool residentGeometry = false;         // draws read resident VERB/POINT data; verbs/points empty
};

// Backend plugin interface (render_context_impl.hpp:26-243, reduced to what
// the GEOMETRY flush needs).
class RenderContextImpl
{
public:
    virtual ~RenderContextImpl() = default;
    // One frame: its logical flushes in order.  The reference maps the
    // frame's buffers once and calls flush(desc) per logical flush, each
    // descriptor naming its records by first* offsets
    // (render_context.cpp:971-1018); a backend sees the same here.
    virtual bool flushFrame(const std::vector<FlushDescriptor>& flushes) = 0;
    // The resource counts of each draw (Draw::resourceCounts, draw.cpp:
    // 1378-1391): tessellation vertices and contours.  A backend that
    // counts on the device answers here; false if it cannot.
    virtual bool countDraws(const std::vector<DrawRecord>& draws, std::vector<uint32_t>* tess,
                            std::vector<uint32_t>* contours) = 0;
    // An encoded image decoded into a texture with its mip chain
    // (RenderContext::decodeImage's Bitmap::decode + makeImageTexture,
    // render_context.cpp:237-262); null when the backend cannot.
    virtual std::shared_ptr<RiveRenderImage> decodeImageTexture(const uint8_t*, size_t) { return nullptr; }
    // Whether large plain fills are laid out for interior triangulation
    // (FrameDescriptor::triangulationThresholds with a budget above zero);
    // off by default (frameBudgetMs = 0).
    virtual bool interiorTriangulation() const { return false; }
};

class RenderContext : public Factory
{
public:
    struct FrameDescriptor // render_context.hpp:102-144
    {
        uint32_t renderTargetWidth = 0, renderTargetHeight = 0;
        uint32_t loadAction = 0; // 0 clear, 1 preserve
        ColorInt clearColor = 0;
        uint32_t ditherMode = 1;
    };
    explicit RenderContext(std::unique_ptr<RenderContextImpl> impl) : m_impl(std::move(impl)) {}
    void beginFrame(const FrameDescriptor& fd);
    // Starts the next canvas of a batched flush: draws from here on render
    // into `target` with this frame's size and clear color.  Each canvas
    // starts with an empty clip buffer.
    void beginCanvas(const FrameDescriptor& fd, rive_hip_target* target);
    // Lays the frame out in logical flushes and hands each to the backend.
    // maxFlushes > 0 stops after that many (a debugging aid); flushCount, if
    // given, receives the frame's number of logical flushes.
    bool flush(rive_hip_target* target, uint32_t maxFlushes = 0, uint32_t* flushCount = nullptr);
    // The batched-canvas flush's host half: lays every canvas begun since
    // beginFrame out into one FlushDescriptor (no device work) and clears the
    // frame; RenderContextImpl::flushFrame then submits it.  False when no
    // canvas was begun.
    bool prepareCanvases(rive_hip_target* target, FlushDescriptor* out);
    const FrameDescriptor& frameDescriptor() const { return m_frame; }
    RenderContextImpl* impl() { return m_impl.get(); }
    // render_context.cpp:237-262: the backend decodes the image.
    std::shared_ptr<RenderImage> decodeImage(const uint8_t* bytes, size_t size) override
    {
        return m_impl->decodeImageTexture(bytes, size);
    }

    // The clip ID whose content the clip buffer holds after the draws pushed
    // so far (RenderContext::getClipContentID, render_context.hpp), as laid
    // out while the frame fits one logical flush; 0 when none.
    uint32_t getClipContentID() const;

    // Used by RiveRenderer.
    uint64_t newClipSerial() { return ++m_lastClipSerial; }
    // The records of one draw (PathDraw::Make + the PathData/PaintData/
    // PaintAuxData it writes, draw.cpp:420-825, gpu.cpp:859-1063).
    // opacity: the renderer's modulated opacity, applied to the paint's
    // color or gradient stops (draw.cpp:727-741, gradient.cpp:191-220).
    DrawRecord makeRecord(const RiveRenderPath* path, const Mat2D& m, const RiveRenderPaint* paint, uint32_t clipID,
                          bool clipUpdate, uint32_t outerClipID, FillRule fillRule,
                          const Mat2D* clipRectInverse = nullptr, float opacity = 1.f);
    // Keeps a path the frame's draws point at (a softened copy) alive until
    // the frame is flushed.
    RiveRenderPath* keepForFrame(std::shared_ptr<RiveRenderPath> path)
    {
        m_framePaths.push_back(std::move(path));
        return m_framePaths.back().get();
    }
    // A draw and the clip stack it sits under; clip draws and clip IDs are
    // assigned per logical flush when the frame is flushed (applyClip,
    // rive_renderer.cpp:631-824).
    void pushDrawBatch(std::vector<ClipRef> clips, const DrawRecord& draw);

private:
    struct Batch
    {
        std::vector<ClipRef> clips;
        DrawRecord draw;
    };
    FlushDescriptor layoutFlush(std::vector<DrawRecord>& draws, rive_hip_target* target, uint32_t loadAction,
                                const std::vector<rive_hip_canvas>* canvases);
    std::unique_ptr<RenderContextImpl> m_impl;
    FrameDescriptor m_frame;
    std::vector<Batch> m_batches;
    struct CanvasRecord
    {
        rive_hip_target* target;
        size_t firstBatch;
        FrameDescriptor frame;
    };
    std::vector<CanvasRecord> m_canvases;
    uint64_t m_lastClipSerial = 0;
    // Opacity-modulated copies of gradients and softened copies of feathered
    // paths, alive until the frame is flushed.
    std::vector<std::shared_ptr<Gradient>> m_modulatedGradients;
    std::vector<std::shared_ptr<RiveRenderPath>> m_framePaths;
    // applyClip's bookkeeping as draws arrive (for getClipContentID): clip
    // IDs handed out so far and the element whose content is current.
    uint32_t m_provLastClipID = 0;
    uint64_t m_provContentSerial = 0;
    std::vector<std::pair<uint64_t, uint32_t>> m_provClipIDs;
};
} // namespace gpu

class RiveRenderer : public Renderer
{
public:
    explicit RiveRenderer(gpu::RenderContext* ctx);
    void save() override;
    void restore() override;
    void transform(const Mat2D& m) override;
    void drawPath(RenderPath* path, RenderPaint* paint) override;
    void clipPath(RenderPath* path) override;
    void drawImage(const RenderImage* image, ImageSampler sampler, BlendMode blendMode, float opacity) override;
    // Image meshes (the reference's ImageMeshDraw) are not drawn by this
    // backend: the call is reported on stderr and counted.
    void drawImageMesh(const RenderImage* image, ImageSampler sampler, std::shared_ptr<RenderBuffer> vertices_f32,
                       std::shared_ptr<RenderBuffer> uvCoords_f32, std::shared_ptr<RenderBuffer> indices_u16,
                       uint32_t vertexCount, uint32_t indexCount, BlendMode blendMode, float opacity) override;
    void modulateOpacity(float opacity) override;
    float currentModulatedOpacity() const { return m_stack.back().modulatedOpacity; }
    uint32_t unsupportedDraws() const { return m_unsupportedDraws; }

    // The current clip rect (rive_renderer.hpp: hasClipRect, getClipRect,
    // getClipRectMatrix): left, top, right, bottom in the space of its matrix.
    struct AABB
    {
        float left, top, right, bottom;
    };
    // rive_renderer.cpp:16-55: is the path an axis-aligned rectangle (a move
    // and three lines, any further points equal to the first)?  rect gets
    // left, top, right, bottom.
    static bool IsAABB(const RawPath& raw, float rect[4]);
    bool hasClipRect() const { return m_stack.back().hasClipRect; }
    AABB getClipRect() const
    {
        const float* r = m_stack.back().clipRect;
        return {r[0], r[1], r[2], r[3]};
    }
    const Mat2D& getClipRectMatrix() const { return m_stack.back().clipRectMatrix; }

private:
    RiveRenderPath m_unitRectPath;
    struct IAABB
    {
        int l = 0, t = 0, r = 0, b = 0;
        bool empty() const { return l >= r || t >= b; }
        IAABB intersect(const IAABB& o) const;
    };
    struct ClipElement
    {
        Mat2D matrix;
        const RiveRenderPath* path;
        FillRule fillRule;
        IAABB pixelBounds;
        uint64_t serial;
    };
    struct RenderState
    {
        Mat2D matrix;
        size_t clipStackHeight = 0;
        IAABB overallClipPixelBounds;
        // rive_renderer.hpp RenderState clip rect: the rect (l, t, r, b) in
        // the space of clipRectMatrix, and ClipRectInverseMatrix.
        bool hasClipRect = false;
        float clipRect[4] = {0, 0, 0, 0};
        Mat2D clipRectMatrix;
        Mat2D clipRectInverse;
        float modulatedOpacity = 1.f;
    };
    uint32_t m_unsupportedDraws = 0;
    void clipRectImpl(float rect[4], RiveRenderPath* path);
    void clipPathImpl(RiveRenderPath* path);
    static IAABB MapBounds(const Mat2D& m, const RawPath& raw, float outsetX, float outsetY);
    gpu::RenderContext* m_context;
    std::vector<RenderState> m_stack;
    std::vector<ClipElement> m_clipStack;
};

// Backend that sends flushes to the gfx950 device (RIVE_HIP_FLUSH_GEOMETRY).
class RenderContextHIPImpl : public gpu::RenderContextImpl
{
public:
    explicit RenderContextHIPImpl(rive_hip_context* ctx) : m_ctx(ctx) {}
    // Uploads every logical flush's records once, into shared buffers, then
    // issues one rive_hip_flush per logical flush with its first_* offsets.
    bool flushFrame(const std::vector<gpu::FlushDescriptor>& flushes) override;
    // Per-draw (tessellation vertices, contours) of `draws`, counted on the
    // device (rive_hip_count_draws).
    bool countDraws(const std::vector<gpu::DrawRecord>& draws, std::vector<uint32_t>* tess,
                    std::vector<uint32_t>* contours) override;
    // rive_hip_image_decode: host entropy decoding, device reconstruction,
    // premultiplication and mips.
    std::shared_ptr<RiveRenderImage> decodeImageTexture(const uint8_t* bytes, size_t size) override
    {
        rive_hip_image* h = nullptr;
        uint32_t format = 0, w = 0, hgt = 0;
        if (rive_image_probe(bytes, size, &format, &w, &hgt) != RIVE_HIP_OK ||
            rive_hip_image_decode(m_ctx, bytes, size, &h) != RIVE_HIP_OK)
            return nullptr;
        return std::make_shared<RiveRenderImage>(h, w, hgt, /*owned=*/true);
    }
    bool interiorTriangulation() const override { return rive_hip_get_interior_triangulation(m_ctx) != 0; }

private:
    bool upload(int kind, const void* data, size_t bytes);
    rive_hip_context* m_ctx;
    size_t m_sizes[RIVE_HIP_BUFFER_KIND_COUNT] = {};
};
// The serving loop's scenes with their geometry resident on the device:
// every path's verbs and points are uploaded once into the VERB and POINT
// buffers, and each batch only records draws (matrices and paints) and
// flushes, one canvas per request.
struct ResidentScene
{
    const uint8_t* verbs;
    const float* points;
    const rive_stream_path* paths;
    uint32_t pathCount;
    const uint32_t* colors; // ColorInt per path, solid srcOver fills
};

struct SceneCanvas
{
    uint32_t scene;
    const float* matrices; // 6 per path: the path's view matrix
    rive_hip_target* target;
    uint32_t width, height;
};

class ResidentSession
{
public:
    int init(rive_hip_context* ctx, const std::vector<ResidentScene>& scenes);
    // Renders the canvases with batched flushes (consecutive canvases share
    // a flush while their contours fit the 16-bit contour IDs).
    int render(const std::vector<SceneCanvas>& canvases);
    // render() in two halves: prepare records and lays the flushes out on
    // the host (no device work, so it can overlap earlier flushes still on
    // the device); submit uploads them and launches their passes.
    int prepare(const std::vector<SceneCanvas>& canvases, std::vector<gpu::FlushDescriptor>* out);
    int submit(const std::vector<gpu::FlushDescriptor>& flushes);

private:
    rive_hip_context* m_ctx = nullptr;
    std::vector<std::vector<RiveRenderPath>> m_paths;
    std::vector<std::vector<RiveRenderPaint>> m_paints;
    std::vector<uint32_t> m_contours;
};

} // namespace rive
