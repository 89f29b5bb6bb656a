// Front-end checks (no device work: the flush goes to a recording backend).
// Replays the call sequences of the reference's RiveRenderer tests
// (tests/unit_tests/renderer/pls_renderer_test.cpp): IsAABB (:825-932),
// clip-rects (:553-738), nan-render-path, stroke-thickness, empty-paths,
// empty-paths2 and invalid-gradient-stops (:740-823, :934-1045, :1113-1174), and
// checks the same expectations; and the modulate-opacity tests
// (modulate_opacity_test.cpp).  Prints "OK" or the failures.
#include "rive_renderer.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

using namespace rive;

namespace
{
int g_failures = 0;

void expect(bool ok, const char* what, int line)
{
    if (!ok)
    {
        std::printf("FAIL line %d: %s\n", line, what);
        ++g_failures;
    }
}
#define EXPECT(c) expect((c), #c, __LINE__)

bool near_rect(const RiveRenderer::AABB& r, float l, float t, float rr, float b)
{
    auto near = [](float a, float x) { return std::fabs(a - x) <= 1e-3f; };
    return near(r.left, l) && near(r.top, t) && near(r.right, rr) && near(r.bottom, b);
}
#define EXPECT_RECT(r, l, t, rr, b) expect(near_rect((r), (l), (t), (rr), (b)), #r " ~ " #l "," #t "," #rr "," #b, __LINE__)

class RecordingImpl : public gpu::RenderContextImpl
{
public:
    bool flushFrame(const std::vector<gpu::FlushDescriptor>& f) override
    {
        last = f;
        return true;
    }
    std::vector<gpu::FlushDescriptor> last;
    // Every draw counts as one small path: these scenes fit one logical flush.
    bool countDraws(const std::vector<gpu::DrawRecord>& draws, std::vector<uint32_t>* tess,
                    std::vector<uint32_t>* contours) override
    {
        tess->assign(draws.size(), 8);
        contours->assign(draws.size(), 1);
        return true;
    }
};

void rect_path(RiveRenderPath& p, float l, float t, float r, float b)
{
    p.moveTo(l, t);
    p.lineTo(r, t);
    p.lineTo(r, b);
    p.lineTo(l, b);
    p.close();
}

void oval_path(RiveRenderPath& p, float l, float t, float r, float b)
{
    // Four cubics through the axis extremes (circle constant 0.5523).
    const float cx = (l + r) * .5f, cy = (t + b) * .5f, rx = (r - l) * .5f, ry = (b - t) * .5f;
    const float k = 0.55228475f;
    p.moveTo(cx + rx, cy);
    p.cubicTo(cx + rx, cy + k * ry, cx + k * rx, cy + ry, cx, cy + ry);
    p.cubicTo(cx - k * rx, cy + ry, cx - rx, cy + k * ry, cx - rx, cy);
    p.cubicTo(cx - rx, cy - k * ry, cx - k * rx, cy - ry, cx, cy - ry);
    p.cubicTo(cx + k * rx, cy - ry, cx + rx, cy - k * ry, cx + rx, cy);
    p.close();
}
} // namespace

bool aabb_is(const RawPath& p, float l, float t, float r, float b)
{
    float rect[4];
    return RiveRenderer::IsAABB(p, rect) && rect[0] == l && rect[1] == t && rect[2] == r && rect[3] == b;
}

bool not_aabb(const RawPath& p)
{
    float rect[4];
    return !RiveRenderer::IsAABB(p, rect);
}

void add_rect(RawPath& p, float l, float t, float r, float b, bool clockwise)
{
    // RawPath::addRect (src/math/raw_path.cpp:133-156)
    p.move({l, t});
    if (clockwise)
    {
        p.line({r, t});
        p.line({r, b});
        p.line({l, b});
    }
    else
    {
        p.line({l, b});
        p.line({r, b});
        p.line({r, t});
    }
    p.close();
}

// The reference's "IsAABB" renderer test (pls_renderer_test.cpp:825-932).
void is_aabb_cases()
{
    for (bool doClose : {true, false})
    {
        RawPath a;
        if (doClose)
            add_rect(a, 0, 0, 10, 10, false);
        else
        {
            a.line({10, 0});
            a.line({10, 10});
            a.line({0, 10});
        }
        EXPECT(aabb_is(a, 0, 0, 10, 10));
        RawPath b;
        if (doClose)
            add_rect(b, 1, 2, -1, 5, true);
        else
        {
            b.move({1, 2});
            b.line({-1, 2});
            b.line({-1, 5});
            b.line({1, 5});
        }
        EXPECT(aabb_is(b, -1, 2, 1, 5));
        RawPath c;
        if (doClose)
        {
            add_rect(c, 0, 0, 10, 10, false);
            add_rect(c, 1, 2, -1, 5, false);
        }
        else
        {
            c.line({10, 0});
            c.line({10, 10});
            c.line({0, 10});
            c.line({0, 0});
            c.line({10, 0});
        }
        EXPECT(not_aabb(c));
    }
    RawPath p;
    p.move({1, 0});
    EXPECT(not_aabb(p));
    p.line({0, 0});
    EXPECT(not_aabb(p));
    p.line({0, 1});
    EXPECT(not_aabb(p));
    p.line({1, 1});
    EXPECT(aabb_is(p, 0, 0, 1, 1));
    p.close();
    EXPECT(aabb_is(p, 0, 0, 1, 1));
    RawPath q;
    q.move({0, 0});
    q.line({1, 0});
    q.line({1, 1});
    q.line({0, 1.1f});
    q.close();
    EXPECT(not_aabb(q));
    RawPath u;
    u.move({0, 0});
    u.line({1, 0});
    u.line({1, 1});
    u.line({0, 1});
    EXPECT(aabb_is(u, 0, 0, 1, 1));
    u.close();
    EXPECT(aabb_is(u, 0, 0, 1, 1));
    RawPath v;
    v.move({1, 1});
    v.line({0, 1});
    v.line({0, 0});
    v.line({1, 0});
    EXPECT(aabb_is(v, 0, 0, 1, 1));
    v.line({1, 1});
    EXPECT(aabb_is(v, 0, 0, 1, 1));
    v.close();
    EXPECT(aabb_is(v, 0, 0, 1, 1));
    v.close();
    EXPECT(aabb_is(v, 0, 0, 1, 1));
    v.cubic({1, 1}, {1, 1}, {1, 1});
    EXPECT(aabb_is(v, 0, 0, 1, 1));
    v.quad({1, 1}, {1, 1});
    EXPECT(aabb_is(v, 0, 0, 1, 1));
}

// "nan-render-path", "stroke-thickness", "empty-paths", "empty-paths2" and
// "invalid-gradient-stops" (pls_renderer_test.cpp:740-823, 934-1045,
// 1113-1174): NaN and infinite geometry, degenerate strokes and move-only paths
// must not break the front end; a NaN clip empties the clip; invalid gradient
// stops make no gradient.
void robustness_cases()
{
    const float nan = std::numeric_limits<float>::quiet_NaN(), inf = std::numeric_limits<float>::infinity();
    gpu::RenderContext ctx(std::make_unique<RecordingImpl>());
    gpu::RenderContext::FrameDescriptor fd;
    fd.renderTargetWidth = 100;
    fd.renderTargetHeight = 100;
    ctx.beginFrame(fd);
    RiveRenderer r(&ctx);
    RiveRenderPaint paint;
    paint.color(0xffffffff);

    RiveRenderPath finite, nanPath;
    finite.moveTo(1, 2);
    finite.lineTo(3, 4);
    finite.cubicTo(5, 6, 7, 8, 9, 10);
    finite.cubicTo(11, 12, 13, 14, 15, 16);
    finite.lineTo(17, 18);
    nanPath.moveTo(nan, nan);
    nanPath.lineTo(nan, nan);
    nanPath.cubicTo(nan, nan, nan, nan, nan, nan);
    nanPath.cubicTo(nan, nan, nan, nan, nan, nan);
    nanPath.lineTo(nan, nan);

    r.drawPath(&nanPath, &paint);
    r.save();
    r.clipPath(&nanPath);
    r.drawPath(&finite, &paint);
    r.restore();
    r.save();
    r.transform(Mat2D(nan, nan, nan, nan, nan, nan));
    r.drawPath(&finite, &paint);
    r.clipPath(&finite);
    r.drawPath(&nanPath, &paint);
    r.restore();
    EXPECT(ctx.getClipContentID() == 0);
    r.save();
    r.clipPath(&nanPath); // a NaN clip empties the clip: nothing is drawn
    r.drawPath(&finite, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    r.clipPath(&finite);
    r.drawPath(&finite, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    r.restore();
    r.save();
    r.clipPath(&finite);
    r.drawPath(&finite, &paint);
    EXPECT(ctx.getClipContentID() != 0);
    r.restore();
    RiveRenderPath moveNaN;
    moveNaN.moveTo(nan, nan);
    moveNaN.close();
    r.drawPath(&moveNaN, &paint);

    RiveRenderPath zig;
    zig.lineTo(100, 100);
    zig.lineTo(100, 0);
    RiveRenderPaint stroke;
    stroke.style(RenderPaintStyle::stroke);
    uint32_t tiny = 1u;
    float tinyF;
    std::memcpy(&tinyF, &tiny, 4);
    for (float t : {tinyF, nan, inf, -inf})
    {
        stroke.thickness(t);
        r.drawPath(&zig, &stroke);
    }

    RiveRenderPath empty, moves, big, moves2;
    r.drawPath(&empty, &paint);
    moves.moveTo(1, 1);
    moves.moveTo(2, 2);
    moves.lineTo(2, 2);
    r.drawPath(&moves, &paint);
    unsigned seed = 1;
    auto rand100 = [&seed]() {
        seed = seed * 1103515245u + 12345u;
        return (float)((seed >> 8) & 0xffffu) * (100.f / 65535.f);
    };
    for (int i = 0; i < 100000; ++i)
        big.cubicTo(rand100(), rand100(), rand100(), rand100(), rand100(), rand100());
    r.drawPath(&big, &paint);
    for (float k : {0.f, 1000.f, 3000.f, 5000.f, 7000.f})
        moves2.moveTo(k, k + 1000.f);
    r.drawPath(&moves2, &paint);
    EXPECT(ctx.flush(nullptr));

    auto grad = [](std::vector<float> stops) {
        std::vector<ColorInt> colors(std::max<size_t>(stops.size(), 1), 0xff00ff00u);
        const bool lin = (bool)Gradient::MakeLinear(0, 0, 0, 0, colors.data(), stops.data(), stops.size());
        const bool rad = (bool)Gradient::MakeRadial(0, 0, 0, colors.data(), stops.data(), stops.size());
        return lin == rad ? (int)lin : -1;
    };
    EXPECT(grad({0, .5f, 1}) == 1);
    EXPECT(grad({0}) == 1);
    EXPECT(grad({1}) == 1);
    EXPECT(grad({}) == 0);
    EXPECT(grad({0, .5f, 1.1f}) == 0);
    EXPECT(grad({0, .5f, inf}) == 0);
    EXPECT(grad({0, .5f, nan}) == 0);
    EXPECT(grad({-.1f, .5f, 1}) == 0);
    EXPECT(grad({-inf, .5f, 1}) == 0);
    EXPECT(grad({nan, .5f, 1}) == 0);
    EXPECT(grad({0, .4f, .5f, 1}) == 1);
    EXPECT(grad({0, .5f, .4f, 1}) == 0);
    EXPECT(grad({.4f, 0, .5f, 1}) == 0);
    EXPECT(grad({0, .4f, 1, .5f}) == 0);
    EXPECT(grad({0, nan, 1}) == 0);
}

// The reference's modulate-opacity renderer tests
// (tests/unit_tests/renderer/modulate_opacity_test.cpp:58-420): stacking over
// save/restore, repeated calls, identity, zero, deep nesting, clips and
// transforms, gradient modulation and its single-entry cache; plus the alpha
// each draw's PaintData carries.
void modulate_opacity_cases()
{
    auto impl = std::make_unique<RecordingImpl>();
    RecordingImpl* rec = impl.get();
    gpu::RenderContext ctx(std::move(impl));
    gpu::RenderContext::FrameDescriptor fd;
    fd.renderTargetWidth = fd.renderTargetHeight = 100;
    ctx.beginFrame(fd);
    RawPath raw;
    add_rect(raw, 0, 0, 100, 100, true);
    auto path = ctx.makeRenderPath(raw, FillRule::nonZero);
    EXPECT(raw.empty());
    auto paint = ctx.makeRenderPaint();
    paint->color(0xffffffff);
    RiveRenderer r(&ctx);
    auto approx = [](float a, float b) { return std::fabs(a - b) <= 1e-5f * std::max(1.f, std::fabs(b)); };
    EXPECT(r.currentModulatedOpacity() == 1.f);
    r.save();
    r.modulateOpacity(.5f);
    EXPECT(approx(r.currentModulatedOpacity(), .5f));
    r.save();
    r.modulateOpacity(.5f);
    EXPECT(approx(r.currentModulatedOpacity(), .25f));
    r.drawPath(path.get(), paint.get()); // alpha 255 * .25 = 63.75 -> 64
    r.restore();
    EXPECT(approx(r.currentModulatedOpacity(), .5f));
    r.drawPath(path.get(), paint.get()); // 128
    r.restore();
    EXPECT(approx(r.currentModulatedOpacity(), 1.f));
    r.drawPath(path.get(), paint.get()); // 255
    r.save();
    r.modulateOpacity(.5f);
    r.modulateOpacity(.5f);
    r.modulateOpacity(.5f);
    EXPECT(approx(r.currentModulatedOpacity(), .125f));
    r.drawPath(path.get(), paint.get()); // 32
    r.restore();
    r.save();
    r.modulateOpacity(0.f);
    EXPECT(r.currentModulatedOpacity() == 0.f);
    r.drawPath(path.get(), paint.get()); // 0
    r.restore();
    for (int i = 0; i < 10; ++i)
    {
        r.save();
        r.modulateOpacity(.9f);
    }
    EXPECT(approx(r.currentModulatedOpacity(), 0.3486784401f));
    r.drawPath(path.get(), paint.get()); // lround(255 * .34868) = 89
    for (int i = 0; i < 10; ++i)
        r.restore();
    EXPECT(approx(r.currentModulatedOpacity(), 1.f));
    // with a clip path and a transform
    RawPath ovalRaw;
    ovalRaw.move({90, 50});
    ovalRaw.cubic({90, 72}, {72, 90}, {50, 90});
    ovalRaw.cubic({28, 90}, {10, 72}, {10, 50});
    ovalRaw.cubic({10, 28}, {28, 10}, {50, 10});
    ovalRaw.cubic({72, 10}, {90, 28}, {90, 50});
    ovalRaw.close();
    auto clip = ctx.makeRenderPath(ovalRaw, FillRule::nonZero);
    r.save();
    r.clipPath(clip.get());
    r.modulateOpacity(.5f);
    r.drawPath(path.get(), paint.get()); // 128, after its clip draw
    r.restore();
    r.save();
    r.translate(25, 25);
    r.scale(.5f, .5f);
    r.modulateOpacity(.75f);
    EXPECT(approx(r.currentModulatedOpacity(), .75f));
    r.drawPath(path.get(), paint.get()); // lround(191.25) = 191
    r.restore();
    // a gradient paint: every stop's alpha is modulated
    const ColorInt gcolors[] = {0xffff0000, 0xff0000ff};
    const float gstops[] = {0.f, 1.f};
    auto gpaint = ctx.makeRenderPaint();
    gpaint->shader(ctx.makeLinearGradient(0, 0, 100, 0, gcolors, gstops, 2));
    r.save();
    r.modulateOpacity(.5f);
    r.drawPath(path.get(), gpaint.get());
    r.restore();
    r.drawPath(path.get(), gpaint.get());
    EXPECT(ctx.flush(nullptr));
    EXPECT(rec->last.size() == 1);
    if (rec->last.size() == 1)
    {
        const gpu::FlushDescriptor& f = rec->last[0];
        // PaintData values are RGBA with alpha in the top byte; record 0 is
        // the clear color, the clip draw sits before its draw.
        std::vector<uint32_t> alphas;
        for (size_t i = 1; i < f.paints.size(); ++i)
            if ((f.paints[i].params & 0xfu) == 1u)
                alphas.push_back(f.paints[i].value >> 24);
        const std::vector<uint32_t> want = {64, 128, 255, 32, 0, 89, 128, 191};
        EXPECT(alphas == want);
        // the two gradient draws: a simple ramp each, the modulated one with
        // alpha 128 stops, the plain one with 255
        EXPECT(f.gradSpans.size() == 2);
        if (f.gradSpans.size() == 2)
        {
            EXPECT((f.gradSpans[0].color0 >> 24) == 128 && (f.gradSpans[0].color1 >> 24) == 128);
            EXPECT((f.gradSpans[1].color0 >> 24) == 255 && (f.gradSpans[1].color1 >> 24) == 255);
        }
    }
    // Gradient::getModulated: identity, colors, cache, stops, radial
    auto g = Gradient::MakeLinear(0, 0, 100, 0, gcolors, gstops, 2);
    EXPECT(g->getModulated(1.f).get() == g.get());
    auto m1 = g->getModulated(.5f);
    EXPECT(m1.get() != g.get() && colorAlpha(m1->colors[0]) == 128 && colorAlpha(m1->colors[1]) == 128);
    EXPECT(g->getModulated(.5f).get() == m1.get());
    auto m3 = g->getModulated(.25f);
    EXPECT(m3.get() != m1.get() && g->getModulated(.25f).get() == m3.get());
    auto m5 = g->getModulated(.5f);
    EXPECT(m5.get() != m1.get() && m5.get() != m3.get());
    const ColorInt c3[] = {0xffff0000, 0xff00ff00, 0xff0000ff};
    const float s3[] = {0.f, .3f, 1.f};
    auto g3 = Gradient::MakeLinear(0, 0, 100, 0, c3, s3, 3);
    auto g3m = g3->getModulated(.5f);
    EXPECT(g3m->count() == g3->count() && g3m->stops[0] <= g3m->stops[1] && g3m->stops[1] <= g3m->stops[2]);
    const ColorInt rc[] = {0x80ff0000, 0x80ffffff};
    auto rg = Gradient::MakeRadial(50, 50, 50, rc, gstops, 2);
    auto rgm = rg->getModulated(.5f);
    EXPECT(rgm->paintType == rg->paintType && colorAlpha(rgm->colors[0]) == 64 && colorAlpha(rgm->colors[1]) == 64);
}

int main()
{
    is_aabb_cases();
    robustness_cases();
    modulate_opacity_cases();

    gpu::RenderContext ctx(std::make_unique<RecordingImpl>());
    gpu::RenderContext::FrameDescriptor fd;
    fd.renderTargetWidth = 100;
    fd.renderTargetHeight = 100;
    ctx.beginFrame(fd);

    RiveRenderPath rectA, rectB, rectC, ovalA, ovalB, ovalC, drawable;
    rect_path(rectA, 0, 0, 100, 500);
    rect_path(rectB, 1, -1, 20, 40);
    rect_path(rectC, 0, 0, 300, 600);
    oval_path(ovalA, 0, 0, 100, 400);
    oval_path(ovalB, 0, 0, 200, 500);
    oval_path(ovalC, 0, 0, 300, 600);
    oval_path(drawable, -100, -100, 700, 800);
    RiveRenderPaint paint;
    paint.color(0xffffffff);

    RiveRenderer r(&ctx);
    // A lone rect clips with a clip rect: no clip content.
    r.save();
    r.clipPath(&rectA);
    EXPECT(ctx.getClipContentID() == 0);
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    r.restore();

    r.save();
    r.clipPath(&ovalA);
    r.drawPath(&drawable, &paint);
    uint32_t clipID = ctx.getClipContentID();
    EXPECT(clipID != 0);
    EXPECT(!r.hasClipRect());
    r.clipPath(&rectA); // nested: still a clip rect
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == clipID);
    EXPECT(r.hasClipRect());
    EXPECT_RECT(r.getClipRect(), 0, 0, 100, 500);
    EXPECT(r.getClipRectMatrix() == Mat2D());

    // Nested rects under one matrix intersect into one clip rect.
    r.clipPath(&ovalB);
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() != clipID);
    clipID = ctx.getClipContentID();
    r.clipPath(&rectB);
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == clipID);
    EXPECT(r.hasClipRect());
    EXPECT_RECT(r.getClipRect(), 1, 0, 20, 40);
    EXPECT(r.getClipRectMatrix() == Mat2D());

    // An incompatible (rotated) matrix turns the rect into a clip path.
    r.save();
    r.rotate(1);
    r.clipPath(&ovalA);
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() != clipID);
    clipID = ctx.getClipContentID();
    r.clipPath(&rectB);
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() != clipID);
    clipID = ctx.getClipContentID();
    r.restore();
    EXPECT(r.hasClipRect());
    EXPECT_RECT(r.getClipRect(), 1, 0, 20, 40);
    EXPECT(r.getClipRectMatrix() == Mat2D());

    // A compatible (scale + translate) matrix still intersects.
    r.save();
    r.scale(1, 2);
    r.translate(3, 4);
    r.clipPath(&ovalC);
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() != clipID);
    clipID = ctx.getClipContentID();
    r.clipPath(&rectC);
    r.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == clipID);
    EXPECT(r.hasClipRect());
    EXPECT_RECT(r.getClipRect(), 3, 8, 20, 40);
    EXPECT(r.getClipRectMatrix() == Mat2D());
    r.restore();
    r.restore();

    // A new frame (the reference flushes here; the clip buffer starts empty).
    ctx.flush(nullptr);
    ctx.beginFrame(fd);
    RiveRenderer r2(&ctx);
    r2.save();
    const Mat2D m = Mat2D::fromRotation(1) * Mat2D(.3f, .3f, .9f, -.7f, .4f, .2f);
    r2.transform(m);
    r2.clipPath(&rectA);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    EXPECT(r2.hasClipRect());
    EXPECT_RECT(r2.getClipRect(), 0, 0, 100, 500);
    EXPECT(r2.getClipRectMatrix() == m);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    r2.translate(.9f, -.1f);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    EXPECT(r2.hasClipRect());
    EXPECT_RECT(r2.getClipRect(), 1.9f, 0, 20, 39.9f);
    EXPECT(r2.getClipRectMatrix() == m);
    r2.scale(10, 2);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    // Quarter turns keep the rect axis aligned.
    const float pi = 3.14159265358979f;
    r2.rotate(pi / 2);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(r2.hasClipRect());
    EXPECT_RECT(r2.getClipRect(), 10.9f, 1.9f, 10.9f, 39.9f);
    EXPECT(r2.getClipRectMatrix() == m);
    EXPECT(ctx.getClipContentID() == 0);
    r2.rotate(pi);
    r2.translate(-10, 0);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(r2.hasClipRect());
    EXPECT_RECT(r2.getClipRect(), 10.9f, 1.9f, 10.9f, 17.9f);
    EXPECT(r2.getClipRectMatrix() == m);
    EXPECT(ctx.getClipContentID() == 0);
    r2.rotate(-pi / 2);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    EXPECT(r2.hasClipRect());
    EXPECT_RECT(r2.getClipRect(), 10.9f, 1.9f, -9.1f, 17.9f);
    EXPECT(r2.getClipRectMatrix() == m);
    // Flips too.
    r2.scale(1, -1);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    EXPECT(r2.hasClipRect());
    EXPECT_RECT(r2.getClipRect(), 10.9f, 17.9f, -9.1f, 17.9f);
    EXPECT(r2.getClipRectMatrix() == m);
    r2.scale(-1, -1);
    r2.translate(2, 0);
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() == 0);
    EXPECT(r2.hasClipRect());
    EXPECT_RECT(r2.getClipRect(), 30.9f, 17.9f, -9.1f, 17.9f);
    EXPECT(r2.getClipRectMatrix() == m);
    // A skew makes them incompatible: a clip path, and the rect stays.
    r2.transform(Mat2D(1, .1f, 0, 1, 0, 0));
    r2.clipPath(&rectB);
    r2.drawPath(&drawable, &paint);
    EXPECT(ctx.getClipContentID() != 0);
    EXPECT_RECT(r2.getClipRect(), 30.9f, 17.9f, -9.1f, 17.9f);
    EXPECT(r2.getClipRectMatrix() == m);
    r2.restore();

    if (g_failures == 0)
        std::printf("OK\n");
    return g_failures == 0 ? 0 : 1;
}
