// Deferred 2D command stream: record rive::Renderer calls as a flat,
// pointer-free byte stream plus a blob arena, and replay it through
// RiveRenderer into a target.
//
// The wire format is the reference's (render_commands.hpp:17-350,
// command_stream.hpp:34-165): each command is one opcode byte followed by its
// plain-data payload, back to back; variable-size data (verbs, points,
// gradient colours and stops) goes to the blob arena, 8-byte aligned, and the
// payload carries its offset.  Opcode values and payload layouts are the
// reference's for the commands used here, so a stream of those commands reads
// the same on both sides.  Resource handles are dense u32 ids with a
// generation.
#include "../../include/rive_hip.h"
#include "rive_renderer.hpp"

#include <cstring>
#include <memory>
#include <vector>

namespace
{
enum class Cmd : uint8_t
{
    makePath = 0,
    makeEmptyPath = 1,
    makePaint = 2,
    makeLinearGradient = 3,
    makeRadialGradient = 4,
    pathFillRule = 10,
    paintStyle = 13,
    paintColor = 14,
    paintThickness = 15,
    paintJoin = 16,
    paintCap = 17,
    paintFeather = 18,
    paintBlendMode = 19,
    paintShader = 20,
    save = 22,
    restore = 23,
    transform = 24,
    drawPath = 25,
    clipPath = 26,
    modulateOpacity = 29,
};
constexpr uint8_t kLastOpcode = 32;
using Handle = uint32_t;
constexpr Handle kNoHandle = ~0u;

struct MakeIdPOD
{
    Handle id;
    uint32_t generation;
};
struct MakePathPOD
{
    Handle id;
    uint32_t generation;
    uint64_t blobOffset;   // verbs
    uint64_t pointsOffset; // points
    uint32_t verbCount, pointCount, fillRule, pad;
};
static_assert(sizeof(MakePathPOD) == 40, "MakePathPOD layout");
struct LinearGradientPOD
{
    Handle id;
    uint32_t generation;
    float sx, sy, ex, ey;
    uint64_t blobOffset;  // colors
    uint64_t stopsOffset; // stops
    uint32_t count, pad;
};
static_assert(sizeof(LinearGradientPOD) == 48, "LinearGradientPOD layout");
struct RadialGradientPOD
{
    Handle id;
    uint32_t generation;
    float cx, cy, radius;
    uint32_t count;
    uint64_t blobOffset;
    uint64_t stopsOffset;
};
static_assert(sizeof(RadialGradientPOD) == 40, "RadialGradientPOD layout");
// (the reference's natural layouts: a u8 after a u32 handle, padded to 8
// bytes; the padding is written as zeros here)
struct PathFillRulePOD
{
    Handle path;
    uint8_t fillRule;
    uint8_t pad[3];
};
static_assert(sizeof(PathFillRulePOD) == 8, "PathFillRulePOD layout");
struct PaintU8POD
{
    Handle paint;
    uint8_t value;
    uint8_t pad[3];
};
static_assert(sizeof(PaintU8POD) == 8, "PaintU8POD layout");
struct PaintColorPOD
{
    Handle paint;
    uint32_t color;
};
struct PaintFloatPOD
{
    Handle paint;
    float value;
};
struct PaintShaderPOD
{
    Handle paint;
    Handle shader;
};
struct TransformPOD
{
    float xx, xy, yx, yy, tx, ty;
};
struct DrawPathPOD
{
    Handle path, paint;
    uint32_t pathVersion, paintVersion;
};
struct ClipPathPOD
{
    Handle path;
    uint32_t version;
};
struct OpacityPOD
{
    float opacity;
};

// Payload sizes by opcode (0: no payload or not supported here).
size_t payload_size(uint8_t op)
{
    switch ((Cmd)op)
    {
        case Cmd::makePath: return sizeof(MakePathPOD);
        case Cmd::makeEmptyPath:
        case Cmd::makePaint: return sizeof(MakeIdPOD);
        case Cmd::makeLinearGradient: return sizeof(LinearGradientPOD);
        case Cmd::makeRadialGradient: return sizeof(RadialGradientPOD);
        case Cmd::pathFillRule: return sizeof(PathFillRulePOD);
        case Cmd::paintStyle:
        case Cmd::paintJoin:
        case Cmd::paintCap:
        case Cmd::paintBlendMode: return sizeof(PaintU8POD);
        case Cmd::paintColor: return sizeof(PaintColorPOD);
        case Cmd::paintThickness:
        case Cmd::paintFeather: return sizeof(PaintFloatPOD);
        case Cmd::paintShader: return sizeof(PaintShaderPOD);
        case Cmd::save:
        case Cmd::restore: return 0;
        case Cmd::transform: return sizeof(TransformPOD);
        case Cmd::drawPath: return sizeof(DrawPathPOD);
        case Cmd::clipPath: return sizeof(ClipPathPOD);
        case Cmd::modulateOpacity: return sizeof(OpacityPOD);
    }
    return ~size_t(0);
}
} // namespace

struct rive_cmd_buffer
{
    std::vector<uint8_t> commands, blobs;
    uint32_t width = 0, height = 0, clearColor = 0, dither = 1;

    template <typename POD> void put(Cmd op, const POD& pod)
    {
        commands.push_back((uint8_t)op);
        const size_t at = commands.size();
        commands.resize(at + sizeof(POD));
        std::memcpy(commands.data() + at, &pod, sizeof(POD));
    }
    void put(Cmd op) { commands.push_back((uint8_t)op); }
    uint64_t blob(const void* data, size_t bytes)
    {
        blobs.resize((blobs.size() + 7) & ~size_t(7), 0);
        const uint64_t at = blobs.size();
        blobs.resize(at + bytes);
        if (bytes)
            std::memcpy(blobs.data() + at, data, bytes);
        return at;
    }
};

extern "C" int rive_cmd_record_stream(const rive_stream* s, rive_cmd_buffer** out)
{
    if (s == nullptr || out == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    if (s->image_count != 0)
        return RIVE_HIP_ERROR_UNSUPPORTED; // images need a decodeImage resource
    auto b = std::make_unique<rive_cmd_buffer>();
    b->width = s->width;
    b->height = s->height;
    b->clearColor = s->clear_color;
    b->dither = s->dither;
    for (uint32_t i = 0; i < s->path_count; ++i)
    {
        const rive_stream_path& p = s->paths[i];
        MakePathPOD m{};
        m.id = i;
        m.blobOffset = b->blob(s->verbs + p.first_verb, p.verb_count);
        m.pointsOffset = b->blob(s->points + 2 * (size_t)p.first_point, 8 * (size_t)p.point_count);
        m.verbCount = p.verb_count;
        m.pointCount = p.point_count;
        m.fillRule = p.fill_rule;
        b->put(Cmd::makePath, m);
    }
    Handle nextShader = 0;
    for (uint32_t i = 0; i < s->draw_count; ++i)
    {
        const rive_stream_draw& d = s->draws[i];
        if (d.image != 0)
            return RIVE_HIP_ERROR_UNSUPPORTED;
        // one paint per draw, configured before use
        const Handle paint = i;
        b->put(Cmd::makePaint, MakeIdPOD{paint, 0});
        b->put(Cmd::paintStyle, PaintU8POD{paint, (uint8_t)d.style, {}});
        b->put(Cmd::paintThickness, PaintFloatPOD{paint, d.thickness});
        b->put(Cmd::paintJoin, PaintU8POD{paint, (uint8_t)d.join, {}});
        b->put(Cmd::paintCap, PaintU8POD{paint, (uint8_t)d.cap, {}});
        b->put(Cmd::paintFeather, PaintFloatPOD{paint, d.feather});
        b->put(Cmd::paintBlendMode, PaintU8POD{paint, (uint8_t)d.blend_mode, {}});
        if (d.paint_type == 0)
            b->put(Cmd::paintColor, PaintColorPOD{paint, d.color});
        else
        {
            const Handle sh = nextShader++;
            const uint64_t colors = b->blob(s->stop_colors + d.first_stop, 4 * (size_t)d.stop_count);
            const uint64_t stops = b->blob(s->stop_offsets + d.first_stop, 4 * (size_t)d.stop_count);
            if (d.paint_type == 1)
            {
                LinearGradientPOD g{};
                g.id = sh;
                g.sx = d.grad[0];
                g.sy = d.grad[1];
                g.ex = d.grad[2];
                g.ey = d.grad[3];
                g.blobOffset = colors;
                g.stopsOffset = stops;
                g.count = d.stop_count;
                b->put(Cmd::makeLinearGradient, g);
            }
            else
            {
                RadialGradientPOD g{};
                g.id = sh;
                g.cx = d.grad[0];
                g.cy = d.grad[1];
                g.radius = d.grad[2];
                g.count = d.stop_count;
                g.blobOffset = colors;
                g.stopsOffset = stops;
                b->put(Cmd::makeRadialGradient, g);
            }
            b->put(Cmd::paintShader, PaintShaderPOD{paint, sh});
        }
        b->put(Cmd::save);
        b->put(Cmd::transform,
               TransformPOD{d.matrix[0], d.matrix[1], d.matrix[2], d.matrix[3], d.matrix[4], d.matrix[5]});
        for (uint32_t c = 0; c < d.clip_count && c < 2; ++c)
            b->put(Cmd::clipPath, ClipPathPOD{d.clip_paths[c], 0});
        b->put(Cmd::drawPath, DrawPathPOD{d.path, paint, 0, 0});
        b->put(Cmd::restore);
    }
    *out = b.release();
    return RIVE_HIP_OK;
}

extern "C" void rive_cmd_free(rive_cmd_buffer* b) { delete b; }

extern "C" int rive_cmd_bytes(const rive_cmd_buffer* b, const uint8_t** commands, uint64_t* command_bytes,
                              const uint8_t** blobs, uint64_t* blob_bytes)
{
    if (b == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    if (commands)
        *commands = b->commands.data();
    if (command_bytes)
        *command_bytes = b->commands.size();
    if (blobs)
        *blobs = b->blobs.data();
    if (blob_bytes)
        *blob_bytes = b->blobs.size();
    return RIVE_HIP_OK;
}

extern "C" int rive_cmd_from_bytes(const uint8_t* commands, uint64_t command_bytes, const uint8_t* blobs,
                                   uint64_t blob_bytes, uint32_t width, uint32_t height, uint32_t clear_color,
                                   uint32_t dither, rive_cmd_buffer** out)
{
    if (out == nullptr || (commands == nullptr && command_bytes) || (blobs == nullptr && blob_bytes))
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    auto b = std::make_unique<rive_cmd_buffer>();
    b->commands.assign(commands, commands + command_bytes);
    b->blobs.assign(blobs, blobs + blob_bytes);
    b->width = width;
    b->height = height;
    b->clearColor = clear_color;
    b->dither = dither;
    *out = b.release();
    return RIVE_HIP_OK;
}

namespace
{
struct Replay
{
    std::vector<std::unique_ptr<rive::RiveRenderPath>> paths;
    std::vector<std::unique_ptr<rive::RiveRenderPaint>> paints;
    std::vector<std::shared_ptr<rive::Gradient>> shaders;

    template <typename T> static T* at(std::vector<std::unique_ptr<T>>& v, Handle h)
    {
        return h < v.size() ? v[h].get() : nullptr;
    }
    template <typename T> static void grow(std::vector<T>& v, Handle h)
    {
        if (h >= v.size())
            v.resize((size_t)h + 1);
    }
};

// Walks the stream, creating resources and issuing Renderer calls; fails on
// an unknown opcode, a truncated payload, a blob out of range or a handle
// that names nothing.
int replay(const rive_cmd_buffer* b, rive::RiveRenderer& r, Replay& res)
{
    const uint8_t* c = b->commands.data();
    const size_t n = b->commands.size();
    size_t pos = 0;
    auto blob = [&](uint64_t off, uint64_t bytes) -> const uint8_t* {
        if (off > b->blobs.size() || bytes > b->blobs.size() - off)
            return nullptr;
        return b->blobs.data() + off;
    };
    while (pos < n)
    {
        const uint8_t op = c[pos++];
        const size_t sz = op <= kLastOpcode ? payload_size(op) : ~size_t(0);
        if (sz == ~size_t(0) || n - pos < sz)
            return RIVE_HIP_ERROR_INVALID_ARGUMENT;
        const uint8_t* p = c + pos;
        pos += sz;
        auto pod = [&](auto& v) { std::memcpy(&v, p, sizeof(v)); };
        switch ((Cmd)op)
        {
            case Cmd::makePath:
            {
                MakePathPOD m;
                pod(m);
                const uint8_t* verbs = blob(m.blobOffset, m.verbCount);
                const uint8_t* pts = blob(m.pointsOffset, 8ull * m.pointCount);
                if ((verbs == nullptr && m.verbCount) || (pts == nullptr && m.pointCount) || m.id == kNoHandle)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                auto path = std::make_unique<rive::RiveRenderPath>((rive::FillRule)m.fillRule);
                std::vector<float> xy(2 * (size_t)m.pointCount);
                if (m.pointCount)
                    std::memcpy(xy.data(), pts, 8ull * m.pointCount);
                uint32_t k = 0;
                for (uint32_t v = 0; v < m.verbCount; ++v)
                {
                    const uint32_t need = verbs[v] == 0 || verbs[v] == 1 ? 1 : verbs[v] == 2 ? 2 : verbs[v] == 4 ? 3 : 0;
                    if (verbs[v] != 5 && need == 0)
                        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                    if (k + need > m.pointCount)
                        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                    const float* q = xy.data() + 2 * (size_t)k;
                    switch (verbs[v])
                    {
                        case 0: path->moveTo(q[0], q[1]); break;
                        case 1: path->lineTo(q[0], q[1]); break;
                        case 2: path->mutableRawPath().quad({q[0], q[1]}, {q[2], q[3]}); break;
                        case 4: path->cubicTo(q[0], q[1], q[2], q[3], q[4], q[5]); break;
                        case 5: path->close(); break;
                    }
                    k += need;
                }
                Replay::grow(res.paths, m.id);
                res.paths[m.id] = std::move(path);
                break;
            }
            case Cmd::makeEmptyPath:
            {
                MakeIdPOD m;
                pod(m);
                if (m.id == kNoHandle)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                Replay::grow(res.paths, m.id);
                res.paths[m.id] = std::make_unique<rive::RiveRenderPath>();
                break;
            }
            case Cmd::makePaint:
            {
                MakeIdPOD m;
                pod(m);
                if (m.id == kNoHandle)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                Replay::grow(res.paints, m.id);
                res.paints[m.id] = std::make_unique<rive::RiveRenderPaint>();
                break;
            }
            case Cmd::makeLinearGradient:
            case Cmd::makeRadialGradient:
            {
                Handle id;
                uint64_t co, so;
                uint32_t count;
                LinearGradientPOD lg{};
                RadialGradientPOD rg{};
                if ((Cmd)op == Cmd::makeLinearGradient)
                {
                    pod(lg);
                    id = lg.id, co = lg.blobOffset, so = lg.stopsOffset, count = lg.count;
                }
                else
                {
                    pod(rg);
                    id = rg.id, co = rg.blobOffset, so = rg.stopsOffset, count = rg.count;
                }
                const uint8_t* colors = blob(co, 4ull * count);
                const uint8_t* stops = blob(so, 4ull * count);
                if (id == kNoHandle || (count && (colors == nullptr || stops == nullptr)))
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                std::vector<uint32_t> cv(count);
                std::vector<float> sv(count);
                if (count)
                {
                    std::memcpy(cv.data(), colors, 4ull * count);
                    std::memcpy(sv.data(), stops, 4ull * count);
                }
                Replay::grow(res.shaders, id);
                res.shaders[id] = (Cmd)op == Cmd::makeLinearGradient
                                      ? rive::Gradient::MakeLinear(lg.sx, lg.sy, lg.ex, lg.ey, cv.data(), sv.data(),
                                                                   count)
                                      : rive::Gradient::MakeRadial(rg.cx, rg.cy, rg.radius, cv.data(), sv.data(),
                                                                   count);
                break;
            }
            case Cmd::pathFillRule:
            {
                PathFillRulePOD m;
                pod(m);
                rive::RiveRenderPath* path = Replay::at(res.paths, m.path);
                if (!path)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                path->fillRule((rive::FillRule)m.fillRule);
                break;
            }
            case Cmd::paintStyle:
            case Cmd::paintJoin:
            case Cmd::paintCap:
            case Cmd::paintBlendMode:
            {
                PaintU8POD m;
                pod(m);
                rive::RiveRenderPaint* pt = Replay::at(res.paints, m.paint);
                if (!pt)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                if ((Cmd)op == Cmd::paintStyle)
                    pt->style(m.value == 1 ? rive::RenderPaintStyle::stroke : rive::RenderPaintStyle::fill);
                else if ((Cmd)op == Cmd::paintJoin)
                    pt->join((rive::StrokeJoin)m.value);
                else if ((Cmd)op == Cmd::paintCap)
                    pt->cap((rive::StrokeCap)m.value);
                else
                    pt->blendMode((rive::BlendMode)m.value);
                break;
            }
            case Cmd::paintColor:
            {
                PaintColorPOD m;
                pod(m);
                rive::RiveRenderPaint* pt = Replay::at(res.paints, m.paint);
                if (!pt)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                pt->color(m.color);
                break;
            }
            case Cmd::paintThickness:
            case Cmd::paintFeather:
            {
                PaintFloatPOD m;
                pod(m);
                rive::RiveRenderPaint* pt = Replay::at(res.paints, m.paint);
                if (!pt)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                if ((Cmd)op == Cmd::paintThickness)
                    pt->thickness(m.value);
                else
                    pt->feather(m.value);
                break;
            }
            case Cmd::paintShader:
            {
                PaintShaderPOD m;
                pod(m);
                rive::RiveRenderPaint* pt = Replay::at(res.paints, m.paint);
                if (!pt)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                if (m.shader == kNoHandle)
                    pt->shader(std::shared_ptr<rive::Gradient>());
                else if (m.shader < res.shaders.size())
                {
                    // a gradient that failed validation paints nothing (its
                    // draw is culled like the stream replay's colour 0)
                    if (res.shaders[m.shader])
                        pt->shader(res.shaders[m.shader]);
                    else
                        pt->color(0);
                }
                else
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                break;
            }
            case Cmd::save: r.save(); break;
            case Cmd::restore: r.restore(); break;
            case Cmd::transform:
            {
                TransformPOD m;
                pod(m);
                r.transform(rive::Mat2D(m.xx, m.xy, m.yx, m.yy, m.tx, m.ty));
                break;
            }
            case Cmd::drawPath:
            {
                DrawPathPOD m;
                pod(m);
                rive::RiveRenderPath* path = Replay::at(res.paths, m.path);
                rive::RiveRenderPaint* pt = Replay::at(res.paints, m.paint);
                if (!path || !pt)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                r.drawPath(path, pt);
                break;
            }
            case Cmd::clipPath:
            {
                ClipPathPOD m;
                pod(m);
                rive::RiveRenderPath* path = Replay::at(res.paths, m.path);
                if (!path)
                    return RIVE_HIP_ERROR_INVALID_ARGUMENT;
                r.clipPath(path);
                break;
            }
            case Cmd::modulateOpacity:
            {
                OpacityPOD m;
                pod(m);
                r.modulateOpacity(m.opacity);
                break;
            }
            default: return RIVE_HIP_ERROR_UNSUPPORTED;
        }
    }
    return RIVE_HIP_OK;
}
} // namespace

extern "C" int rive_hip_render_commands(rive_hip_context* ctx, const rive_cmd_buffer* b, rive_hip_target* target)
{
    if (ctx == nullptr || b == nullptr || target == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    rive::gpu::RenderContext rc(std::make_unique<rive::RenderContextHIPImpl>(ctx));
    rive::gpu::RenderContext::FrameDescriptor fd;
    fd.renderTargetWidth = b->width;
    fd.renderTargetHeight = b->height;
    fd.clearColor = b->clearColor;
    fd.ditherMode = b->dither;
    rc.beginFrame(fd);
    Replay res;
    rive::RiveRenderer r(&rc);
    const int st = replay(b, r, res);
    if (st != RIVE_HIP_OK)
        return st;
    return rc.flush(target) ? RIVE_HIP_OK : RIVE_HIP_ERROR_DEVICE;
}
