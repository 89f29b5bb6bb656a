// Seeded synthetic draw streams: recorded sequences of rive::Renderer calls
// used by the benchmark and the parity tests.
//
// Random numbers: std::mt19937_64 with the float and bounded-int mappings of
// the reference's test RNG (tests/common/rand.hpp:13-54):
//   f32(a, b)  = a + (u64 >> 40) * 2^-24 * (b - a)
//   u32(a, b)  = Lemire's nearly-divisionless bounded integer.
//
// Draw k of a stream consumes random numbers in this order:
//   kind, cubic count, p0.xy, then 3 control points (x then y) per cubic,
//   [stroke: width, join, cap], paint selector,
//   [gradient: type, stop count, interior stops, stop colors, geometry |
//    solid: r, g, b, alpha], clip selector, [2 clip paths of 8 cubics].
#include "../../include/rive_hip.h"
#include "owned_stream.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace
{
using rivehip_host::OwnedStream;
using rivehip_host::Rand;

void identity(float m[6])
{
    m[0] = 1, m[1] = 0, m[2] = 0, m[3] = 1, m[4] = 0, m[5] = 0;
}

// Config 1: W x H (1024x1024 by default), 200 closed paths of 8-32 cubics,
// control points uniform over the canvas, nonZero fill, solid premultiplied
// srcOver paints with alpha U[0.5, 1], no clips.
void gen_config1(OwnedStream* o, Rand& r, uint32_t W, uint32_t H, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i)
    {
        uint32_t cubics = r.u32(8, 32);
        uint32_t path = o->addPath(r, cubics, true, 0, (float)W, (float)H);
        rive_stream_draw d{};
        d.path = path;
        identity(d.matrix);
        d.style = 0;
        d.blend_mode = 3;
        d.paint_type = 0;
        d.color = o->randomColor(r);
        o->draws.push_back(d);
    }
}

// Config 2's paint and clip choices for draw d (gradient 30%, clips 10%).
void paint_and_clips(OwnedStream* o, Rand& r, rive_stream_draw* d, float fw, float fh);

// Config 2: W x H (1920x1080 by default), 5000 paths of 4-64 cubics: 60%
// nonZero fills, 20% evenOdd fills (both closed), 20% open strokes (width
// log-uniform 1..20 px, random join and cap); 30% of paints are linear or
// radial gradients with 2..8 stops; 10% of draws sit under two nested clip
// paths (closed, 8 cubics, nonZero).
void gen_config2(OwnedStream* o, Rand& r, uint32_t W, uint32_t H, uint32_t n)
{
    const float fw = (float)W, fh = (float)H;
    for (uint32_t i = 0; i < n; ++i)
    {
        float kind = r.f32(0, 1);
        uint32_t cubics = r.u32(4, 64);
        rive_stream_draw d{};
        identity(d.matrix);
        d.blend_mode = 3;
        if (kind < .6f)
        {
            d.path = o->addPath(r, cubics, true, 0, fw, fh);
            d.style = 0;
        }
        else if (kind < .8f)
        {
            d.path = o->addPath(r, cubics, true, 1, fw, fh);
            d.style = 0;
        }
        else
        {
            d.path = o->addPath(r, cubics, false, 0, fw, fh);
            d.style = 1;
            d.thickness = expf(r.f32(0.f, logf(20.f)));
            d.join = r.u32(0, 2);
            d.cap = r.u32(0, 2);
        }
        paint_and_clips(o, r, &d, fw, fh);
        o->draws.push_back(d);
    }
}

void paint_and_clips(OwnedStream* o, Rand& r, rive_stream_draw* dp, float fw, float fh)
{
    rive_stream_draw& d = *dp;
    {
        if (r.f32(0, 1) < .3f)
        {
            d.paint_type = 1 + r.u32(0, 1);
            uint32_t nstops = r.u32(2, 8);
            d.first_stop = (uint32_t)o->stopColors.size();
            d.stop_count = nstops;
            std::vector<float> stops(nstops);
            stops[0] = 0;
            stops[nstops - 1] = 1;
            for (uint32_t k = 1; k + 1 < nstops; ++k)
                stops[k] = r.f32(0, 1);
            std::sort(stops.begin() + 1, stops.end() - 1);
            for (uint32_t k = 0; k < nstops; ++k)
            {
                o->stopOffsets.push_back(stops[k]);
                o->stopColors.push_back(o->randomColor(r));
            }
            if (d.paint_type == 1)
            {
                d.grad[0] = r.f32(0, fw);
                d.grad[1] = r.f32(0, fh);
                d.grad[2] = r.f32(0, fw);
                d.grad[3] = r.f32(0, fh);
            }
            else
            {
                d.grad[0] = r.f32(0, fw);
                d.grad[1] = r.f32(0, fh);
                d.grad[2] = r.f32(.05f, .5f) * std::max(fw, fh);
                d.grad[3] = 0;
            }
        }
        else
        {
            d.paint_type = 0;
            d.color = o->randomColor(r);
        }
        if (r.f32(0, 1) < .1f)
        {
            d.clip_count = 2;
            d.clip_paths[0] = o->addPath(r, 8, true, 0, fw, fh);
            d.clip_paths[1] = o->addPath(r, 8, true, 0, fw, fh);
        }
    }
}

// Config 3: W x H (3840x2160 by default), 20000 draws: config 2's mix with
// 25% of the draws feathered clockwise fills (feather U[1, 40] px) and 20% of
// the draws blended with multiply, screen, overlay or color-dodge; plus 64
// 256x256 RGBA8 images (a gradient under seeded noise, opaque or with
// varying alpha) drawn with drawImage at evenly spaced places in the stream
// under random affine transforms (scale U[.25, 4], rotation U[0, 2pi),
// translation over the canvas).
void gen_config3(OwnedStream* o, Rand& r, uint32_t W, uint32_t H, uint32_t n)
{
    const float fw = (float)W, fh = (float)H;
    const uint32_t kImages = 64, kSide = 256;
    for (uint32_t k = 0; k < kImages; ++k)
    {
        std::vector<uint8_t> px((size_t)kSide * kSide * 4);
        const bool translucent = (k & 1) != 0;
        for (uint32_t y = 0; y < kSide; ++y)
            for (uint32_t x = 0; x < kSide; ++x)
            {
                const uint32_t noise = r.u32(0, 63);
                float rr = ((float)x / kSide) * 192.f + (float)noise, gg = ((float)y / kSide) * 192.f + (float)noise;
                float bb = (float)((x ^ y) & 0xff) * .5f + (float)noise;
                float a = translucent ? 255.f * (.25f + .75f * (float)((x + y) & 0xff) / 255.f) : 255.f;
                uint8_t* q = &px[((size_t)y * kSide + x) * 4];
                q[0] = (uint8_t)lrintf(std::min(rr, 255.f) * a / 255.f);
                q[1] = (uint8_t)lrintf(std::min(gg, 255.f) * a / 255.f);
                q[2] = (uint8_t)lrintf(std::min(bb, 255.f) * a / 255.f);
                q[3] = (uint8_t)lrintf(a);
            }
        o->imagePixels.push_back(std::move(px));
        o->images.push_back({kSide, kSide, nullptr});
    }
    const uint32_t blends[4] = {24, 14, 15, 18}; // rive::BlendMode multiply, screen, overlay, colorDodge
    const uint32_t every = std::max<uint32_t>(n / kImages, 1);
    uint32_t nextImage = 0;
    for (uint32_t i = 0; i < n; ++i)
    {
        if (i % every == 0 && nextImage < kImages)
        {
            rive_stream_draw d{};
            d.image = ++nextImage;
            d.image_sampler = 9; // bilinear, clamped
            d.opacity = 1.f;
            d.blend_mode = 3;
            const float s = r.f32(.25f, 4.f), a = r.f32(0.f, 6.28318530718f);
            const float tx = r.f32(0, fw), ty = r.f32(0, fh);
            const float c = cosf(a) * s, sn = sinf(a) * s;
            d.matrix[0] = c, d.matrix[1] = sn, d.matrix[2] = -sn, d.matrix[3] = c;
            d.matrix[4] = tx, d.matrix[5] = ty;
            o->draws.push_back(d);
        }
        const uint32_t cubics = r.u32(4, 64);
        rive_stream_draw d{};
        identity(d.matrix);
        d.blend_mode = 3;
        if (r.f32(0, 1) < .25f)
        {
            d.path = o->addPath(r, cubics, true, 2, fw, fh); // clockwise
            d.style = 0;
            d.feather = r.f32(1.f, 40.f);
        }
        else
        {
            const float kind = r.f32(0, 1);
            if (kind < .6f)
            {
                d.path = o->addPath(r, cubics, true, 0, fw, fh);
                d.style = 0;
            }
            else if (kind < .8f)
            {
                d.path = o->addPath(r, cubics, true, 1, fw, fh);
                d.style = 0;
            }
            else
            {
                d.path = o->addPath(r, cubics, false, 0, fw, fh);
                d.style = 1;
                d.thickness = expf(r.f32(0.f, logf(20.f)));
                d.join = r.u32(0, 2);
                d.cap = r.u32(0, 2);
            }
        }
        if (r.f32(0, 1) < .2f)
            d.blend_mode = blends[r.u32(0, 3)];
        paint_and_clips(o, r, &d, fw, fh);
        o->draws.push_back(d);
    }
}
} // namespace

extern "C" int rive_stream_generate(int config, uint64_t seed, uint32_t width, uint32_t height,
                                    uint32_t draw_count, rive_stream** out)
{
    if (out == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    auto* o = new OwnedStream;
    Rand r(seed);
    if (config == 1)
    {
        uint32_t W = width ? width : 1024, H = height ? height : 1024;
        gen_config1(o, r, W, H, draw_count ? draw_count : 200);
        o->finish(W, H);
    }
    else if (config == 2)
    {
        uint32_t W = width ? width : 1920, H = height ? height : 1080;
        gen_config2(o, r, W, H, draw_count ? draw_count : 5000);
        o->finish(W, H);
    }
    else if (config == 3)
    {
        uint32_t W = width ? width : 3840, H = height ? height : 2160;
        gen_config3(o, r, W, H, draw_count ? draw_count : 20000);
        o->finish(W, H);
    }
    else
    {
        delete o;
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    }
    *out = &o->s;
    return RIVE_HIP_OK;
}

extern "C" int rive_stream_generate_batch(uint64_t seed, uint32_t n, rive_stream** out)
{
    // The batched workload: canvases of log-uniform size and aspect, each
    // with a log-uniform number of config-1 paths over its own area.
    if (out == nullptr || n == 0)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    Rand r(seed);
    for (uint32_t i = 0; i < n; ++i)
    {
        const float side = expf(r.f32(logf(64.f), logf(1024.f)));
        const float aspect = expf(r.f32(logf(.5f), logf(2.f)));
        const uint32_t W = std::max<uint32_t>(1, (uint32_t)lroundf(side * sqrtf(aspect)));
        const uint32_t H = std::max<uint32_t>(1, (uint32_t)lroundf(side / sqrtf(aspect)));
        const uint32_t paths = (uint32_t)lroundf(expf(r.f32(logf(10.f), logf(2000.f))));
        auto* o = new OwnedStream;
        gen_config1(o, r, W, H, paths);
        o->finish(W, H);
        out[i] = &o->s;
    }
    return RIVE_HIP_OK;
}

extern "C" void rive_stream_free(rive_stream* stream)
{
    if (stream == nullptr)
        return;
    // rive_stream is the first member of OwnedStream.
    delete reinterpret_cast<OwnedStream*>(stream);
}
