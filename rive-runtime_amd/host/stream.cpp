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
