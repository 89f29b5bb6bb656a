// Seeded draw-stream building blocks shared by the stream generators and the
// animated scenes: the reference test RNG's mappings and a stream that owns
// its arrays.
//
// Random numbers: std::mt19937_64 with the float and bounded-int mappings of
// the reference's test RNG (tests/common/rand.hpp:13-54):
//   f32(a, b)  = a + (u64 >> 40) * 2^-24 * (b - a)
//   u32(a, b)  = Lemire's nearly-divisionless bounded integer.
#pragma once
#include "../../include/rive_hip.h"

#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace rivehip_host
{
class Rand
{
public:
    explicit Rand(uint64_t seed) : m_impl(seed) {}
    uint64_t u64() { return m_impl(); }
    uint32_t u32() { return static_cast<uint32_t>(m_impl()); }
    float f32(float start, float end)
    {
        float u = (m_impl() >> 40) * (1.0f / (1 << 24));
        return start + u * (end - start);
    }
    uint32_t u32(uint32_t start, uint32_t end)
    {
        if (start == 0 && end == UINT32_MAX)
            return u32();
        uint32_t range = end - start + 1;
        uint64_t m = uint64_t(u32()) * range;
        uint32_t l = uint32_t(m);
        if (l < range)
        {
            uint32_t threshold = (0u - range) % range;
            while (l < threshold)
            {
                m = uint64_t(u32()) * range;
                l = uint32_t(m);
            }
        }
        return start + uint32_t(m >> 32);
    }

private:
    std::mt19937_64 m_impl;
};

struct OwnedStream
{
    rive_stream s{};
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    std::vector<rive_stream_path> paths;
    std::vector<rive_stream_draw> draws;
    std::vector<uint32_t> stopColors;
    std::vector<float> stopOffsets;
    std::vector<std::vector<uint8_t>> imagePixels; // premultiplied RGBA8
    std::vector<rive_stream_image> images;

    uint32_t addPath(Rand& r, uint32_t cubics, bool close, uint32_t fillRule, float W, float H)
    {
        rive_stream_path p{};
        p.first_verb = (uint32_t)verbs.size();
        p.first_point = (uint32_t)(points.size() / 2);
        p.fill_rule = fillRule;
        verbs.push_back(0); // move
        points.push_back(r.f32(0, W));
        points.push_back(r.f32(0, H));
        for (uint32_t i = 0; i < cubics; ++i)
        {
            verbs.push_back(4); // cubic
            for (int k = 0; k < 3; ++k)
            {
                points.push_back(r.f32(0, W));
                points.push_back(r.f32(0, H));
            }
        }
        if (close)
            verbs.push_back(5);
        p.verb_count = (uint32_t)verbs.size() - p.first_verb;
        p.point_count = (uint32_t)(points.size() / 2) - p.first_point;
        paths.push_back(p);
        return (uint32_t)paths.size() - 1;
    }

    uint32_t randomColor(Rand& r)
    {
        uint32_t cr = r.u32(0, 255), cg = r.u32(0, 255), cb = r.u32(0, 255);
        uint32_t a = (uint32_t)lrintf(255.f * r.f32(.5f, 1.f));
        return (a << 24) | (cr << 16) | (cg << 8) | cb;
    }

    void finish(uint32_t W, uint32_t H)
    {
        s.width = W;
        s.height = H;
        s.clear_color = 0x00000000;
        s.dither = 1;
        s.verbs = verbs.data();
        s.verb_count = (uint32_t)verbs.size();
        s.points = points.data();
        s.point_count = (uint32_t)(points.size() / 2);
        s.paths = paths.data();
        s.path_count = (uint32_t)paths.size();
        s.draws = draws.data();
        s.draw_count = (uint32_t)draws.size();
        s.stop_colors = stopColors.data();
        s.stop_offsets = stopOffsets.data();
        s.stop_count = (uint32_t)stopColors.size();
        for (size_t i = 0; i < images.size(); ++i)
            images[i].rgba_premul = imagePixels[i].data();
        s.images = images.empty() ? nullptr : images.data();
        s.image_count = (uint32_t)images.size();
    }
};

} // namespace rivehip_host
