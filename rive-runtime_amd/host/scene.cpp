// Animated scenes and the serving loop.
//
// A scene is a transform tree of nodes with keyframed affine animation and
// closed cubic paths hanging off the nodes.  Evaluated at a time t it yields
// a draw stream (one rive::Renderer drawPath per path under its node's world
// transform) for any output size.  The node transform and its composition
// follow the reference runtime: local = rotation(r) with translation (x, y),
// then scaleByValues(sx, sy) (transform_component.cpp:63-69); world =
// parent.world * local (:73-84); each keyed value interpolates as
// from + (to - from) * ease(f) (cubic_ease_interpolator.cpp:5-18).
//
// The serving loop renders a stream of timed requests (scene, time, output
// size) with continuous batching: whenever the device is free it admits every
// request that has arrived, in arrival order, while the batch's estimated
// HBM footprint stays within the budget and its size within the slot count,
// renders the whole batch as one batched flush (one canvas per request, all
// canvases in the same kernel launches), and records arrival, admission and
// completion times.  While a batch is on the device the host already admits
// and lays out the next one.  Every scene's path geometry is uploaded to the device
// once and stays resident; a batch records only the paths' matrices and
// paints.  Output targets of each size stay resident in a pool and are
// reused across batches, as do the context's pipeline buffers.
#include "../../include/rive_hip.h"
#include "owned_stream.hpp"
#include "rive_renderer.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

using rivehip_host::OwnedStream;
using rivehip_host::Rand;

namespace
{
// Cubic-bezier easing (the CSS/After-Effects curve through (0,0), (x1,y1),
// (x2,y2), (1,1)): x(t) is inverted by a table of 11 samples, then four
// Newton steps where the slope is at least 1e-3, else bisection to 1e-7 in
// at most 10 steps, and y(t) returned.  This is the published bezier-easing
// method (G. Renaudeau) that cubic_interpolator_solver.cpp:12-95 implements.
struct CubicEase
{
    static constexpr int kSamples = 11;
    float x1 = 0, y1 = 0, x2 = 1, y2 = 1;
    float table[kSamples] = {};

    static float bezier(float t, float a1, float a2)
    {
        // ((1 - 3 a2 + 3 a1) t + (3 a2 - 6 a1)) t + 3 a1) t
        return (((1.0f - 3.0f * a2 + 3.0f * a1) * t + (3.0f * a2 - 6.0f * a1)) * t + (3.0f * a1)) * t;
    }
    static float slope(float t, float a1, float a2)
    {
        return 3.0f * (1.0f - 3.0f * a2 + 3.0f * a1) * t * t + 2.0f * (3.0f * a2 - 6.0f * a1) * t + (3.0f * a1);
    }
    void build()
    {
        for (int i = 0; i < kSamples; ++i)
            table[i] = bezier(i * (1.0f / (kSamples - 1)), x1, x2);
    }
    float solve_t(float x) const
    {
        const float step = 1.0f / (kSamples - 1);
        float start = 0.0f;
        int k = 1;
        for (; k != kSamples - 1 && table[k] <= x; ++k)
            start += step;
        --k;
        const float dist = (x - table[k]) / (table[k + 1] - table[k]);
        float t = start + dist * step;
        const float s0 = slope(t, x1, x2);
        if (s0 >= 0.001f)
        {
            for (int i = 0; i < 4; ++i)
            {
                const float s = slope(t, x1, x2);
                if (s == 0.0f)
                    return t;
                t -= (bezier(t, x1, x2) - x) / s;
            }
            return t;
        }
        if (s0 == 0.0f)
            return t;
        float hi = start + step, cur = 0.f;
        int i = 0;
        float dx;
        do
        {
            cur = start + (hi - start) / 2.0f;
            dx = bezier(cur, x1, x2) - x;
            if (dx > 0.0f)
                hi = cur;
            else
                start = cur;
        } while (std::fabs(dx) > 0.0000001f && ++i < 10);
        return cur;
    }
    float apply(float f) const { return bezier(solve_t(f), y1, y2); }
};

constexpr int kChannels = 5; // x, y, rotation, scaleX, scaleY

struct Key
{
    float time;
    float value[kChannels];
    uint32_t cubic; // 0 linear, 1 cubic-bezier ease towards the next key
    CubicEase ease;
};

struct Node
{
    int32_t parent; // -1 for the root; parents precede their children
    std::vector<Key> keys;
};

// 2x3 affine matrices as [a, b, c, d, tx, ty]: x' = a x + c y + tx.
void mat_mul(const float A[6], const float B[6], float out[6])
{
    const float r[6] = {A[0] * B[0] + A[2] * B[1], A[1] * B[0] + A[3] * B[1],
                        A[0] * B[2] + A[2] * B[3], A[1] * B[2] + A[3] * B[3],
                        A[0] * B[4] + A[2] * B[5] + A[4], A[1] * B[4] + A[3] * B[5] + A[5]};
    std::memcpy(out, r, sizeof(r));
}

} // namespace

struct rive_scene
{
    std::vector<Node> nodes;
    float duration = 0;
    float extent = 512; // paths live in [-extent/2, extent/2]^2 of the root's space
    // paths (closed cubic contours) and the node each hangs off
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    std::vector<rive_stream_path> paths;
    std::vector<uint32_t> pathNode;
    std::vector<uint32_t> colors;
    uint64_t cubics = 0;

    void local(uint32_t n, float t, float out[6]) const
    {
        const std::vector<Key>& ks = nodes[n].keys;
        float v[kChannels];
        if (t <= ks.front().time)
            std::memcpy(v, ks.front().value, sizeof(v));
        else if (t >= ks.back().time)
            std::memcpy(v, ks.back().value, sizeof(v));
        else
        {
            size_t k = 0;
            while (k + 1 < ks.size() && ks[k + 1].time <= t)
                ++k;
            const Key& a = ks[k];
            const Key& b = ks[k + 1];
            const float f = (t - a.time) / (b.time - a.time);
            const float e = a.cubic ? a.ease.apply(f) : f;
            for (int c = 0; c < kChannels; ++c)
                v[c] = a.value[c] + (b.value[c] - a.value[c]) * e;
        }
        // rotation with translation, then scaleByValues
        float s = 0.f, c = 1.f;
        if (v[2] != 0.f)
        {
            s = std::sin(v[2]);
            c = std::cos(v[2]);
        }
        out[0] = c * v[3];
        out[1] = s * v[3];
        out[2] = -s * v[4];
        out[3] = c * v[4];
        out[4] = v[0];
        out[5] = v[1];
    }

    void world(float t, std::vector<float>& out) const
    {
        out.resize(nodes.size() * 6);
        for (uint32_t n = 0; n < nodes.size(); ++n)
        {
            float L[6];
            local(n, t, L);
            if (nodes[n].parent < 0)
                std::memcpy(&out[6 * n], L, sizeof(L));
            else
                mat_mul(&out[6 * (size_t)nodes[n].parent], L, &out[6 * n]);
        }
    }
};

extern "C" int rive_scene_generate(uint64_t seed, rive_scene** out)
{
    if (out == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    auto sc = std::make_unique<rive_scene>();
    Rand r(seed);
    const uint32_t nodeCount = r.u32(50, 500);
    const uint32_t pathCount = r.u32(20, 1500);
    sc->duration = 4.f;
    std::vector<uint32_t> depth(nodeCount, 0);
    for (uint32_t n = 0; n < nodeCount; ++n)
    {
        Node nd;
        if (n == 0)
            nd.parent = -1;
        else
        {
            // a parent among the earlier nodes, keeping depth <= 6
            uint32_t p;
            do
                p = r.u32(0, n - 1);
            while (depth[p] >= 5);
            nd.parent = (int32_t)p;
            depth[n] = depth[p] + 1;
        }
        const uint32_t keyCount = r.u32(2, 10);
        std::vector<float> times(keyCount);
        times[0] = 0.f;
        times[keyCount - 1] = sc->duration;
        for (uint32_t k = 1; k + 1 < keyCount; ++k)
            times[k] = r.f32(0.f, sc->duration);
        std::sort(times.begin(), times.end());
        for (uint32_t k = 0; k < keyCount; ++k)
        {
            Key key{};
            key.time = times[k];
            const float reach = n == 0 ? 0.f : 48.f;
            key.value[0] = r.f32(-reach, reach);
            key.value[1] = r.f32(-reach, reach);
            key.value[2] = r.f32(-0.6f, 0.6f);
            key.value[3] = r.f32(0.8f, 1.2f);
            key.value[4] = r.f32(0.8f, 1.2f);
            key.cubic = r.u32(0, 1);
            if (key.cubic)
            {
                key.ease.x1 = r.f32(0.f, 1.f);
                key.ease.y1 = r.f32(-0.5f, 1.5f);
                key.ease.x2 = r.f32(0.f, 1.f);
                key.ease.y2 = r.f32(-0.5f, 1.5f);
                key.ease.build();
            }
            nd.keys.push_back(key);
        }
        // equal key times would divide by zero: nudge them apart
        for (uint32_t k = 1; k < keyCount; ++k)
            if (!(nd.keys[k].time > nd.keys[k - 1].time))
                nd.keys[k].time = std::nextafter(nd.keys[k - 1].time, 1e9f);
        sc->nodes.push_back(nd);
    }
    // paths from the config-1 distribution over the scene extent: closed,
    // 8-32 cubics, control points uniform; solid colours with alpha U[.5, 1]
    const float h = sc->extent * .5f;
    for (uint32_t i = 0; i < pathCount; ++i)
    {
        rive_stream_path p{};
        p.first_verb = (uint32_t)sc->verbs.size();
        p.first_point = (uint32_t)(sc->points.size() / 2);
        p.fill_rule = 0;
        const uint32_t cubics = r.u32(8, 32);
        sc->verbs.push_back(0);
        sc->points.push_back(r.f32(-h, h));
        sc->points.push_back(r.f32(-h, h));
        for (uint32_t c = 0; c < cubics; ++c)
        {
            sc->verbs.push_back(4);
            for (int k = 0; k < 6; ++k)
                sc->points.push_back(r.f32(-h, h));
        }
        sc->verbs.push_back(5);
        p.verb_count = (uint32_t)sc->verbs.size() - p.first_verb;
        p.point_count = (uint32_t)(sc->points.size() / 2) - p.first_point;
        sc->paths.push_back(p);
        sc->pathNode.push_back(r.u32(0, nodeCount - 1));
        const uint32_t cr = r.u32(0, 255), cg = r.u32(0, 255), cb = r.u32(0, 255);
        const uint32_t a = (uint32_t)lrintf(255.f * r.f32(.5f, 1.f));
        sc->colors.push_back((a << 24) | (cr << 16) | (cg << 8) | cb);
        sc->cubics += cubics;
    }
    *out = sc.release();
    return RIVE_HIP_OK;
}

extern "C" void rive_scene_free(rive_scene* scene) { delete scene; }

extern "C" int rive_scene_info(const rive_scene* scene, rive_scene_info_t* info)
{
    if (scene == nullptr || info == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    info->node_count = (uint32_t)scene->nodes.size();
    info->path_count = (uint32_t)scene->paths.size();
    info->cubic_count = scene->cubics;
    info->duration = scene->duration;
    info->extent = scene->extent;
    return RIVE_HIP_OK;
}

extern "C" int rive_scene_node(const rive_scene* scene, uint32_t node, int32_t* parent, rive_scene_key* keys,
                               uint32_t capacity, uint32_t* key_count)
{
    if (scene == nullptr || node >= scene->nodes.size())
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    const Node& nd = scene->nodes[node];
    if (parent)
        *parent = nd.parent;
    if (key_count)
        *key_count = (uint32_t)nd.keys.size();
    for (uint32_t k = 0; keys != nullptr && k < capacity && k < nd.keys.size(); ++k)
    {
        const Key& s = nd.keys[k];
        keys[k].time = s.time;
        std::memcpy(keys[k].value, s.value, sizeof(s.value));
        keys[k].cubic = s.cubic;
        keys[k].ease[0] = s.ease.x1;
        keys[k].ease[1] = s.ease.y1;
        keys[k].ease[2] = s.ease.x2;
        keys[k].ease[3] = s.ease.y2;
    }
    return RIVE_HIP_OK;
}

extern "C" int rive_scene_world_transforms(const rive_scene* scene, float t, float* out)
{
    if (scene == nullptr || out == nullptr)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    std::vector<float> w;
    scene->world(t, w);
    std::memcpy(out, w.data(), w.size() * sizeof(float));
    return RIVE_HIP_OK;
}

extern "C" float rive_ease_cubic(float x1, float y1, float x2, float y2, float f)
{
    CubicEase e;
    e.x1 = x1;
    e.y1 = y1;
    e.x2 = x2;
    e.y2 = y2;
    e.build();
    return e.apply(f);
}

namespace
{
// Each path's matrix at time t for a width x height target: the scene
// extent fits the shorter side, centred, under the path's node's world
// transform.
void view_matrices(const rive_scene* sc, float t, uint32_t W, uint32_t H, std::vector<float>& world,
                   std::vector<float>& out)
{
    sc->world(t, world);
    const float s = (float)std::min(W, H) / sc->extent;
    const float view[6] = {s, 0.f, 0.f, s, W * .5f, H * .5f};
    out.resize(6 * sc->paths.size());
    for (uint32_t i = 0; i < sc->paths.size(); ++i)
        mat_mul(view, &world[6 * (size_t)sc->pathNode[i]], &out[6 * (size_t)i]);
}

// The scene at time t as a draw stream for width x height.
OwnedStream* scene_stream(const rive_scene* sc, float t, uint32_t W, uint32_t H, std::vector<float>& world)
{
    auto* o = new OwnedStream;
    o->verbs = sc->verbs;
    o->points = sc->points;
    o->paths = sc->paths;
    std::vector<float> m;
    view_matrices(sc, t, W, H, world, m);
    o->draws.resize(sc->paths.size());
    for (uint32_t i = 0; i < sc->paths.size(); ++i)
    {
        rive_stream_draw& d = o->draws[i];
        std::memset(&d, 0, sizeof(d));
        d.path = i;
        std::memcpy(d.matrix, &m[6 * (size_t)i], sizeof(d.matrix));
        d.blend_mode = 3; // srcOver
        d.color = sc->colors[i];
        d.opacity = 1.f;
    }
    o->finish(W, H);
    return o;
}

// Estimated HBM footprint of rendering a request: the RGBA8 target, the path
// cell grids (24 bytes per (path, tile) at most) and the expanded geometry
// (about 64 tessellated vertices of 120 bytes of pipeline state per cubic).
uint64_t request_bytes(const rive_scene* sc, uint32_t W, uint32_t H)
{
    const uint64_t tiles = (uint64_t)((W + 15) / 16) * ((H + 15) / 16);
    return 4ull * W * H + 24ull * sc->paths.size() * tiles + 120ull * 64ull * sc->cubics;
}

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
} // namespace

extern "C" int rive_scene_stream(const rive_scene* scene, float t, uint32_t width, uint32_t height,
                                 rive_stream** out)
{
    if (scene == nullptr || out == nullptr || width == 0 || height == 0)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    std::vector<float> world;
    *out = &scene_stream(scene, t, width, height, world)->s;
    return RIVE_HIP_OK;
}

extern "C" uint64_t rive_scene_request_bytes(const rive_scene* scene, uint32_t width, uint32_t height)
{
    return scene ? request_bytes(scene, width, height) : 0;
}

extern "C" int rive_hip_serve(rive_hip_context* ctx, const rive_scene* const* scenes, uint32_t scene_count,
                              const rive_hip_request* requests, uint32_t n, const rive_hip_serve_config* config,
                              rive_hip_request_result* results, uint8_t* const* pixels)
{
    if (ctx == nullptr || scenes == nullptr || requests == nullptr || results == nullptr || config == nullptr ||
        config->max_batch == 0)
        return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i)
    {
        const rive_hip_request& q = requests[i];
        if (q.scene >= scene_count || scenes[q.scene] == nullptr || q.width == 0 || q.height == 0 ||
            (i > 0 && q.arrival_s < requests[i - 1].arrival_s))
            return RIVE_HIP_ERROR_INVALID_ARGUMENT;
    }
    // resident output targets, one free list per size
    std::map<std::pair<uint32_t, uint32_t>, std::vector<rive_hip_target*>> pool;
    std::vector<rive_hip_target*> owned;
    auto take = [&](uint32_t w, uint32_t h, rive_hip_target** t) -> int {
        auto& v = pool[{w, h}];
        if (!v.empty())
        {
            *t = v.back();
            v.pop_back();
            return RIVE_HIP_OK;
        }
        int st = rive_hip_target_create(ctx, w, h, t);
        if (st == RIVE_HIP_OK)
            owned.push_back(*t);
        return st;
    };
    // every scene's geometry resident on the device for the whole loop
    rive::ResidentSession session;
    {
        std::vector<rive::ResidentScene> rs;
        for (uint32_t i = 0; i < scene_count; ++i)
        {
            const rive_scene* sc = scenes[i];
            if (sc == nullptr)
                return RIVE_HIP_ERROR_INVALID_ARGUMENT;
            rs.push_back({sc->verbs.data(), sc->points.data(), sc->paths.data(), (uint32_t)sc->paths.size(),
                          sc->colors.data()});
        }
        const int st = session.init(ctx, rs);
        if (st != RIVE_HIP_OK)
            return st;
    }
    int status = RIVE_HIP_OK;
    const auto t0 = std::chrono::steady_clock::now();
    std::deque<uint32_t> pending;
    uint32_t next = 0, done = 0, batchIndex = 0;
    std::vector<float> world;
    // A batch laid out on the host: its requests, targets and flushes.
    struct Prepared
    {
        std::vector<uint32_t> requests;
        std::vector<std::vector<float>> matrices;
        std::vector<rive_hip_target*> targets;
        std::vector<rive::gpu::FlushDescriptor> flushes;
        double admit = 0;
        uint64_t bytes = 0;
    };
    // Admits the requests that have arrived (in arrival order, under the slot
    // count and the HBM budget; a request larger than the budget alone is
    // admitted alone) and lays them out; an empty batch when none waits.
    auto form = [&](Prepared& p) -> int {
        p = Prepared{};
        const double now = now_s(t0);
        while (next < n && requests[next].arrival_s <= now)
            pending.push_back(next++);
        while (!pending.empty() && p.requests.size() < config->max_batch)
        {
            const rive_hip_request& q = requests[pending.front()];
            const uint64_t b = request_bytes(scenes[q.scene], q.width, q.height);
            if (!p.requests.empty() && p.bytes + b > config->hbm_budget_bytes)
                break;
            p.bytes += b;
            p.requests.push_back(pending.front());
            pending.pop_front();
        }
        if (p.requests.empty())
            return RIVE_HIP_OK;
        p.admit = now_s(t0);
        // only the transforms change per request: the paths' view matrices
        p.matrices.resize(p.requests.size());
        std::vector<rive::SceneCanvas> canvases;
        for (size_t k = 0; k < p.requests.size(); ++k)
        {
            const rive_hip_request& q = requests[p.requests[k]];
            view_matrices(scenes[q.scene], q.time, q.width, q.height, world, p.matrices[k]);
            rive_hip_target* t = nullptr;
            int st = take(q.width, q.height, &t);
            if (st != RIVE_HIP_OK)
                return st;
            p.targets.push_back(t);
            canvases.push_back({q.scene, p.matrices[k].data(), t, q.width, q.height});
        }
        return session.prepare(canvases, &p.flushes);
    };
    Prepared cur, following;
    while (done < n && status == RIVE_HIP_OK)
    {
        if (cur.requests.empty())
        {
            status = form(cur);
            if (status != RIVE_HIP_OK)
                break;
            if (cur.requests.empty())
            {
                std::this_thread::sleep_for(std::chrono::duration<double>(requests[next].arrival_s - now_s(t0)));
                continue;
            }
        }
        status = session.submit(cur.flushes);
        // lay the next batch out on the host while this one is on the device
        if (status == RIVE_HIP_OK)
            status = form(following);
        if (status == RIVE_HIP_OK)
            status = rive_hip_sync(ctx);
        const double finish = now_s(t0);
        for (size_t k = 0; k < cur.requests.size() && status == RIVE_HIP_OK; ++k)
        {
            const uint32_t i = cur.requests[k];
            if (pixels != nullptr && pixels[i] != nullptr)
                status = rive_hip_read_pixels(ctx, cur.targets[k], pixels[i]);
            rive_hip_request_result& res = results[i];
            res.arrival_s = requests[i].arrival_s;
            res.admit_s = cur.admit;
            res.done_s = finish;
            res.batch = batchIndex;
            res.batch_size = (uint32_t)cur.requests.size();
            res.batch_bytes = cur.bytes;
            ++done;
        }
        for (size_t k = 0; k < cur.targets.size(); ++k)
            pool[{requests[cur.requests[k]].width, requests[cur.requests[k]].height}].push_back(cur.targets[k]);
        ++batchIndex;
        cur = std::move(following);
        following = Prepared{};
    }
    for (rive_hip_target* t : owned)
        rive_hip_target_destroy(t);
    return status;
}
