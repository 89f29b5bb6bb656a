// Softening of feathered fills on the host (RiveRenderPath::
// makeSoftenedCopyForFeathering, rive_render_path.cpp:233-449): curvature is
// what breaks a one-dimensional feather along the normal, so every cubic is
// chopped convex, cusps are crossed with short lines, and each piece is
// chopped again wherever its tangent has turned by a fixed angle that shrinks
// as the feather grows on screen.  The chop maths is the device's (shared
// header), so the host and the counting kernels agree on every chop.
#include "rive_renderer.hpp"
#include "../csrc/curve_math.hpp"

#include <algorithm>
#include <cmath>

namespace rive
{
namespace
{
using rivehip::V2d;

constexpr float kFeatherPolarMinAngle = 3.14159265f / 16; // gpu.hpp:69

// The cubic p[4] chopped where its tangent has turned by each successive
// multiple of the rotation (numChops of them), appended to `out`
// (rive_render_path.cpp:233-321).  The tangent's direction at T is the root of
// a quadratic; roots out of order or closer than 1e-4 are dropped.
void chop_at_uniform_rotation(RawPath* out, const V2d p[4], V2d tan0, int numChops, const float rot[4])
{
    const V2d C = p[1] - p[0], D = p[2] - p[1], E = p[3] - p[0];
    const V2d B = D - C;
    const V2d A = D * -3.f + E;
    V2d tangent = tan0;
    V2d rem[4] = {p[0], p[1], p[2], p[3]};
    float remainingT = 0;
    V2d chops[10];
    for (int i = 0; i < numChops; i += 4)
    {
        const int n = std::min(numChops - i, 4);
        float roots[4] = {0, 0, 0, 0};
        for (int j = 0; j < n; ++j)
        {
            tangent = V2d{rot[0] * tangent.x + rot[2] * tangent.y, rot[1] * tangent.x + rot[3] * tangent.y};
            const float a = A.x * tangent.y - A.y * tangent.x;
            const float b_over_2 = B.x * tangent.y - B.y * tangent.x;
            const float c = C.x * tangent.y - C.y * tangent.x;
            float q = sqrtf(b_over_2 * b_over_2 - a * c);
            q = -b_over_2 - copysignf(q, b_over_2);
            roots[j] = c / q;
        }
        float t[4];
        int numT = 0;
        float maxT = remainingT;
        for (int j = 0; j < n; ++j)
            if (roots[j] > maxT + 1e-4f && roots[j] < 1 - 1e-4f)
                t[numT++] = maxT = roots[j];
        auto local = [&](float tv) { return fminf(fmaxf((tv - remainingT) / (1 - remainingT), 0.f), 1.f); };
        for (int j = 0; j < numT; j += 2)
        {
            if (j + 1 < numT)
            {
                rivehip::chop_at2(rem, chops, local(t[j]), local(t[j + 1]));
                out->cubic({chops[1].x, chops[1].y}, {chops[2].x, chops[2].y}, {chops[3].x, chops[3].y});
                out->cubic({chops[4].x, chops[4].y}, {chops[5].x, chops[5].y}, {chops[6].x, chops[6].y});
                for (int k = 0; k < 4; ++k)
                    rem[k] = chops[6 + k];
                remainingT = t[j + 1];
            }
            else
            {
                rivehip::chop_at(rem, chops, local(t[j]));
                out->cubic({chops[1].x, chops[1].y}, {chops[2].x, chops[2].y}, {chops[3].x, chops[3].y});
                for (int k = 0; k < 4; ++k)
                    rem[k] = chops[3 + k];
                remainingT = t[j];
            }
        }
    }
    out->cubic({rem[1].x, rem[1].y}, {rem[2].x, rem[2].y}, {rem[3].x, rem[3].y});
}

// math::measure_angle_between_vectors (bezier_utils.cpp:161-170).
float angle_between(V2d a, V2d b)
{
    float cosTheta = rivehip::vdot(a, b) / sqrtf(rivehip::vdot(a, a) * rivehip::vdot(b, b));
    cosTheta = std::max(std::min(1.f, cosTheta), -1.f); // NaN -> 1
    return acosf(cosTheta);
}
} // namespace

std::shared_ptr<RiveRenderPath> RiveRenderPath::makeSoftenedCopyForFeathering(float feather,
                                                                              float matrixMaxScale) const
{
    // The angle between chops: the polar density of a radius a quarter of
    // the feather's on screen (precision 2), mirrored back up past pi/16 and
    // capped at pi/2 (the flattening in the tessellator needs it).
    const float r_ = feather * (3.f / 2) * matrixMaxScale * .25f;
    const float cosTheta = 1.f - (1.f / 2) / r_;
    const float polarSegmentsPerRadian = .5f / acosf(std::max(cosTheta, -1.f));
    float rotationBetweenJoins = 1 / polarSegmentsPerRadian;
    if (rotationBetweenJoins < kFeatherPolarMinAngle)
    {
        const float d = (kFeatherPolarMinAngle - rotationBetweenJoins) * 5.f;
        rotationBetweenJoins = kFeatherPolarMinAngle + d * d * d;
    }
    rotationBetweenJoins = std::min(rotationBetweenJoins, 3.14159265f / 2);
    // Mat2D::fromRotation: {cos, sin, -sin, cos}
    const float s = (float)sin((double)rotationBetweenJoins), c = (float)cos((double)rotationBetweenJoins);
    const float rot[4] = {c, s, -s, c};
    const float rev[4] = {c, -s, s, c};

    RawPath out;
    const auto& verbs = m_raw.verbs();
    const auto& pts = m_raw.points();
    size_t pi = 0;
    for (PathVerb v : verbs)
    {
        switch (v)
        {
            case PathVerb::move:
                out.move(pts[pi]);
                pi += 1;
                break;
            case PathVerb::line:
                out.line(pts[pi]);
                pi += 1;
                break;
            case PathVerb::cubic:
            {
                const V2d p4[4] = {{pts[pi - 1].x, pts[pi - 1].y},
                                   {pts[pi].x, pts[pi].y},
                                   {pts[pi + 1].x, pts[pi + 1].y},
                                   {pts[pi + 2].x, pts[pi + 2].y}};
                float T[4] = {0, 0, 0, 0};
                bool areCusps = false;
                int n = rivehip::convex_180_chops(p4, T, &areCusps);
                if (areCusps)
                {
                    // cross each cusp with a short line (padding 1e-2 in T)
                    for (int i = n - 1; i >= 0; --i)
                    {
                        const float minT = i == 0 ? 0.f : (T[i - 1] + T[i]) * .5f;
                        const float maxT = i + 1 == n ? 1.f : (T[i + 1] + T[i]) * .5f;
                        T[i * 2 + 1] = fminf(T[i] + 1e-2f, maxT);
                        T[i * 2 + 0] = fmaxf(T[i] - 1e-2f, minT);
                    }
                    n *= 2;
                }
                V2d chops[16];
                rivehip::chop_at_n(p4, chops, T, n);
                const V2d* p = chops;
                for (int i = 0; i <= n; ++i, p += 3)
                {
                    if (areCusps && (i & 1) == 1)
                    {
                        out.line({p[3].x, p[3].y});
                        continue;
                    }
                    const V2d t0 = rivehip::tan0(p), t1 = rivehip::tan1(p);
                    const int numChops = (int)(angle_between(t0, t1) / rotationBetweenJoins);
                    float turn = rivehip::vcross(p[2] - p[0], p[3] - p[1]);
                    if (turn == 0)
                        turn = rivehip::vcross(t0, t1);
                    chop_at_uniform_rotation(&out, p, t0, numChops, turn >= 0 ? rot : rev);
                }
                pi += 3;
                break;
            }
            case PathVerb::quad:
                pi += 2; // stored as cubics
                break;
            case PathVerb::close:
                out.close();
                break;
        }
    }
    return std::make_shared<RiveRenderPath>(std::move(out), m_fillRule);
}
} // namespace rive
