// CPU ORACLE (test infrastructure only): restatement of the tessellation
// shader (renderer/src/shaders/tessellate.glsl:61-567), evaluated at the pixel
// centers a TessVertexSpan covers in the 2048-wide tessellation texture.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace oracle
{
namespace
{
constexpr float PI_ = 3.14159265359f;   // common.glsl:7
constexpr float _2PI_ = 6.28318530718f; // common.glsl:8

struct F2
{
    float x, y;
};
inline F2 sub(F2 a, F2 b) { return {a.x - b.x, a.y - b.y}; }
inline float dot2(F2 a, F2 b) { return a.x * b.x + a.y * b.y; }
inline bool neq(F2 a, F2 b) { return a.x != b.x || a.y != b.y; }
// GLSL clamp() = min(max(x, lo), hi) with GPU min/max, which return the
// non-NaN operand (IEEE maxNum/minNum): a NaN input clamps to lo.
inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
inline F2 normalize2(F2 v)
{
    float l = sqrtf(v.x * v.x + v.y * v.y);
    return {v.x / l, v.y / l};
}
inline F2 umix(F2 a, F2 b, float t) // common.glsl:185-188
{
    return {(b.x - a.x) * t + a.x, (b.y - a.y) * t + a.y};
}
inline float det2(F2 c0, F2 c1) { return c0.x * c1.y - c1.x * c0.y; }

// bezier_utils.glsl:24-31
float cosine_between_vectors(F2 a, F2 b)
{
    float ab_cosTheta = dot2(a, b);
    float ab_pow2 = dot2(a, a) * dot2(b, b);
    return (ab_pow2 == 0.f) ? 1.f : clampf(ab_cosTheta * (1.f / sqrtf(ab_pow2)), -1.f, 1.f);
}
// bezier_utils.glsl:63-69
void find_cubic_tangents(F2 p0, F2 p1, F2 p2, F2 p3, F2 t[2])
{
    t[0] = sub(neq(p0, p1) ? p1 : neq(p1, p2) ? p2 : p3, p0);
    t[1] = sub(p3, neq(p3, p2) ? p2 : neq(p2, p1) ? p1 : p0);
}
// common.glsl:197-202
float atan2v(F2 v)
{
    v = normalize2(v);
    float theta = acosf(clampf(v.x, -1.f, 1.f));
    return v.y >= 0.f ? theta : -theta;
}
inline float glsl_mod(float x, float y) { return x - y * floorf(x / y); }
inline uint32_t fbits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
} // namespace

void tessellate_span(const TessVertexSpan& s, const ContourData* contours,
                     const float* pathRecords, TessVertex* texture, uint32_t textureRows)
{
    F2 P0{s.pts[0].x, s.pts[0].y}, P1{s.pts[1].x, s.pts[1].y};
    F2 P2{s.pts[2].x, s.pts[2].y}, P3{s.pts[3].x, s.pts[3].y};
    // The span's path (tessellate.glsl:125-144): a feathered fill's cubic is
    // flattened in proportion to its curvature first (:146-193).
    {
        const uint32_t contourID = s.contourIDWithFlags & CONTOUR_ID_MASK;
        const uint32_t pathID = contourID > 0u && contours != nullptr ? contours[contourID - 1].pathID : 0u;
        if (pathID != 0u && pathRecords != nullptr)
        {
            const float strokeRadius = pathRecords[pathID * 16 + 6];
            const float featherRadius = pathRecords[pathID * 16 + 7];
            if (featherRadius != 0.f && strokeRadius == 0.f)
            {
                V2 q0{P0.x, P0.y}, q1{P1.x, P1.y}, q2{P2.x, P2.y}, q3{P3.x, P3.y};
                soften_feathered_cubic(&q0, &q1, &q2, &q3, featherRadius);
                P1 = F2{q1.x, q1.y};
                P2 = F2{q2.x, q2.y};
            }
        }
    }
    const uint32_t segmentCounts = s.segmentCounts;
    uint32_t parametricSegmentCount = segmentCounts & 0x3ffu;
    if ((s.contourIDWithFlags & CULL_EXCESS_TESSELLATION_SEGMENTS_CONTOUR_FLAG) != 0u)
    {
        // An outer-cubic patch has a fixed span: Wang's formula again on the
        // path's matrix, the excess segments collapsed to T = 0
        // (tessellate.glsl:195-211).
        const uint32_t contourID = s.contourIDWithFlags & CONTOUR_ID_MASK;
        const uint32_t pathID = contourID > 0u && contours != nullptr ? contours[contourID - 1].pathID : 0u;
        const float* m = pathRecords != nullptr ? pathRecords + pathID * 16 : nullptr;
        if (m != nullptr)
        {
            const F2 a{(-2.f * P1.x + P2.x) + P0.x, (-2.f * P1.y + P2.y) + P0.y};
            const F2 b{(-2.f * P2.x + P3.x) + P1.x, (-2.f * P2.y + P3.y) + P1.y};
            const F2 d0{m[0] * a.x + m[2] * a.y, m[1] * a.x + m[3] * a.y};
            const F2 d1{m[0] * b.x + m[2] * b.y, m[1] * b.x + m[3] * b.y};
            const float mm = std::max(d0.x * d0.x + d0.y * d0.y, d1.x * d1.x + d1.y * d1.y);
            const float n = std::max(ceilf(sqrtf(.75f * 4.f * sqrtf(mm))), 1.f);
            parametricSegmentCount = std::min((uint32_t)n, parametricSegmentCount);
        }
    }
    const uint32_t polarSegmentCount = (segmentCounts >> 10) & 0x3ffu;
    const uint32_t joinSegmentCount = segmentCounts >> 20;
    const uint32_t totalVertexCount =
        parametricSegmentCount + polarSegmentCount + joinSegmentCount - 1u;
    const F2 joinTangent{s.joinTangent.x, s.joinTangent.y};

    // ---- vertex stage (tessellate.glsl:61-286) ----
    F2 tangents[2];
    find_cubic_tangents(P0, P1, P2, P3, tangents);
    float theta0v = acosf(cosine_between_vectors(tangents[0], tangents[1]));
    float radsPerPolarSegment = theta0v / float(polarSegmentCount);
    float turn = det2(sub(P2, P0), sub(P3, P1));
    if (turn == 0.f)
        turn = det2(tangents[0], tangents[1]);
    if (turn < 0.f)
        radsPerPolarSegment = -radsPerPolarSegment;
    float radsPerJoinSegment = 0;
    const uint32_t flagsIn = s.contourIDWithFlags;
    if (joinSegmentCount > 1u)
    {
        float joinTheta = acosf(cosine_between_vectors(tangents[1], joinTangent));
        float joinSpan = float(joinSegmentCount);
        if ((flagsIn & (JOIN_TYPE_MASK | EMULATED_STROKE_CAP_CONTOUR_FLAG)) ==
            (ROUND_JOIN_CONTOUR_FLAG | EMULATED_STROKE_CAP_CONTOUR_FLAG))
            joinSpan -= 2.f;
        radsPerJoinSegment = joinTheta / joinSpan;
        if (det2(tangents[1], joinTangent) < 0.f)
            radsPerJoinSegment = -radsPerJoinSegment;
    }

    for (int rect = 0; rect < 2; ++rect)
    {
        float yf = rect == 0 ? s.y : s.reflectionY;
        if (!(yf >= 0.f) || yf >= float(textureRows))
            continue; // NaN/offscreen reflection is discarded
        int32_t packed = rect == 0 ? s.x0x1 : s.reflectionX0X1;
        int32_t x0 = (int32_t)((uint32_t)packed << 16) >> 16;
        int32_t x1 = packed >> 16;
        uint32_t flags = flagsIn;
        if (x1 < x0)
            flags |= MIRRORED_CONTOUR_CONTOUR_FLAG;
        int32_t lo = std::max(std::min(x0, x1), 0);
        int32_t hi = std::min(std::max(x0, x1), (int32_t)kTessTextureWidth);
        uint32_t row = (uint32_t)yf;
        for (int32_t x = lo; x < hi; ++x)
        {
            // ---- fragment stage (tessellate.glsl:289-567) ----
            float px = float(x) + .5f;
            float vArgsX = float(totalVertexCount) - fabsf(float(x1) - px);
            F2 p0 = P0, p1 = P1, p2 = P2, p3 = P3;
            F2 tans[2] = {tangents[0], tangents[1]};
            float vertexIdx = std::max(floorf(vArgsX), 0.f);
            float totalVC = float(totalVertexCount);
            float parametric = float(parametricSegmentCount);
            float joinCount = float(joinSegmentCount);
            float rads = radsPerPolarSegment;
            uint32_t cf = flags;
            float mergedSegmentCount = totalVC - joinCount;
            float mergedVertexID = vertexIdx;
            if (mergedVertexID <= mergedSegmentCount)
            {
                cf &= ~JOIN_TYPE_MASK;
            }
            else
            {
                p0 = p1 = p2 = p3;
                tans[0] = tangents[1];
                tans[1] = joinTangent;
                parametric = 1.f;
                mergedVertexID -= mergedSegmentCount;
                mergedSegmentCount = joinCount;
                rads = radsPerJoinSegment;
                if ((cf & JOIN_TYPE_MASK) > ROUND_JOIN_CONTOUR_FLAG)
                {
                    if (mergedVertexID < 2.5f)
                        cf |= JOIN_TANGENT_0_CONTOUR_FLAG;
                    if (mergedVertexID > 1.5f && mergedVertexID < 3.5f)
                        cf |= JOIN_TANGENT_INNER_CONTOUR_FLAG;
                }
                else if ((cf & EMULATED_STROKE_CAP_CONTOUR_FLAG) != 0u ||
                         (cf & JOIN_TYPE_MASK) == FEATHER_JOIN_CONTOUR_FLAG)
                {
                    mergedSegmentCount -= 2.f;
                    mergedVertexID -= 1.f;
                }
                cf |= rads < 0.f ? LEFT_JOIN_CONTOUR_FLAG : RIGHT_JOIN_CONTOUR_FLAG;
            }

            F2 tessCoord;
            float theta = 0.f;
            if (mergedVertexID == 0.f || mergedVertexID == mergedSegmentCount ||
                (cf & JOIN_TYPE_MASK) > ROUND_JOIN_CONTOUR_FLAG)
            {
                bool isTan0 = mergedVertexID < mergedSegmentCount * .5f;
                tessCoord = isTan0 ? p0 : p3;
                theta = atan2v(isTan0 ? tans[0] : tans[1]);
            }
            else if ((cf & RETROFIT_TRI_STRIP_CONTOUR_FLAG) != 0u)
            {
                tessCoord = p0; // outerCurve retrofits are not produced here
            }
            else
            {
                float T, polarT;
                if (parametric == mergedSegmentCount)
                {
                    T = mergedVertexID / parametric;
                    polarT = 0.f;
                }
                else
                {
                    F2 C = sub(p1, p0);
                    F2 D = sub(p3, p0);
                    F2 E = sub(p2, p1);
                    F2 B = sub(E, C);
                    F2 A{-3.f * E.x + D.x, -3.f * E.y + D.y};
                    F2 B_{B.x * (parametric * 2.f), B.y * (parametric * 2.f)};
                    float pp = parametric * parametric;
                    F2 C_{C.x * pp, C.y * pp};
                    float lastParametricVertexID = 0.f;
                    float maxParametricVertexID = std::min(parametric - 1.f, mergedVertexID);
                    F2 tan0norm = normalize2(tans[0]);
                    float negAbsRadsPerSegment = -fabsf(rads);
                    float maxRotation0 = (1.f + mergedVertexID) * fabsf(rads);
                    for (int p = 9; p >= 0; --p)
                    {
                        float testParametricID = lastParametricVertexID + exp2f(float(p));
                        if (testParametricID <= maxParametricVertexID)
                        {
                            F2 testTan{testParametricID * A.x + B_.x,
                                       testParametricID * A.y + B_.y};
                            testTan = F2{testParametricID * testTan.x + C_.x,
                                         testParametricID * testTan.y + C_.y};
                            float cosRotation = dot2(normalize2(testTan), tan0norm);
                            float maxRotation =
                                testParametricID * negAbsRadsPerSegment + maxRotation0;
                            maxRotation = std::min(maxRotation, PI_);
                            if (cosRotation >= cosf(maxRotation))
                                lastParametricVertexID = testParametricID;
                        }
                    }
                    float parametricT = lastParametricVertexID / parametric;
                    float lastPolarVertexID = mergedVertexID - lastParametricVertexID;
                    float theta0 = acosf(clampf(tan0norm.x, -1.f, 1.f));
                    theta0 = tan0norm.y >= 0.f ? theta0 : -theta0;
                    theta = lastPolarVertexID * rads + theta0;
                    F2 norm{sinf(theta), -cosf(theta)};
                    float a = dot2(norm, A), b_over_2 = dot2(norm, B), c = dot2(norm, C);
                    float discr_over_4 = std::max(b_over_2 * b_over_2 - a * c, 0.f);
                    float q = sqrtf(discr_over_4);
                    if (b_over_2 > 0.f)
                        q = -q;
                    q -= b_over_2;
                    float _5qa = -.5f * q * a;
                    F2 root = (fabsf(q * q + _5qa) < fabsf(a * c + _5qa)) ? F2{q, a} : F2{c, q};
                    polarT = (root.y != 0.f) ? root.x / root.y : 0.f;
                    polarT = clampf(polarT, 0.f, 1.f);
                    if (lastPolarVertexID == 0.f)
                        polarT = 0.f;
                    T = std::max(parametricT, polarT);
                }
                F2 ab = umix(p0, p1, T);
                F2 bc = umix(p1, p2, T);
                F2 cd = umix(p2, p3, T);
                F2 abc = umix(ab, bc, T);
                F2 bcd = umix(bc, cd, T);
                tessCoord = umix(abc, bcd, T);
                if (T != polarT)
                    theta = atan2v(sub(bcd, abc));
            }
            TessVertex out;
            out.x = tessCoord.x;
            out.y = tessCoord.y;
            if ((cf & JOIN_TYPE_MASK) == FEATHER_JOIN_CONTOUR_FLAG)
                out.z = ((uint32_t)mergedSegmentCount << 16) | (uint32_t)mergedVertexID;
            else
                out.z = fbits(glsl_mod(theta, _2PI_));
            out.contourIDWithFlags = cf;
            texture[(size_t)row * kTessTextureWidth + (uint32_t)x] = out;
        }
    }
}
} // namespace oracle
