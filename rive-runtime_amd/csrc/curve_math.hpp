// Curve maths of the reference's CPU path (bezier_utils.cpp, wangs_formula.hpp,
// simd.hpp, draw.cpp), shared by the device's counting kernels and the host's
// softening of feathered paths, so both see the same chops.  Plain IEEE fp32
// with no contraction (every file including it is built with
// -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define RIVE_HD __host__ __device__

namespace rivehip
{
RIVE_HD inline uint32_t float_bits(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}
struct V2d
{
    float x, y;
};
RIVE_HD __forceinline__ V2d operator+(V2d a, V2d b) { return {a.x + b.x, a.y + b.y}; }
RIVE_HD __forceinline__ V2d operator-(V2d a, V2d b) { return {a.x - b.x, a.y - b.y}; }
RIVE_HD __forceinline__ V2d operator*(V2d a, float s) { return {a.x * s, a.y * s}; }
RIVE_HD __forceinline__ V2d operator*(float s, V2d a) { return {s * a.x, s * a.y}; }
RIVE_HD __forceinline__ bool veq(V2d a, V2d b) { return a.x == b.x && a.y == b.y; }
RIVE_HD __forceinline__ bool vne(V2d a, V2d b) { return !veq(a, b); }
RIVE_HD __forceinline__ V2d vmix(V2d a, V2d b, float t) { return (b - a) * t + a; } // simd.hpp:746-753
RIVE_HD __forceinline__ float vcross(V2d a, V2d b) { return a.x * b.y - a.y * b.x; }
RIVE_HD __forceinline__ float vdot(V2d a, V2d b) { return a.x * b.x + a.y * b.y; }
// simd::clamp ignores NaN (simd.hpp:211-254): fminf/fmaxf have the same rule.
RIVE_HD __forceinline__ float sclamp(float x, float lo, float hi) { return fminf(fmaxf(lo, x), hi); }

constexpr float kTessEpsilon = 1.f / (1 << 10); // bezier_utils.cpp:178
constexpr float kMathEpsilon = 1.f / (1 << 12); // math_types.hpp:20-21
constexpr float kMathPI = 3.14159265f;          // math_types.hpp:18
constexpr uint32_t kMiterBevelJoinSegments = 5;  // draw.cpp:40
constexpr float kPolarPrecisionF = 8.f;

RIVE_HD __forceinline__ float fast_acos_d(float x) // simd.hpp:496-506
{
    const float a = -0.939115566365855f;
    const float b = 0.9217841528914573f;
    const float c = -1.2845906244690837f;
    const float d = 0.295624144969963174f;
    const float pi_over_2 = 1.5707963267948966f;
    float xx = x * x;
    float numer = b * xx + a;
    float denom = xx * (d * xx + c) + 1.f;
    return x * (numer / denom) + pi_over_2;
}

// wangs_formula.hpp:157-169 (cubic, raised to the 4th power).
RIVE_HD __forceinline__ float wang_pow4(const V2d p[4], float precision, const float m[6])
{
    float v0 = (-2.f * p[1].x + p[0].x) + p[2].x;
    float v1 = (-2.f * p[1].y + p[0].y) + p[2].y;
    float v2 = (-2.f * p[2].x + p[1].x) + p[3].x;
    float v3 = (-2.f * p[2].y + p[1].y) + p[3].y;
    float w0 = m[0] * v0 + m[2] * v1;
    float w1 = m[3] * v1 + m[1] * v0;
    float w2 = m[0] * v2 + m[2] * v3;
    float w3 = m[3] * v3 + m[1] * v2;
    const float k = ((3 * 3) * (2 * 2) / 64.f) * (precision * precision);
    return fmaxf(w0 * w0 + w1 * w1, w2 * w2 + w3 * w3) * k;
}
RIVE_HD __forceinline__ uint32_t parametric_count(float n4) // draw.cpp:1195-1203
{
    float n = ceilf(sqrtf(sqrtf(n4)));
    n = sclamp(n, 1.f, 1023.f);
    return (uint32_t)n;
}
RIVE_HD __forceinline__ uint32_t polar_count(V2d t0, V2d t1, float spr) // draw.cpp:1219-1238
{
    float numer = t0.x * t1.x + t0.y * t1.y;
    float denom_pow2 = (t0.x * t0.x + t0.y * t0.y) * (t1.x * t1.x + t1.y * t1.y);
    float cosTheta = numer / sqrtf(denom_pow2);
    cosTheta = sclamp(cosTheta, -1.f, 1.f);
    float theta = fast_acos_d(cosTheta);
    float n = ceilf(theta * spr);
    n = sclamp(n, 1.f, 1023.f);
    return (uint32_t)n;
}
// feather_join_segment_count (draw.cpp:330-340): half a turn of polar
// segments plus the 3 helper vertices' 4 segments, at least 6.
RIVE_HD __forceinline__ uint32_t feather_join_segments(float spr)
{
    const uint32_t n = (uint32_t)ceilf(spr * kMathPI) + 4u;
    return n > 6u ? n : 6u;
}
RIVE_HD __forceinline__ V2d tan0(const V2d p[4]) // bezier_utils.hpp:147-151
{
    return (vne(p[0], p[1]) ? p[1] : vne(p[1], p[2]) ? p[2] : p[3]) - p[0];
}
RIVE_HD __forceinline__ V2d tan1(const V2d p[4])
{
    return p[3] - (vne(p[3], p[2]) ? p[2] : vne(p[2], p[1]) ? p[1] : p[0]);
}

RIVE_HD inline void chop_at(const V2d* src, V2d* dst, float t) // bezier_utils.cpp:34-60
{
    if (t == 1.f)
    {
        V2d s0 = src[0], s1 = src[1], s2 = src[2], s3 = src[3];
        dst[0] = s0;
        dst[1] = s1;
        dst[2] = s2;
        dst[3] = s3;
        dst[4] = dst[5] = dst[6] = src[3];
        return;
    }
    V2d p0 = src[0], p1 = src[1], p2 = src[2], p3 = src[3];
    V2d ab = vmix(p0, p1, t), bc = vmix(p1, p2, t), cd = vmix(p2, p3, t);
    V2d abc = vmix(ab, bc, t), bcd = vmix(bc, cd, t);
    V2d abcd = vmix(abc, bcd, t);
    dst[0] = p0;
    dst[1] = ab;
    dst[2] = abc;
    dst[3] = abcd;
    dst[4] = bcd;
    dst[5] = cd;
    dst[6] = p3;
}
RIVE_HD inline void chop_at2(const V2d* src, V2d* dst, float t0, float t1) // bezier_utils.cpp:62-101
{
    if (t1 == 1.f)
    {
        chop_at(src, dst, t0);
        dst[7] = dst[8] = dst[9] = src[3]; // src may alias dst, as in the reference
        return;
    }
    V2d p0 = src[0], p1 = src[1], p2 = src[2], p3 = src[3];
    V2d ab0 = vmix(p0, p1, t0), ab1 = vmix(p0, p1, t1);
    V2d bc0 = vmix(p1, p2, t0), bc1 = vmix(p1, p2, t1);
    V2d cd0 = vmix(p2, p3, t0), cd1 = vmix(p2, p3, t1);
    V2d abc0 = vmix(ab0, bc0, t0), abc1 = vmix(ab1, bc1, t1);
    V2d bcd0 = vmix(bc0, cd0, t0), bcd1 = vmix(bc1, cd1, t1);
    V2d abcd0 = vmix(abc0, bcd0, t0), abcd1 = vmix(abc1, bcd1, t1);
    V2d mid0 = vmix(abc0, bcd0, t1), mid1 = vmix(abc1, bcd1, t0);
    dst[0] = p0;
    dst[1] = ab0;
    dst[2] = abc0;
    dst[3] = abcd0;
    dst[4] = mid0;
    dst[5] = mid1;
    dst[6] = abcd1;
    dst[7] = bcd1;
    dst[8] = cd1;
    dst[9] = p3;
}
RIVE_HD inline void chop_at_n(const V2d* src, V2d* dst, const float* tv, int n) // bezier_utils.cpp:103-158
{
    if (n == 0)
    {
        V2d s0 = src[0], s1 = src[1], s2 = src[2], s3 = src[3];
        dst[0] = s0;
        dst[1] = s1;
        dst[2] = s2;
        dst[3] = s3;
        return;
    }
    int i = 0;
    float lastT = 0.f;
    for (; i < n - 1; i += 2)
    {
        float tt0 = sclamp((tv[i] - lastT) / (1.f - lastT), 0.f, 1.f);
        float tt1 = sclamp((tv[i + 1] - lastT) / (1.f - lastT), 0.f, 1.f);
        lastT = tv[i + 1];
        chop_at2(src, dst, tt0, tt1);
        src = dst = dst + 6;
    }
    if (i < n)
    {
        float t = sclamp((tv[i] - lastT) / (1.f - lastT), 0.f, 1.f);
        chop_at(src, dst, t);
    }
}
RIVE_HD inline V2d eval_cubic(const V2d p[4], float t) // bezier_utils.cpp:21-32
{
    V2d a = (p[3] + 3.f * (p[1] - p[2])) - p[0];
    V2d b = 3.f * ((p[2] - 2.f * p[1]) + p[0]);
    V2d c = 3.f * (p[1] - p[0]);
    return ((a * t + b) * t + c) * t + p[0];
}
RIVE_HD inline int convex_180_chops(const V2d pts[4], float T[2], bool* areCusps) // bezier_utils.cpp:180-330
{
    const uint32_t kOneMinus2Eps = (127u << 23) - 2u * (1u << (24 - 10));
    V2d p0 = pts[0], p1 = pts[1], p2 = pts[2], p3 = pts[3];
    V2d C = p1 - p0;
    V2d D = p2 - p1;
    V2d E = p3 - p0;
    V2d B = D - C;
    V2d A = -3.f * D + E;
    float a = vcross(A, B);
    float b = vcross(A, C);
    float c = vcross(B, C);
    float b_over_minus_2 = -.5f * b;
    float discr_over_4 = b_over_minus_2 * b_over_minus_2 - a * c;
    float cuspThreshold = a * (kTessEpsilon / 2);
    cuspThreshold *= cuspThreshold;
    if (discr_over_4 < -cuspThreshold)
    {
        *areCusps = false;
        float root = c / b_over_minus_2;
        if (float_bits(root - kTessEpsilon) < kOneMinus2Eps)
        {
            T[0] = root;
            return 1;
        }
        return 0;
    }
    *areCusps = discr_over_4 <= cuspThreshold;
    if (*areCusps)
    {
        if (a != 0.f || b_over_minus_2 != 0.f || c != 0.f)
        {
            float root = b_over_minus_2 / a;
            if (float_bits(root - kTessEpsilon) < kOneMinus2Eps)
            {
                T[0] = root;
                return 1;
            }
            *areCusps = false;
            return 0;
        }
        V2d base = p3 - p0;
        float dp0 = pts[0].x * base.x + pts[0].y * base.y;
        float dp1 = pts[1].x * base.x + pts[1].y * base.y;
        float dp2 = pts[2].x * base.x + pts[2].y * base.y;
        float dp3 = pts[3].x * base.x + pts[3].y * base.y;
        if (dp1 > dp0 && dp2 > dp1 && dp3 > dp2)
        {
            *areCusps = false;
            return 0;
        }
        V2d t0v = (C.x != 0.f || C.y != 0.f) ? C : p2 - p0;
        a = vdot(t0v, A);
        b_over_minus_2 = -vdot(t0v, B);
        c = vdot(t0v, C);
        discr_over_4 = fmaxf(b_over_minus_2 * b_over_minus_2 - a * c, 0.f);
    }
    float q = sqrtf(discr_over_4);
    q = copysignf(q, b_over_minus_2);
    q = q + b_over_minus_2;
    float r0 = q / a, r1 = c / q;
    bool in0 = r0 > kTessEpsilon && r0 < (1.f - kTessEpsilon);
    bool in1 = r1 > kTessEpsilon && r1 < (1.f - kTessEpsilon);
    if (in0)
    {
        if (in1 && r0 != r1)
        {
            if (r0 > r1)
            {
                float t = r0;
                r0 = r1;
                r1 = t;
            }
            T[0] = r0;
            T[1] = r1;
            return 2;
        }
        T[0] = r0;
        return 1;
    }
    if (in1)
    {
        T[0] = r1;
        return 1;
    }
    return 0;
}
RIVE_HD inline void chop_around_cusps(const V2d p[4], V2d* dst, const float* cuspT, int n,
                                  float matrixMaxScale) // draw.cpp:139-174
{
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < n; ++i)
    {
        float minT = i == 0 ? 0.f : (cuspT[i - 1] + cuspT[i]) * .5f;
        float maxT = i + 1 == n ? 1.f : (cuspT[i + 1] + cuspT[i]) * .5f;
        t[i * 2 + 0] = fmaxf(cuspT[i] - kMathEpsilon, minT);
        t[i * 2 + 1] = fminf(cuspT[i] + kMathEpsilon, maxT);
    }
    chop_at_n(p, dst, t, n * 2);
    for (int i = 0; i < n; ++i)
    {
        V2d* chops = dst + i * 6;
        V2d cusp = eval_cubic(p, cuspT[i]);
        chops[3] = chops[6] = cusp;
        V2d pivot = (chops[2] + chops[7]) * .5f;
        V2d dd = cusp - pivot;
        float len2 = dd.x * dd.x + dd.y * dd.y;
        float scale = len2 > 0.f ? (1.f / sqrtf(len2)) : 1.f;
        V2d nrm = dd * scale;
        float denom = matrixMaxScale * kPolarPrecisionF * 2.f;
        pivot = V2d{nrm.x / denom, nrm.y / denom} + cusp;
        chops[4] = chops[5] = pivot;
    }
}

// Chops a stroked cubic into convex pieces (draw.cpp:1062-1101).  Returns the
// number of sub-cubics written to out[0..3n].
RIVE_HD inline int stroke_subcubics(const V2d* p, float maxScale, V2d* out /*[16]*/)
{
    float t[2] = {0.f, 0.f};
    bool areCusps = false;
    int numChops = convex_180_chops(p, t, &areCusps);
    if (areCusps && numChops > 0)
    {
        chop_around_cusps(p, out, t, numChops, maxScale);
        return numChops * 2 + 1;
    }
    if (numChops == 2)
    {
        chop_at2(p, out, t[0], t[1]);
        return 3;
    }
    if (numChops == 1)
    {
        chop_at(p, out, t[0]);
        return 2;
    }
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
    out[3] = p[3];
    return 1;
}

} // namespace rivehip
