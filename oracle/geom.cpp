// CPU ORACLE (test infrastructure only): per-curve integer counting, record
// emission, span wrapping and the tessellation shader, restated from the
// reference.  Compiled with -ffp-contract=off.  See oracle.hpp.
#include "oracle.hpp"

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstring>

namespace oracle
{
namespace
{
constexpr float kPI = 3.14159265f;                 // math_types.hpp:18
constexpr float kEPSILON = 1.f / (1 << 12);        // math_types.hpp:20-21
constexpr float TESS_EPSILON = 1.f / (1 << 10);    // bezier_utils.cpp:178
constexpr uint32_t NUM_SEGMENTS_IN_MITER_OR_BEVEL_JOIN = 5; // draw.cpp:40

inline uint32_t fbits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
inline float ubits(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}
// simd::min/max ignore NaN operands (simd.hpp:211-243); simd::clamp is
// min(max(lo, x), hi) (simd.hpp:251-254).
inline float smin(float a, float b) { return (b < a || std::isnan(a)) ? b : a; }
inline float smax(float a, float b) { return (a < b || std::isnan(a)) ? b : a; }
inline float sclamp(float x, float lo, float hi) { return smin(smax(lo, x), hi); }
inline float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; } // simd.hpp:732-736
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }   // simd.hpp:727-730
inline V2 mix(V2 a, V2 b, float t) { return (b - a) * t + a; }  // simd.hpp:746-753
} // namespace

float fast_acos(float x)
{
    // simd.hpp:496-506 (a rational approximation with published coefficients).
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

float wang_cubic_pow4(const V2 p[4], float precision, const Mat& m)
{
    // wangs_formula.hpp:157-169 with VectorXform (:86-111).
    float v[4];
    v[0] = (-2.f * p[1].x + p[0].x) + p[2].x;
    v[1] = (-2.f * p[1].y + p[0].y) + p[2].y;
    v[2] = (-2.f * p[2].x + p[1].x) + p[3].x;
    v[3] = (-2.f * p[2].y + p[1].y) + p[3].y;
    // m_scale = [xx, yy], m_skew = [yx, xy]
    float w[4];
    w[0] = m.m[0] * v[0] + m.m[2] * v[1];
    w[1] = m.m[3] * v[1] + m.m[1] * v[0];
    w[2] = m.m[0] * v[2] + m.m[2] * v[3];
    w[3] = m.m[3] * v[3] + m.m[1] * v[2];
    float vv[4] = {w[0] * w[0], w[1] * w[1], w[2] * w[2], w[3] * w[3]};
    // length_term_pow2<3>(precision) (wangs_formula.hpp:41-45)
    const float k = ((3 * 3) * (2 * 2) / 64.f) * (precision * precision);
    return std::max(vv[0] + vv[1], vv[2] + vv[3]) * k;
}

float wang_cubic(const V2 p[4], float precision, const Mat& m)
{
    return sqrtf(sqrtf(wang_cubic_pow4(p, precision, m)));
}

float calc_polar_segments_per_radian(float approxDevStrokeRadius)
{
    // bezier_utils.hpp:108-113, Precision = kPolarPrecision (8).
    float cosTheta = 1.f - (1.f / kPolarPrecision) / approxDevStrokeRadius;
    return .5f / acosf(std::max(cosTheta, -1.f));
}

float find_max_scale(const Mat& M)
{
    // src/math/mat2d_find_max_scale.cpp:24-67.
    const float xx = M.m[0], xy = M.m[1], yx = M.m[2], yy = M.m[3];
    if (xy == 0 && yx == 0)
    {
        return std::max(fabsf(xx), fabsf(yy));
    }
    float a = xx * xx + xy * xy;
    float b = xx * yx + yy * xy;
    float c = yx * yx + yy * yy;
    float bSqd = b * b;
    float result;
    if (bSqd <= kEPSILON * kEPSILON)
    {
        result = std::max(a, c);
    }
    else
    {
        float aminusc = a - c;
        float apluscdiv2 = (a + c) * .5f;
        float x = sqrtf(aminusc * aminusc + 4 * bSqd) * .5f;
        result = apluscdiv2 + x;
    }
    if (!std::isfinite(result))
    {
        result = 0;
    }
    result = std::max(result, 0.f);
    return sqrtf(result);
}

V2 eval_cubic_at(const V2 p[4], float t)
{
    // bezier_utils.cpp:21-32
    V2 a = (p[3] + 3.f * (p[1] - p[2])) - p[0];
    V2 b = 3.f * ((p[2] - 2.f * p[1]) + p[0]);
    V2 c = 3.f * (p[1] - p[0]);
    V2 d = p[0];
    return ((a * t + b) * t + c) * t + d;
}

void chop_cubic_at(const V2 src[4], V2 dst[7], float t)
{
    // bezier_utils.cpp:34-60.  Loads happen before stores so that src may
    // alias dst, as in the reference.
    if (t == 1)
    {
        V2 s[4] = {src[0], src[1], src[2], src[3]};
        memmove(dst, s, sizeof(s));
        dst[4] = dst[5] = dst[6] = src[3];
        return;
    }
    V2 p0 = src[0], p1 = src[1], p2 = src[2], p3 = src[3];
    V2 ab = mix(p0, p1, t), bc = mix(p1, p2, t), cd = mix(p2, p3, t);
    V2 abc = mix(ab, bc, t), bcd = mix(bc, cd, t);
    V2 abcd = mix(abc, bcd, t);
    dst[0] = p0;
    dst[1] = ab;
    dst[2] = abc;
    dst[3] = abcd;
    dst[4] = bcd;
    dst[5] = cd;
    dst[6] = p3;
}

void chop_cubic_at2(const V2 src[4], V2 dst[10], float t0, float t1)
{
    // bezier_utils.cpp:62-101
    if (t1 == 1)
    {
        chop_cubic_at(src, dst, t0);
        // NOTE: src may alias dst here; the reference reads src[3] after the
        // chop above has stored into dst.
        dst[7] = dst[8] = dst[9] = src[3];
        return;
    }
    V2 p0 = src[0], p1 = src[1], p2 = src[2], p3 = src[3];
    V2 ab0 = mix(p0, p1, t0), ab1 = mix(p0, p1, t1);
    V2 bc0 = mix(p1, p2, t0), bc1 = mix(p1, p2, t1);
    V2 cd0 = mix(p2, p3, t0), cd1 = mix(p2, p3, t1);
    V2 abc0 = mix(ab0, bc0, t0), abc1 = mix(ab1, bc1, t1);
    V2 bcd0 = mix(bc0, cd0, t0), bcd1 = mix(bc1, cd1, t1);
    V2 abcd0 = mix(abc0, bcd0, t0), abcd1 = mix(abc1, bcd1, t1);
    // middle = mix(abc, bcd, T.zwxy): lanes xy use t1 on the t0-chop's abc/bcd,
    // lanes zw use t0 on the t1-chop's abc/bcd.
    V2 middle0 = mix(abc0, bcd0, t1);
    V2 middle1 = mix(abc1, bcd1, t0);
    dst[0] = p0;
    dst[1] = ab0;
    dst[2] = abc0;
    dst[3] = abcd0;
    dst[4] = middle0;
    dst[5] = middle1;
    dst[6] = abcd1;
    dst[7] = bcd1;
    dst[8] = cd1;
    dst[9] = p3;
}

void chop_cubic_at_n(const V2 src[4], V2 dst[], const float tValues[], int tCount)
{
    // bezier_utils.cpp:103-158
    if (tCount == 0)
    {
        V2 s[4] = {src[0], src[1], src[2], src[3]};
        memmove(dst, s, sizeof(s));
        return;
    }
    int i = 0;
    float lastT = 0;
    for (; i < tCount - 1; i += 2)
    {
        float tt0 = sclamp((tValues[i] - lastT) / (1 - lastT), 0, 1);
        float tt1 = sclamp((tValues[i + 1] - lastT) / (1 - lastT), 0, 1);
        lastT = tValues[i + 1];
        chop_cubic_at2(src, dst, tt0, tt1);
        src = dst = dst + 6;
    }
    if (i < tCount)
    {
        float t = tValues[i];
        t = sclamp((t - lastT) / (1 - lastT), 0, 1);
        chop_cubic_at(src, dst, t);
    }
}

V2 find_cubic_tan0(const V2 p[4])
{
    return (p[0] != p[1] ? p[1] : p[1] != p[2] ? p[2] : p[3]) - p[0];
}
V2 find_cubic_tan1(const V2 p[4])
{
    return p[3] - (p[3] != p[2] ? p[2] : p[2] != p[1] ? p[1] : p[0]);
}

int find_cubic_convex_180_chops(const V2 pts[4], float T[2], bool* areCusps)
{
    // bezier_utils.cpp:180-330
    const uint32_t kIEEE_one_minus_2_epsilon = (127 << 23) - 2 * (1 << (24 - 10));
    V2 p0 = pts[0], p1 = pts[1], p2 = pts[2], p3 = pts[3];
    // CubicCoeffs (bezier_utils.hpp:44-52)
    V2 C = p1 - p0;
    V2 D = p2 - p1;
    V2 E = p3 - p0;
    V2 B = D - C;
    V2 A = -3.f * D + E;

    float a = cross(A, B);
    float b = cross(A, C);
    float c = cross(B, C);
    float b_over_minus_2 = -.5f * b;
    float discr_over_4 = b_over_minus_2 * b_over_minus_2 - a * c;
    float cuspThreshold = a * (TESS_EPSILON / 2);
    cuspThreshold *= cuspThreshold;

    if (discr_over_4 < -cuspThreshold)
    {
        *areCusps = false;
        float root = c / b_over_minus_2;
        if (fbits(root - TESS_EPSILON) < kIEEE_one_minus_2_epsilon)
        {
            T[0] = root;
            return 1;
        }
        return 0;
    }

    *areCusps = discr_over_4 <= cuspThreshold;
    if (*areCusps)
    {
        if (a != 0 || b_over_minus_2 != 0 || c != 0)
        {
            float root = b_over_minus_2 / a;
            if (fbits(root - TESS_EPSILON) < kIEEE_one_minus_2_epsilon)
            {
                T[0] = root;
                return 1;
            }
            *areCusps = false;
            return 0;
        }
        V2 base = p3 - p0;
        float dp[4];
        for (int i = 0; i < 4; ++i)
            dp[i] = pts[i].x * base.x + pts[i].y * base.y;
        if (dp[1] > dp[0] && dp[2] > dp[1] && dp[3] > dp[2])
        {
            *areCusps = false;
            return 0;
        }
        V2 tan0 = (C.x != 0.f || C.y != 0.f) ? C : p2 - p0;
        a = dot(tan0, A);
        b_over_minus_2 = -dot(tan0, B);
        c = dot(tan0, C);
        discr_over_4 = std::max(b_over_minus_2 * b_over_minus_2 - a * c, 0.f);
    }

    float q = sqrtf(discr_over_4);
    q = copysignf(q, b_over_minus_2);
    q = q + b_over_minus_2;
    float roots[2] = {q / a, c / q};
    bool inside0 = roots[0] > TESS_EPSILON && roots[0] < (1 - TESS_EPSILON);
    bool inside1 = roots[1] > TESS_EPSILON && roots[1] < (1 - TESS_EPSILON);
    if (inside0)
    {
        if (inside1 && roots[0] != roots[1])
        {
            if (roots[0] > roots[1])
                std::swap(roots[0], roots[1]);
            T[0] = roots[0];
            T[1] = roots[1];
            return 2;
        }
        T[0] = roots[0];
        return 1;
    }
    if (inside1)
    {
        T[0] = roots[1];
        return 1;
    }
    return 0;
}

void chop_cubic_around_cusps(const V2 p[4], V2 dst[], const float cuspT[], int n,
                             float matrixMaxScale)
{
    // draw.cpp:139-174
    float t[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i)
    {
        float minT = i == 0 ? 0.f : (cuspT[i - 1] + cuspT[i]) * .5f;
        float maxT = i + 1 == n ? 1.f : (cuspT[i + 1] + cuspT[i]) * .5f;
        t[i * 2 + 0] = fmaxf(cuspT[i] - kEPSILON, minT);
        t[i * 2 + 1] = fminf(cuspT[i] + kEPSILON, maxT);
    }
    chop_cubic_at_n(p, dst, t, n * 2);
    for (int i = 0; i < n; ++i)
    {
        V2* chops = dst + i * 6;
        V2 cusp = eval_cubic_at(p, cuspT[i]);
        chops[3] = chops[6] = cusp;
        V2 pivot = (chops[2] + chops[7]) * .5f;
        V2 d = cusp - pivot;
        // Vec2D::normalized() (src/math/vec2d.cpp:20-25)
        float len2 = d.x * d.x + d.y * d.y;
        float scale = len2 > 0 ? (1 / std::sqrt(len2)) : 1;
        V2 nrm = d * scale;
        float denom = matrixMaxScale * kPolarPrecision * 2;
        pivot = V2{nrm.x / denom, nrm.y / denom} + cusp;
        chops[4] = chops[5] = pivot;
    }
}

void derive_directions(PathDrawDesc* d)
{
    // draw.cpp:651-719 for CoverageType::pixelLocalStorage (rasterOrdering),
    // clockwiseFillOverride = false.
    d->contourFlags = 0;
    if (d->strokeRadius > 0)
    {
        d->dirs = Dirs::forward;
    }
    else if (d->fillRule == FillRule::clockwise)
    {
        const float* m = d->matrix.m;
        float det = m[0] * m[3] - m[2] * m[1];
        if (det < 0)
        {
            d->dirs = Dirs::forwardThenReverse;
            d->contourFlags |= NEGATE_PATH_FILL_COVERAGE_FLAG;
        }
        else
        {
            d->dirs = Dirs::reverseThenForward;
        }
    }
    else
    {
        d->dirs = Dirs::reverseThenForward;
    }
}

namespace
{
// Raw-path iteration helpers (include/rive/math/raw_path.hpp:104-200):
// move/line consume 1 point, quad 2, cubic 3, close 0.  Line/cubic point
// accessors back up one point to include the previous endpoint.
int pts_advance(uint8_t v)
{
    switch (v)
    {
        case kMove: return 1;
        case kLine: return 1;
        case kQuad: return 2;
        case kCubic: return 3;
        default: return 0;
    }
}

V2 find_starting_tangent(const V2* pts, const V2* end) // draw.cpp:178-191
{
    const V2 p0 = pts[0];
    while (++pts < end)
    {
        if (*pts != p0)
            return *pts - p0;
    }
    return {1, 0};
}
V2 find_ending_tangent(const V2* pts, const V2* end) // draw.cpp:195-208
{
    const V2 endpoint = end[-1];
    while (--end > pts)
    {
        V2 p = end[-1];
        if (p != endpoint)
            return endpoint - p;
    }
    return {-1, 0};
}
V2 find_join_tangent(const V2* joinPoint, const V2* end, bool closed,
                     const V2* p0) // draw.cpp:210-251
{
    const V2* nextPoint = joinPoint + 1;
    nextPoint = nextPoint != end ? nextPoint : p0;
    V2 tangent = *nextPoint - *joinPoint;
    if (tangent != V2{0, 0})
        return tangent;
    for (const V2* p = joinPoint + 1; p != end; ++p)
        if (*p != *joinPoint)
            return *p - *joinPoint;
    if (closed)
        for (const V2* p = p0; p != joinPoint; ++p)
            if (*p != *joinPoint)
                return *p - *joinPoint;
    return {0, 0}; // unreachable for well-formed paths
}
Cap empty_stroke_cap(bool closed, Join join, Cap cap) // draw.cpp:253-281
{
    if (closed)
    {
        switch (join)
        {
            case Join::round: return Cap::round;
            case Join::miter: return Cap::square;
            case Join::bevel: return Cap::butt;
        }
    }
    return cap;
}
uint32_t join_type_flags(Join join) // draw.cpp:289-302
{
    switch (join)
    {
        case Join::miter: return MITER_REVERT_JOIN_CONTOUR_FLAG;
        case Join::round: return ROUND_JOIN_CONTOUR_FLAG;
        case Join::bevel: return BEVEL_JOIN_CONTOUR_FLAG;
    }
    return 0;
}
uint32_t parametric_count_from_pow4(float n4)
{
    // draw.cpp:1195-1203: ceil(sqrt(sqrt(n4))), clamped to [1, 1023].
    float n = ceilf(sqrtf(sqrtf(n4)));
    n = sclamp(n, 1, (float)kMaxParametricSegments);
    return (uint32_t)n;
}
// feather_join_segment_count (draw.cpp:330-340): half a turn of polar
// segments plus the helper vertices.
uint32_t feather_join_segment_count(float polarSegmentsPerRadian)
{
    uint32_t n = (uint32_t)ceilf(polarSegmentsPerRadian * kPI) + (FEATHER_JOIN_HELPER_VERTEX_COUNT + 1);
    return std::max(n, FEATHER_JOIN_MIN_SEGMENT_COUNT);
}

uint32_t polar_count(V2 t0, V2 t1, float polarSegmentsPerRadian)
{
    // draw.cpp:1219-1238
    float numer = t0.x * t1.x + t0.y * t1.y;
    float denom_pow2 = (t0.x * t0.x + t0.y * t0.y) * (t1.x * t1.x + t1.y * t1.y);
    float cosTheta = numer / sqrtf(denom_pow2);
    cosTheta = sclamp(cosTheta, -1, 1);
    float theta = fast_acos(cosTheta);
    float n = ceilf(theta * polarSegmentsPerRadian);
    n = sclamp(n, 1, (float)kMaxPolarSegments);
    return (uint32_t)n;
}
std::array<V2, 4> line_to_cubic(V2 p0, V2 p1) // draw.cpp:110-121
{
    // mix(endPts, endPts.zwxy, 1/3): control points at p0 + (p1-p0)/3 and
    // p1 + (p0-p1)/3.
    const float third = 1 / 3.f;
    std::array<V2, 4> c;
    c[0] = p0;
    c[1] = (p1 - p0) * third + p0;
    c[2] = (p0 - p1) * third + p1;
    c[3] = p1;
    return c;
}

struct ContourInfo // draw.hpp:445-462
{
    uint32_t endVerb, endPt; // iterator at the end of the contour
    uint32_t startVerb, startPt;
    size_t endLineIdx;
    size_t firstCurveIdx, endCurveIdx;
    size_t firstRotationIdx, endRotationIdx;
    V2 midpoint;
    bool closed;
    size_t strokeJoinCount;
    uint32_t strokeCapSegmentCount;
    uint32_t paddingVertexCount;
};

// Writer that assigns tessellation locations (render_context.cpp:3079-3229).
struct Writer
{
    Dirs dirs = Dirs::forward;
    uint32_t pathTessLocation = 0;
    uint32_t pathMirroredTessLocation = 0;
    uint32_t nextCubicPaddingVertexCount = 0;
    std::vector<CubicRecord>* out;

    uint32_t nextVertexIndex() const
    {
        return dirs != Dirs::reverse ? pathTessLocation : pathMirroredTessLocation - 1;
    }
    void pushCubic(const V2 pts[4], V2 joinTangent, uint32_t parametric,
                   uint32_t polar, uint32_t join, uint32_t contourIDWithFlags)
    {
        uint32_t total = nextCubicPaddingVertexCount + parametric + polar + join - 1;
        nextCubicPaddingVertexCount = 0;
        CubicRecord r;
        memcpy(r.pts, pts, sizeof(r.pts));
        r.joinTangent = joinTangent;
        r.parametric = parametric;
        r.polar = polar;
        r.join = join;
        r.contourIDWithFlags = contourIDWithFlags;
        r.totalVertexCount = total;
        r.forwardLocation = ~0u;
        r.mirroredLocation = ~0u;
        if (dirs != Dirs::reverse)
        {
            r.forwardLocation = pathTessLocation;
            pathTessLocation += total;
        }
        if (dirs != Dirs::forward)
        {
            r.mirroredLocation = pathMirroredTessLocation;
            pathMirroredTessLocation -= total;
        }
        out->push_back(r);
    }
};
} // namespace

namespace
{
bool g_interiorTriangulation = false;

// gpu::find_transformed_area (gpu.cpp:1160-1173): the area of the path's
// local bounds mapped to the screen.
float transformed_area(const PathDrawDesc& d)
{
    if (d.ptCount == 0)
        return 0;
    float l = d.pts[0].x, t = d.pts[0].y, r = l, b = t;
    for (uint32_t i = 1; i < d.ptCount; ++i)
    {
        l = std::min(l, d.pts[i].x), r = std::max(r, d.pts[i].x);
        t = std::min(t, d.pts[i].y), b = std::max(b, d.pts[i].y);
    }
    const float* m = d.matrix.m;
    auto map = [m](float x, float y) { return V2{m[0] * x + m[2] * y + m[4], m[1] * x + m[3] * y + m[5]}; };
    const V2 q[4] = {map(l, t), map(r, t), map(r, b), map(l, b)};
    const V2 v[3] = {q[1] - q[0], q[2] - q[0], q[3] - q[0]};
    return (fabsf(cross(v[0], v[1])) + fabsf(cross(v[1], v[2]))) * .5f;
}

// iterateOuterCubics (draw.cpp:2401-2650) for an interior-triangulated fill:
// every line and every cubic (chopped uniformly in T into pieces of at most
// 16 parametric segments) becomes one outer-cubic patch of 16 segments plus a
// bowtie join, flagged to cull the segments it does not need; the gaps
// between a chopped cubic's chord and its pieces' chords are filled by
// triangle strips retrofitted into patches.  The triangulation of the
// interior itself (GrInnerFanTriangulator) is not restated: it adds no
// tessellation besides its grout patches, which are left out (none here).
PathTess tessellate_outer_cubic_records(const PathDrawDesc& d, uint32_t firstContourID, uint32_t tessLocation)
{
    constexpr uint32_t kSpan = 16, kSpanPlusJoin = 17;
    constexpr uint32_t kMaxSubdivisions = (1023 + kSpan - 1) / kSpan;
    PathTess result;
    auto subdivisions = [&](const V2* p) {
        float n = ceilf(wang_cubic(p, kParametricPrecision, d.matrix) * (1.f / kSpan));
        return (uint32_t)sclamp(n, 1, (float)kMaxSubdivisions);
    };
    // counting pass
    uint32_t patchCount = 0, contourCount = 0;
    {
        V2 p0{0, 0};
        const V2* pts = d.pts;
        for (uint32_t v = 0; v < d.verbCount; ++v)
        {
            switch (d.verbs[v])
            {
                case 0: // move
                    if (contourCount != 0 && !(pts[-1] == p0))
                        ++patchCount;
                    p0 = pts[0];
                    ++contourCount;
                    break;
                case 1:
                    ++patchCount;
                    break;
                case 4:
                {
                    const V2 c[4] = {pts[-1], pts[0], pts[1], pts[2]};
                    const uint32_t n = subdivisions(c);
                    patchCount += n + (n - 1 + 2) / 3;
                    break;
                }
                default:
                    break;
            }
            pts += pts_advance(d.verbs[v]);
        }
        if (contourCount != 0 && !(d.pts[d.ptCount - 1] == p0))
            ++patchCount;
    }
    if (patchCount == 0)
        return result;
    result.tessVertexCount = patchCount * kSpanPlusJoin;
    result.midpointFanTessVertexCount = double_sided(d.dirs) ? result.tessVertexCount * 2 : result.tessVertexCount;
    result.maxTessellatedSegmentCount = patchCount;

    Writer w;
    w.dirs = d.dirs;
    w.out = &result.records;
    const uint32_t count = result.midpointFanTessVertexCount;
    switch (d.dirs)
    {
        case Dirs::forward:
            w.pathTessLocation = tessLocation;
            break;
        case Dirs::reverse:
            w.pathMirroredTessLocation = tessLocation + count;
            break;
        case Dirs::reverseThenForward:
            w.pathTessLocation = w.pathMirroredTessLocation = tessLocation + count / 2;
            break;
        case Dirs::forwardThenReverse:
            w.pathTessLocation = tessLocation;
            w.pathMirroredTessLocation = tessLocation + count;
            break;
    }
    uint32_t contourID = firstContourID;
    uint32_t cf = 0;
    auto pushContour = [&]() {
        ContourData cd;
        cd.midpoint = V2{0, 0};
        cd.pathID = d.pathID;
        cd.vertexIndex0 = w.nextVertexIndex();
        result.contours.push_back(cd);
        result.contourFirstRecord.push_back((uint32_t)result.records.size());
        result.contourPadding.push_back(0);
        cf = d.contourFlags | contourID++;
    };
    auto pushPatch = [&](const V2* p) {
        w.pushCubic(p, V2{0, 0}, kSpan, 1, 1, cf | CULL_EXCESS_TESSELLATION_SEGMENTS_CONTOUR_FLAG);
    };
    // pushRetrofitCubicTriStrip (render_context.cpp:3231-3257)
    auto pushStrip = [&](const V2* p, size_t n) {
        const V2 strip[4] = {p[0], p[1], p[std::min<size_t>(3, n - 1)], p[2]};
        w.pushCubic(strip, p[std::min<size_t>(4, n - 1)], kSpan, 1, 1, cf | RETROFIT_TRI_STRIP_CONTOUR_FLAG);
    };
    V2 p0{0, 0};
    uint32_t contours = 0;
    const V2* pts = d.pts;
    for (uint32_t v = 0; v < d.verbCount; ++v)
    {
        switch (d.verbs[v])
        {
            case 0:
                if (contours != 0 && !(pts[-1] == p0))
                {
                    const auto c = line_to_cubic(pts[-1], p0);
                    pushPatch(c.data());
                }
                pushContour();
                p0 = pts[0];
                ++contours;
                break;
            case 1:
            {
                const auto c = line_to_cubic(pts[-1], pts[0]);
                pushPatch(c.data());
                break;
            }
            case 4:
            {
                const V2 c[4] = {pts[-1], pts[0], pts[1], pts[2]};
                const uint32_t n = subdivisions(c);
                if (n == 1)
                {
                    pushPatch(c);
                    break;
                }
                // chop uniformly in T (bezier_utils.cpp:104-158, null tValues)
                std::vector<V2> chops(3 * n + 1);
                {
                    const V2* src = c;
                    V2* dst = chops.data();
                    int i = 0;
                    const int tCount = (int)n - 1;
                    for (; i < tCount - 1; i += 2)
                    {
                        const float k = (float)(tCount + 1 - i);
                        V2 tmp[10];
                        chop_cubic_at2(src, tmp, 1.f / k, 2.f / k);
                        memcpy(dst, tmp, sizeof(tmp));
                        src = dst = dst + 6;
                    }
                    if (i < tCount)
                    {
                        V2 tmp[7];
                        chop_cubic_at(src, tmp, .5f);
                        memcpy(dst, tmp, sizeof(tmp));
                    }
                }
                for (uint32_t i = 0; i < n; ++i)
                    pushPatch(chops.data() + 3 * i);
                for (uint32_t a = 1; a < n; a += 3)
                {
                    const uint32_t remaining = n - a;
                    const uint32_t b = remaining == 1 ? a : a + 1;
                    V2 strip[5];
                    size_t m = 0;
                    strip[m++] = chops[b * 3];
                    strip[m++] = chops[(b + 1) * 3];
                    if (remaining > 1)
                        strip[m++] = chops[(b - 1) * 3];
                    if (remaining > 2)
                        strip[m++] = chops[(b + 2) * 3];
                    strip[m++] = chops[0];
                    pushStrip(strip, m);
                }
                break;
            }
            default:
                break;
        }
        pts += pts_advance(d.verbs[v]);
    }
    if (contours != 0 && !(d.pts[d.ptCount - 1] == p0))
    {
        const auto c = line_to_cubic(d.pts[d.ptCount - 1], p0);
        pushPatch(c.data());
    }
    return result;
}
} // namespace

void set_interior_triangulation(bool enabled) { g_interiorTriangulation = enabled; }

bool uses_interior_triangulation(const PathDrawDesc& d)
{
    // PathDraw::Make (draw.cpp:510-538) with TriangulationController's
    // guards alone (frameBudgetMs = infinity): a plain fill of at least
    // 512 x 512 px on screen and at most 256 verbs.
    return g_interiorTriangulation && d.strokeRadius == 0 && d.featherRadius == 0 && d.verbCount <= 256 &&
           transformed_area(d) >= 512.f * 512.f;
}

PathTess tessellate_path_records(const PathDrawDesc& d, uint32_t firstContourID,
                                 uint32_t tessLocation)
{
    if (uses_interior_triangulation(d))
        return tessellate_outer_cubic_records(d, firstContourID, tessLocation);
    PathTess result;
    const bool isStroke = d.strokeRadius > 0;
    // Feathered fills are tessellated like strokes, with feather joins
    // (draw.cpp:776-806).
    const bool isStrokeOrFeather = isStroke || d.featherRadius > 0;
    const uint8_t* verbs = d.verbs;
    const V2* P = d.pts;
    const uint32_t nVerbs = d.verbCount;

    // ---- initForMidpointFan (draw.cpp:768-1392) ----
    float strokeMatrixMaxScale = 0;
    float polarSegmentsPerRadian = 0;
    if (isStrokeOrFeather)
    {
        strokeMatrixMaxScale = find_max_scale(d.matrix);
        float r_ = 0;
        if (d.featherRadius != 0)
        {
            r_ = d.featherRadius * strokeMatrixMaxScale;
            // The polar density needed at FEATHER_POLAR_SEGMENT_MIN_ANGLE
            // (draw.cpp:784-802).
            const float FEATHER_MAX_SCREEN_SPACE_RADIUS =
                1 / (kPolarPrecision * (1 - COS_FEATHER_POLAR_SEGMENT_MIN_ANGLE_OVER_2));
            r_ = std::min(r_, FEATHER_MAX_SCREEN_SPACE_RADIUS);
        }
        if (isStroke)
            r_ += d.strokeRadius * strokeMatrixMaxScale;
        polarSegmentsPerRadian = calc_polar_segments_per_radian(r_);
    }
    float parametricPrecision = kParametricPrecision;
    if (d.featherRadius > 1)
    {
        // Wide feathers tessellate to a coarser tolerance (draw.cpp:879-889).
        parametricPrecision = std::min(parametricPrecision * 100.f / (d.featherRadius * strokeMatrixMaxScale),
                                       parametricPrecision);
    }
    const bool roundJoinStroked = isStroke && d.join == Join::round;

    std::vector<ContourInfo> contours;
    std::vector<float> n4s;          // per curve (after chops)
    std::vector<uint32_t> paramCounts;
    std::vector<std::pair<V2, V2>> tangentPairs; // per rotation
    std::vector<uint32_t> polarCounts;
    std::vector<uint8_t> chopKeys;   // per stroked cubic (before chops)
    std::vector<std::array<float, 2>> chopTs;
    std::vector<std::array<V2, 10>> chopPts;

    size_t lineCount = 0;
    size_t curveIdx = 0, rotationIdx = 0;
    size_t contourFirstCurveIdx = 0, contourFirstRotationIdx = 0;
    int preChopVerbCount = 0;
    V2 endpointsSum{0, 0};
    bool closed = !isStroke;
    V2 lastTangent{0, 1}, firstTangent{0, 1};
    uint32_t startVerb = 0, startPt = 0;

    auto finishAndAppendContour = [&](uint32_t endVerb, uint32_t endPt) {
        if (closed)
        {
            V2 finalPt = P[endPt - 1];
            V2 movePt = P[startPt];
            if (fbits(movePt.x) != fbits(finalPt.x) || fbits(movePt.y) != fbits(finalPt.y))
            {
                if (roundJoinStroked)
                {
                    V2 tangent = movePt - finalPt;
                    tangentPairs.push_back({lastTangent, tangent});
                    ++rotationIdx;
                    lastTangent = tangent;
                }
                ++lineCount;
                ++preChopVerbCount;
                endpointsSum = endpointsSum + movePt;
            }
            if (roundJoinStroked && preChopVerbCount != 0)
            {
                tangentPairs.push_back({lastTangent, firstTangent});
                ++rotationIdx;
            }
        }
        size_t strokeJoinCount = preChopVerbCount;
        if (!closed)
            strokeJoinCount = std::max<size_t>(strokeJoinCount, 1) - 1;
        ContourInfo ci;
        ci.endVerb = endVerb;
        ci.endPt = endPt;
        ci.startVerb = startVerb;
        ci.startPt = startPt;
        ci.endLineIdx = lineCount;
        ci.firstCurveIdx = contourFirstCurveIdx;
        ci.endCurveIdx = curveIdx;
        ci.firstRotationIdx = contourFirstRotationIdx;
        ci.endRotationIdx = rotationIdx;
        ci.midpoint = isStroke ? V2{0, 0} : endpointsSum * (1.f / preChopVerbCount);
        ci.closed = closed;
        ci.strokeJoinCount = strokeJoinCount;
        ci.strokeCapSegmentCount = 0;
        ci.paddingVertexCount = 0;
        contours.push_back(ci);
        contourFirstCurveIdx = curveIdx;
        contourFirstRotationIdx = rotationIdx;
    };

    uint32_t pi = 0; // point cursor
    for (uint32_t vi = 0; vi < nVerbs; ++vi)
    {
        uint8_t verb = verbs[vi];
        switch (verb)
        {
            case kMove:
                if (vi != startVerb)
                {
                    finishAndAppendContour(vi, pi);
                    startVerb = vi;
                    startPt = pi;
                }
                preChopVerbCount = 0;
                endpointsSum = {0, 0};
                closed = !isStroke;
                lastTangent = {0, 1};
                firstTangent = {0, 1};
                break;
            case kClose:
                closed = true;
                break;
            case kLine:
            {
                const V2* p = P + pi - 1;
                if (roundJoinStroked)
                {
                    V2 tangent = p[1] - p[0];
                    if (preChopVerbCount == 0)
                        firstTangent = tangent;
                    else
                    {
                        tangentPairs.push_back({lastTangent, tangent});
                        ++rotationIdx;
                    }
                    lastTangent = tangent;
                }
                ++preChopVerbCount;
                endpointsSum = endpointsSum + p[1];
                ++lineCount;
                break;
            }
            case kCubic:
            {
                const V2* p = P + pi - 1;
                if (!isStrokeOrFeather)
                {
                    ++preChopVerbCount;
                    endpointsSum = endpointsSum + p[3];
                    n4s.push_back(wang_cubic_pow4(p, parametricPrecision, d.matrix));
                    ++curveIdx;
                    break;
                }
                if (roundJoinStroked)
                {
                    V2 t0 = find_cubic_tan0(p), t1 = find_cubic_tan1(p);
                    if (preChopVerbCount == 0)
                        firstTangent = t0;
                    else
                    {
                        tangentPairs.push_back({lastTangent, t0});
                        ++rotationIdx;
                    }
                    lastTangent = t1;
                }
                ++preChopVerbCount;
                endpointsSum = endpointsSum + p[3];
                float t[2] = {0, 0};
                bool areCusps = false;
                // Feathered fills were already chopped (draw.cpp:1062-1067).
                uint8_t numChops = isStroke ? (uint8_t)find_cubic_convex_180_chops(p, t, &areCusps) : 0;
                uint8_t chopKey = (uint8_t)((numChops << 1) | (areCusps ? 1 : 0));
                chopKeys.push_back(chopKey);
                chopTs.push_back({t[0], t[1]});
                std::array<V2, 10> buf{};
                V2 local[16];
                const V2* q = p;
                int subCount = 1;
                if (areCusps && numChops > 0)
                {
                    chop_cubic_around_cusps(p, local, t, numChops, strokeMatrixMaxScale);
                    q = local;
                    subCount = numChops * 2 + 1;
                }
                else if (numChops == 2)
                {
                    chop_cubic_at2(p, local, t[0], t[1]);
                    q = local;
                    subCount = 3;
                }
                else if (numChops == 1)
                {
                    chop_cubic_at(p, local, t[0]);
                    q = local;
                    subCount = 2;
                }
                (void)buf;
                for (int s = 0; s < subCount; ++s, q += 3)
                {
                    n4s.push_back(wang_cubic_pow4(q, parametricPrecision, d.matrix));
                    ++curveIdx;
                    // Feathered fills have no polar segments: a zero
                    // rotation (draw.cpp:1111-1128).
                    if (isStroke)
                        tangentPairs.push_back({find_cubic_tan0(q), find_cubic_tan1(q)});
                    else
                        tangentPairs.push_back({V2{0, 1}, V2{0, 1}});
                    ++rotationIdx;
                }
                break;
            }
            default:
                break; // quads are converted to cubics by RiveRenderPath
        }
        pi += pts_advance(verb);
    }
    if (startVerb != nVerbs)
        finishAndAppendContour(nVerbs, pi);

    // Pass 2: finish counts (draw.cpp:1180-1376).
    paramCounts.resize(n4s.size());
    for (size_t j = 0; j < n4s.size(); ++j)
        paramCounts[j] = parametric_count_from_pow4(n4s[j]);
    polarCounts.resize(tangentPairs.size());
    for (size_t j = 0; j < tangentPairs.size(); ++j)
        polarCounts[j] = polar_count(tangentPairs[j].first, tangentPairs[j].second,
                                     polarSegmentsPerRadian);

    size_t contourFirstLineIdx = 0;
    uint32_t tessVertexCount = 0;
    size_t unpaddedCurveCount = curveIdx;
    size_t emptyStrokeCountForCaps = 0;
    for (ContourInfo& c : contours)
    {
        size_t contourLineCount = c.endLineIdx - contourFirstLineIdx;
        uint32_t v = (uint32_t)(contourLineCount * 2);
        for (size_t j = c.firstCurveIdx; j < c.endCurveIdx; ++j)
            v += paramCounts[j];
        if (isStrokeOrFeather)
        {
            for (size_t j = c.firstRotationIdx; j < c.endRotationIdx; ++j)
                v += polarCounts[j];
            if (!isStroke)
                v += (uint32_t)c.strokeJoinCount * (feather_join_segment_count(polarSegmentsPerRadian) - 1);
            else if (d.join == Join::round)
                v -= (uint32_t)c.strokeJoinCount;
            else
                v += (uint32_t)c.strokeJoinCount * (NUM_SEGMENTS_IN_MITER_OR_BEVEL_JOIN - 1);
            bool empty = c.endLineIdx == contourFirstLineIdx &&
                         c.endCurveIdx == c.firstCurveIdx;
            Cap cap = d.cap;
            bool needsCaps = false;
            if (!empty)
            {
                cap = d.cap;
                needsCaps = !c.closed;
            }
            else if (isStroke)
            {
                cap = empty_stroke_cap(c.closed, d.join, d.cap);
                needsCaps = cap != Cap::butt;
            }
            if (needsCaps)
            {
                if (cap == Cap::round)
                {
                    float n = ceilf(polarSegmentsPerRadian * kPI);
                    n += 2;
                    n = fminf(n, (float)kMaxPolarSegments);
                    c.strokeCapSegmentCount = (uint32_t)n;
                }
                else
                {
                    c.strokeCapSegmentCount = NUM_SEGMENTS_IN_MITER_OR_BEVEL_JOIN;
                }
                emptyStrokeCountForCaps += empty ? 2 : 1;
                v += (c.strokeCapSegmentCount - 1) * 2;
            }
        }
        else
        {
            v += (uint32_t)(c.endCurveIdx - c.firstCurveIdx);
        }
        c.paddingVertexCount = (kMidpointFanPatchSegmentSpan -
                                (v % kMidpointFanPatchSegmentSpan)) %
                               kMidpointFanPatchSegmentSpan;
        v += c.paddingVertexCount;
        tessVertexCount += v;
        contourFirstLineIdx = c.endLineIdx;
    }
    result.tessVertexCount = tessVertexCount;
    result.midpointFanTessVertexCount =
        tessVertexCount == 0 ? 0 : (double_sided(d.dirs) ? tessVertexCount * 2 : tessVertexCount);
    result.maxTessellatedSegmentCount =
        (uint32_t)(lineCount + unpaddedCurveCount + emptyStrokeCountForCaps);
    if (tessVertexCount == 0)
        return result;

    // ---- pushTessellationData (draw.cpp:1899-1990) ----
    Writer w;
    w.dirs = d.dirs;
    w.out = &result.records;
    const uint32_t count = result.midpointFanTessVertexCount;
    switch (d.dirs)
    {
        case Dirs::forward:
            w.pathTessLocation = tessLocation;
            w.pathMirroredTessLocation = 0;
            break;
        case Dirs::reverse:
            w.pathTessLocation = 0;
            w.pathMirroredTessLocation = tessLocation + count;
            break;
        case Dirs::reverseThenForward:
            w.pathTessLocation = w.pathMirroredTessLocation = tessLocation + count / 2;
            break;
        case Dirs::forwardThenReverse:
            w.pathTessLocation = tessLocation;
            w.pathMirroredTessLocation = tessLocation + count;
            break;
    }

    // ---- pushMidpointFanTessellationData (draw.cpp:1992-2375) ----
    size_t chopCursor = 0;
    uint32_t contourID = firstContourID;
    for (const ContourInfo& c : contours)
    {
        const V2* pts = P + c.startPt;
        const V2* endPts = P + c.endPt;
        size_t cIdx = c.firstCurveIdx;
        size_t rIdx = c.firstRotationIdx;
        uint32_t joinTypeFlags = 0;
        bool needsFirstEmulatedCapAsJoin = false;
        uint32_t emulatedCapAsJoinFlags = 0;
        if (isStrokeOrFeather)
        {
            joinTypeFlags = isStroke ? join_type_flags(d.join) : FEATHER_JOIN_CONTOUR_FLAG;
            if (c.strokeCapSegmentCount != 0)
            {
                Cap cap = !c.closed ? d.cap : empty_stroke_cap(true, d.join, d.cap);
                switch (cap)
                {
                    case Cap::butt: emulatedCapAsJoinFlags = BEVEL_JOIN_CONTOUR_FLAG; break;
                    case Cap::square: emulatedCapAsJoinFlags = MITER_CLIP_JOIN_CONTOUR_FLAG; break;
                    case Cap::round: emulatedCapAsJoinFlags = ROUND_JOIN_CONTOUR_FLAG; break;
                }
                emulatedCapAsJoinFlags |= EMULATED_STROKE_CAP_CONTOUR_FLAG;
                needsFirstEmulatedCapAsJoin = true;
            }
        }

        // pushContour (render_context.cpp:3117-3158, 3132-3156)
        w.nextCubicPaddingVertexCount = c.paddingVertexCount;
        ContourData cd;
        cd.midpoint = c.midpoint;
        if (isStroke)
        {
            cd.midpoint.x = c.closed ? 1 : 0;
        }
        cd.pathID = d.pathID;
        cd.vertexIndex0 = w.nextVertexIndex();
        result.contours.push_back(cd);
        result.contourFirstRecord.push_back((uint32_t)result.records.size());
        result.contourPadding.push_back(c.paddingVertexCount);
        uint32_t contourIDWithFlags = d.contourFlags | contourID;
        ++contourID;

        const uint32_t numSegmentsInNotRoundJoin = isStroke ? NUM_SEGMENTS_IN_MITER_OR_BEVEL_JOIN
                                                            : feather_join_segment_count(polarSegmentsPerRadian);
        V2 joinTangent{0, 1};
        uint32_t joinSegmentCount = 1;
        V2 implicitClose[2];
        auto pushCapBefore = [&](const V2* cubic) {
            // pushEmulatedStrokeCapAsJoinBeforeCubic (draw.cpp:2377-2399)
            V2 rev[4] = {cubic[3], cubic[2], cubic[1], cubic[0]};
            w.pushCubic(rev, find_cubic_tan0(cubic), 0, 0, c.strokeCapSegmentCount,
                        contourIDWithFlags | emulatedCapAsJoinFlags);
        };
        uint32_t ptc = c.startPt;
        for (uint32_t vi = c.startVerb; vi < c.endVerb; ++vi)
        {
            uint8_t verb = verbs[vi];
            bool isFinalVerb = vi + 1 == c.endVerb;
            switch (verb)
            {
                case kMove:
                    implicitClose[1] = P[ptc];
                    break;
                case kClose:
                    break;
                case kLine:
                {
                    const V2* lp = P + ptc - 1;
                    if (isStrokeOrFeather)
                    {
                        if (c.closed || !isFinalVerb)
                        {
                            if (roundJoinStroked)
                            {
                                joinTangent = tangentPairs[rIdx].second;
                                joinSegmentCount = polarCounts[rIdx];
                                ++rIdx;
                            }
                            else
                            {
                                joinTangent = find_join_tangent(lp + 1, endPts, c.closed, pts);
                                joinSegmentCount = numSegmentsInNotRoundJoin;
                            }
                        }
                        else
                        {
                            V2 et = find_ending_tangent(pts, endPts);
                            joinTangent = V2{-et.x, -et.y};
                            joinTypeFlags = emulatedCapAsJoinFlags;
                            joinSegmentCount = c.strokeCapSegmentCount;
                        }
                    }
                    auto cubic = line_to_cubic(lp[0], lp[1]);
                    if (needsFirstEmulatedCapAsJoin)
                    {
                        pushCapBefore(cubic.data());
                        needsFirstEmulatedCapAsJoin = false;
                    }
                    w.pushCubic(cubic.data(), joinTangent, 1, 1, joinSegmentCount,
                                contourIDWithFlags | joinTypeFlags);
                    break;
                }
                case kCubic:
                {
                    const V2* cp = P + ptc - 1;
                    if (!isStrokeOrFeather)
                    {
                        uint32_t n = paramCounts[cIdx++];
                        w.pushCubic(cp, V2{0, 0}, n, 1, 1, contourIDWithFlags);
                        break;
                    }
                    uint8_t chopKey = chopKeys[chopCursor];
                    std::array<float, 2> ts = chopTs[chopCursor];
                    ++chopCursor;
                    V2 local[16];
                    const V2* p = cp;
                    uint32_t numChops = 0;
                    bool cusps = chopKey & 1;
                    uint32_t nch = chopKey >> 1;
                    if (cusps && nch > 0)
                    {
                        chop_cubic_around_cusps(cp, local, ts.data(), (int)nch,
                                                strokeMatrixMaxScale);
                        p = local;
                        numChops = nch * 2;
                    }
                    else if (nch == 2)
                    {
                        chop_cubic_at2(cp, local, ts[0], ts[1]);
                        p = local;
                        numChops = 2;
                    }
                    else if (nch == 1)
                    {
                        chop_cubic_at(cp, local, ts[0]);
                        p = local;
                        numChops = 1;
                    }
                    if (needsFirstEmulatedCapAsJoin)
                    {
                        pushCapBefore(p);
                        needsFirstEmulatedCapAsJoin = false;
                    }
                    for (uint32_t k = 0; k < numChops; ++k, p += 3)
                    {
                        uint32_t np = paramCounts[cIdx++];
                        uint32_t nq = polarCounts[rIdx++];
                        w.pushCubic(p, joinTangent, np, nq, 1, contourIDWithFlags | joinTypeFlags);
                    }
                    uint32_t np = paramCounts[cIdx++];
                    uint32_t nq = polarCounts[rIdx++];
                    if (c.closed || !isFinalVerb)
                    {
                        if (roundJoinStroked)
                        {
                            joinTangent = tangentPairs[rIdx].second;
                            joinSegmentCount = polarCounts[rIdx];
                            ++rIdx;
                        }
                        else
                        {
                            joinTangent = find_join_tangent(cp + 3, endPts, c.closed, pts);
                            joinSegmentCount = numSegmentsInNotRoundJoin;
                        }
                    }
                    else
                    {
                        V2 et = find_ending_tangent(pts, endPts);
                        joinTangent = V2{-et.x, -et.y};
                        joinTypeFlags = emulatedCapAsJoinFlags;
                        joinSegmentCount = c.strokeCapSegmentCount;
                    }
                    w.pushCubic(p, joinTangent, np, nq, joinSegmentCount,
                                contourIDWithFlags | joinTypeFlags);
                    break;
                }
                default:
                    break;
            }
            ptc += pts_advance(verb);
        }

        if (needsFirstEmulatedCapAsJoin)
        {
            V2 p0 = pts[0], left = {p0.x - 1, p0.y}, right = {p0.x + 1, p0.y};
            V2 a[4] = {p0, right, right, right};
            V2 b[4] = {p0, left, left, left};
            pushCapBefore(a);
            pushCapBefore(b);
        }
        else if (c.closed)
        {
            implicitClose[0] = endPts[-1];
            if (fbits(implicitClose[0].x) != fbits(implicitClose[1].x) ||
                fbits(implicitClose[0].y) != fbits(implicitClose[1].y))
            {
                auto cubic = line_to_cubic(implicitClose[0], implicitClose[1]);
                if (roundJoinStroked)
                {
                    joinTangent = tangentPairs[rIdx].second;
                    joinSegmentCount = polarCounts[rIdx];
                    ++rIdx;
                }
                else if (isStrokeOrFeather)
                {
                    joinTangent = find_starting_tangent(pts, endPts);
                    joinSegmentCount = numSegmentsInNotRoundJoin;
                }
                w.pushCubic(cubic.data(), joinTangent, 1, 1, joinSegmentCount,
                            contourIDWithFlags | joinTypeFlags);
            }
        }
        assert(cIdx == c.endCurveIdx);
    }
    return result;
}

void write_spans(const CubicRecord& r, Dirs dirs, std::vector<TessVertexSpan>* out)
{
    // render_context.cpp:3271-3402
    auto emit = [&](float y, int32_t x0, int32_t x1, float ry, int32_t rx0, int32_t rx1) {
        TessVertexSpan s;
        memcpy(s.pts, r.pts, sizeof(s.pts));
        s.joinTangent = r.joinTangent;
        s.y = y;
        s.reflectionY = ry;
        s.x0x1 = (int32_t)(((uint32_t)x1 << 16) | ((uint32_t)x0 & 0xffff));
        s.reflectionX0X1 = (int32_t)(((uint32_t)rx1 << 16) | ((uint32_t)rx0 & 0xffff));
        s.segmentCounts = (r.join << 20) | (r.polar << 10) | r.parametric;
        s.contourIDWithFlags = r.contourIDWithFlags;
        out->push_back(s);
    };
    const int32_t W = (int32_t)kTessTextureWidth;
    const float qnan = ubits(0x7fc00000u);
    const uint32_t total = r.totalVertexCount;
    if (dirs == Dirs::forward)
    {
        uint32_t y = r.forwardLocation / kTessTextureWidth;
        int32_t x0 = r.forwardLocation % kTessTextureWidth;
        int32_t x1 = x0 + (int32_t)total;
        for (;;)
        {
            emit((float)y, x0, x1, qnan, -1, -1);
            if (x1 > W)
            {
                ++y;
                x0 -= W;
                x1 -= W;
                continue;
            }
            break;
        }
    }
    else if (dirs == Dirs::reverse)
    {
        uint32_t ry = (r.mirroredLocation - 1) / kTessTextureWidth;
        int32_t rx0 = (r.mirroredLocation - 1) % kTessTextureWidth + 1;
        int32_t rx1 = rx0 - (int32_t)total;
        for (;;)
        {
            emit((float)ry, rx0, rx1, qnan, -1, -1);
            if (rx1 < 0)
            {
                --ry;
                rx0 += W;
                rx1 += W;
                continue;
            }
            break;
        }
    }
    else
    {
        int32_t y = (int32_t)(r.forwardLocation / kTessTextureWidth);
        int32_t x0 = r.forwardLocation % kTessTextureWidth;
        int32_t x1 = x0 + (int32_t)total;
        uint32_t ry = (r.mirroredLocation - 1) / kTessTextureWidth;
        int32_t rx0 = (r.mirroredLocation - 1) % kTessTextureWidth + 1;
        int32_t rx1 = rx0 - (int32_t)total;
        for (;;)
        {
            emit((float)y, x0, x1, (float)ry, rx0, rx1);
            if (x1 > W || rx1 < 0)
            {
                ++y;
                x0 -= W;
                x1 -= W;
                --ry;
                rx0 += W;
                rx1 += W;
                continue;
            }
            break;
        }
    }
}

void write_padding_span(uint32_t count, uint32_t location, std::vector<TessVertexSpan>* out)
{
    // render_context.cpp:3404-3424: an empty cubic, forward only.
    CubicRecord r{};
    r.parametric = 0;
    r.polar = 0;
    r.join = 1;
    r.contourIDWithFlags = 0;
    r.totalVertexCount = count;
    r.forwardLocation = location;
    write_spans(r, Dirs::forward, out);
}
} // namespace oracle
