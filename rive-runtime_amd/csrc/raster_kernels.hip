// gfx950 kernels of the RenderContext::flush hot path, after counting:
// color ramps, tessellation, segment expansion, binning and the tile raster.
//
// Geometry semantics follow the reference's shaders:
//   tessellation      renderer/src/shaders/tessellate.glsl:61-567
//   patch expansion   renderer/src/shaders/draw_path_common.glsl:275-789
//   pixel resolve     renderer/src/shaders/draw_raster_order_path.frag:14-238
//   color ramps       renderer/src/shaders/color_ramp.glsl:40-106
//
// The raster is not a triangle pipeline.  A midpoint-fan fill drawn forward
// and mirrored with counterclockwise culling adds +coverage for clockwise fan
// triangles and -coverage for counterclockwise ones, so its interior term is
// the winding number of the polygon through the inner AA points.  Each 16x16
// tile gets that winding number as
//     W(q) = C_T + (crossings of the tile's left edge above q)
//              + (crossings of q's row between the left edge and q)
// where C_T, the winding number at the tile's top-left corner, comes from a
// prefix sum of per-path corner crossings.  The 1px AA border quads and stroke
// quads are evaluated per pixel as triangles.  Every inside/outside decision is
// an exact integer predicate on a 1/256-pixel grid with a symbolic
// perturbation tie rule, so tile decompositions agree exactly.
#include "device_common.hpp"

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>

namespace rivehip
{
// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

__device__ __forceinline__ int64_t floor_div128(int64_t a) { return a >> 7; }
__device__ __forceinline__ int64_t ceil_div128(int64_t a) { return -((-a) >> 7); }
__device__ __forceinline__ int32_t clamp_thr(int64_t t)
{
    const int64_t lim = (int64_t)1 << 30;
    return (int32_t)(t < -lim ? -lim : (t > lim ? lim : t));
}

// ---------------------------------------------------------------------------
// Color ramps (color_ramp.glsl:40-106).  One block per GradientSpan.
// ---------------------------------------------------------------------------
__global__ void ramp_kernel(const uint4* __restrict__ spans, uint32_t spanCount, uint32_t rows,
                            uint32_t* __restrict__ ramps)
{
    const uint32_t s = blockIdx.x;
    if (s >= spanCount)
        return;
    uint4 gs = spans[s];
    float x0 = float(gs.x & 0xffffu) * (1.f / 65536.f);
    float x1 = float(gs.x >> 16) * (1.f / 65536.f);
    uint32_t yflags = gs.y;
    uint32_t y = yflags & ~GRAD_SPAN_FLAGS_MASK;
    if (y >= rows)
        return;
    float left = x0, right = x1;
    if (yflags & GRAD_SPAN_FLAG_LEFT_BORDER)
        left = (yflags & GRAD_SPAN_FLAG_COMPLEX_BORDER) ? 0.f : x0 - 1.f / kGradWidth;
    if (yflags & GRAD_SPAN_FLAG_RIGHT_BORDER)
        right = (yflags & GRAD_SPAN_FLAG_COMPLEX_BORDER) ? 1.f : x1 + 1.f / kGradWidth;
    float c0[4], c1[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        int shift = k < 3 ? 16 - 8 * k : 24;
        c0[k] = float((gs.z >> shift) & 0xffu) * (1.f / 255.f);
        c1[k] = float((gs.w >> shift) & 0xffu) * (1.f / 255.f);
    }
    for (uint32_t x = threadIdx.x; x < kGradWidth; x += blockDim.x)
    {
        float u = (float(x) + .5f) / kGradWidth;
        if (!(u >= left && u < right))
            continue;
        float c[4];
        if (u < x0)
        {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                c[k] = c0[k];
        }
        else if (u >= x1)
        {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                c[k] = c1[k];
        }
        else
        {
            float t = (u - x0) / (x1 - x0);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                c[k] = (c1[k] - c0[k]) * t + c0[k];
        }
        uint32_t px = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            float v = fminf(fmaxf(c[k], 0.f), 1.f);
            px |= (uint32_t)floorf(v * 255.f + .5f) << (8 * k);
        }
        ramps[(size_t)y * kGradWidth + x] = px;
    }
}

// ---------------------------------------------------------------------------
// Tessellation (tessellate.glsl).  One thread per texel of each span's first
// rectangle.  Double-sided spans also carry a mirrored rectangle; the raster
// reads forward texels only, so mirrored copies are not written.
// ---------------------------------------------------------------------------
struct SpanV
{
    float p[8];
    float jt[2];
    float y, ry;
    int32_t x0x1, rx0x1;
    uint32_t counts, flags;
};

__device__ __forceinline__ void span_first_rect(const SpanV& s, int32_t* lo, int32_t* hi,
                                                int32_t* x0, int32_t* x1)
{
    *x0 = (int32_t)((uint32_t)s.x0x1 << 16) >> 16;
    *x1 = s.x0x1 >> 16;
    *lo = max(min(*x0, *x1), 0);
    *hi = min(max(*x0, *x1), (int32_t)kTessTextureWidth);
}

__global__ void span_texel_count_kernel(const SpanV* __restrict__ spans, uint32_t n,
                                        uint32_t rows, uint32_t* __restrict__ counts)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    SpanV s = spans[i];
    int32_t lo, hi, x0, x1;
    span_first_rect(s, &lo, &hi, &x0, &x1);
    uint32_t c = 0;
    if (s.y >= 0.f && s.y < float(rows) && hi > lo)
        c = (uint32_t)(hi - lo);
    counts[i] = c;
}

struct F2
{
    float x, y;
};
__device__ __forceinline__ F2 f2sub(F2 a, F2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float f2dot(F2 a, F2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ bool f2neq(F2 a, F2 b) { return a.x != b.x || a.y != b.y; }
__device__ __forceinline__ F2 f2norm(F2 v)
{
    float l = sqrtf(v.x * v.x + v.y * v.y);
    return {v.x / l, v.y / l};
}
__device__ __forceinline__ F2 f2mix(F2 a, F2 b, float t)
{
    return {(b.x - a.x) * t + a.x, (b.y - a.y) * t + a.y};
}
__device__ __forceinline__ float det2(F2 c0, F2 c1) { return c0.x * c1.y - c1.x * c0.y; }
__device__ __forceinline__ float clampf(float x, float lo, float hi)
{
    return fminf(fmaxf(x, lo), hi);
}
__device__ __forceinline__ float cosine_between(F2 a, F2 b) // bezier_utils.glsl:24-31
{
    float ab = f2dot(a, b);
    float ab2 = f2dot(a, a) * f2dot(b, b);
    return ab2 == 0.f ? 1.f : clampf(ab * (1.f / sqrtf(ab2)), -1.f, 1.f);
}
__device__ __forceinline__ float atan2v(F2 v) // common.glsl:197-202
{
    v = f2norm(v);
    float th = acosf(clampf(v.x, -1.f, 1.f));
    return v.y >= 0.f ? th : -th;
}

// ---- feathered-fill softening (tessellate.glsl:146-193) ----
__device__ __forceinline__ float clamped_divide(float a, float b) // bezier_utils.glsl:158-169
{
    a = b < 0.f ? -a : a;
    b = fabsf(b);
    return a > 0.f ? (a < b ? a / b : 1.f) : 0.f;
}
// find_cubic_max_height (bezier_utils.glsl:171-246): three Newton steps on
// the height above the baseline p0 -> p3 of a convex cubic.
__device__ float cubic_max_height(F2 p0, F2 p1, F2 p2, F2 p3, float* outT)
{
    const F2 base = f2sub(p3, p0);
    const float lengthBase = sqrtf(f2dot(base, base));
    if (lengthBase == 0.f)
    {
        *outT = .5f;
        return 0.f;
    }
    const F2 norm{-base.y / lengthBase, base.x / lengthBase};
    const float h2 = f2dot(norm, f2sub(p2, p0));
    const float h1 = f2dot(norm, f2sub(p1, p0));
    const float dh = h1 - h2;
    const float _3A = 3.f * dh, B = -h1 - dh, C = h1;
    float t = .5f;
    for (int i = 0; i < 3; ++i)
    {
        const float _3At = _3A * t;
        t = clamped_divide(_3At * t - C, 2.f * (_3At + B));
    }
    *outT = t;
    return fabsf(t * (t * (t * _3A + 3.f * B) + 3.f * C));
}
// measure_cubic_local_curvature (bezier_utils.glsl:72-156): the tangent's
// rotation across a spread along the tangent, centered at T.
__device__ float cubic_local_curvature(F2 p0, F2 p1, F2 p2, F2 p3, float T, float desiredSpread)
{
    constexpr float PI_ = 3.14159265359f;
    const F2 C = f2sub(p1, p0), D = f2sub(p2, p1), E = f2sub(p3, p0);
    const F2 B = f2sub(D, C);
    const F2 A{-3.f * D.x + E.x, -3.f * D.y + E.y};
    F2 tangent{((A.x * T) + 2.f * B.x) * T + C.x, ((A.y * T) + 2.f * B.y) * T + C.y};
    tangent = F2{3.f * tangent.x, 3.f * tangent.y};
    const float lengthTan = sqrtf(f2dot(tangent, tangent));
    if (lengthTan == 0.f)
        return 0.f;
    tangent = F2{tangent.x * (1.f / lengthTan), tangent.y * (1.f / lengthTan)};
    const float A_ = 2.f * f2dot(A, tangent);
    const float C_ = 3.f * (A_ * T + 4.f * f2dot(B, tangent)) * T + 6.f * f2dot(C, tangent);
    const float maxDT = fminf(T, 1.f - T);
    const float maxSpread = (A_ * maxDT * maxDT + C_) * maxDT;
    const float targetSpread = fminf(desiredSpread, maxSpread * .9999f);
    float dt;
    if (A_ == 0.f)
        dt = targetSpread / C_;
    else
    {
        const float r = 1.f / A_;
        const float b = C_ * r, c = -targetSpread * r;
        const float Q = (-1.f / 3.f) * b, R = .5f * c;
        const float discr = R * R - Q * Q * Q;
        if (discr < 0.f)
        {
            const float sqrtQ = sqrtf(Q);
            const float theta = acosf(R / (sqrtQ * sqrtQ * sqrtQ));
            dt = -2.f * sqrtQ * cosf(theta * (1.f / 3.f) + (-PI_ * 2.f / 3.f));
        }
        else
        {
            float Ar = powf(fabsf(R) + sqrtf(discr), 1.f / 3.f);
            if (R < 0.f)
                Ar = -Ar;
            dt = Ar != 0.f ? Ar + Q / Ar : 0.f;
        }
    }
    dt = fabsf(dt);
    const float t0 = T - dt, t1 = T + dt;
    const F2 tan0d{(A.x * t0 + 2.f * B.x) * t0 + C.x, (A.y * t0 + 2.f * B.y) * t0 + C.y};
    const F2 tan1d{(A.x * t1 + 2.f * B.x) * t1 + C.x, (A.y * t1 + 2.f * B.y) * t1 + C.y};
    const F2 e0 = f2sub(f2neq(p0, p1) ? p1 : f2neq(p1, p2) ? p2 : p3, p0);
    const F2 e1 = f2sub(p3, f2neq(p3, p2) ? p2 : f2neq(p2, p1) ? p1 : p0);
    return acosf(cosine_between(t0 < 1e-3f ? e0 : tan0d, t1 > 1.f - 1e-3f ? e1 : tan1d));
}
// Flattens a feathered fill's cubic in proportion to its curvature across one
// standard deviation, so the feather's center dims the way a convolution of a
// curved edge does.
__device__ void soften_feathered_cubic(F2 p0, F2* p1, F2* p2, F2 p3, float featherRadius,
                                       const float* __restrict__ inverseFeather)
{
    constexpr float PI_ = 3.14159265359f;
    float maxHeightT;
    const float height = cubic_max_height(p0, *p1, *p2, p3, &maxHeightT);
    const float oneStddev = featherRadius * (1.f / 3.f);
    const float curvature = cubic_local_curvature(p0, *p1, *p2, p3, maxHeightT, oneStddev);
    float dimming = 1.f - curvature * (1.f / PI_);
    const F2 d = f2sub(p3, p0);
    const float stddevsPow2 = f2dot(d, d) / (oneStddev * oneStddev);
    dimming = fminf(dimming, (stddevsPow2 - 1.f) * .5f);
    dimming = fminf(dimming, .99f);
    const float x = feather_table(inverseFeather, .5f * dimming) * -2.f + 1.f;
    const float softness = clamped_divide(x * featherRadius, height);
    const F2 f1{(p3.x - p0.x) * (1.f / 3.f) + p0.x, (p3.y - p0.y) * (1.f / 3.f) + p0.y};
    const F2 f2{(p3.x - p0.x) * (2.f / 3.f) + p0.x, (p3.y - p0.y) * (2.f / 3.f) + p0.y};
    *p1 = F2{(f1.x - p1->x) * softness + p1->x, (f1.y - p1->y) * softness + p1->y};
    *p2 = F2{(f2.x - p2->x) * softness + p2->x, (f2.y - p2->y) * softness + p2->y};
}

__global__ void tessellate_kernel(const SpanV* __restrict__ spans, uint32_t spanCount,
                                  const uint32_t* __restrict__ texelOffsets, uint32_t totalTexels,
                                  TessTexel* __restrict__ tex, const uint4* __restrict__ contours,
                                  uint32_t contourCount, const float* __restrict__ pathData, uint32_t pathCount,
                                  const float* __restrict__ featherTables)
{
    constexpr float PI_ = 3.14159265359f;
    constexpr float TWO_PI = 6.28318530718f;
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= totalTexels)
        return;
    // Find the span: last s with texelOffsets[s] <= g.
    uint32_t lo = 0, hi = spanCount;
    while (hi - lo > 1)
    {
        uint32_t mid = (lo + hi) >> 1;
        if (texelOffsets[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    const SpanV s = spans[lo];
    int32_t rlo, rhi, x0, x1;
    span_first_rect(s, &rlo, &rhi, &x0, &x1);
    const int32_t x = rlo + (int32_t)(g - texelOffsets[lo]);
    const uint32_t row = (uint32_t)s.y;

    F2 P0{s.p[0], s.p[1]}, P1{s.p[2], s.p[3]}, P2{s.p[4], s.p[5]}, P3{s.p[6], s.p[7]};
    {
        // the span's path (tessellate.glsl:125-144): feathered fills soften
        const uint32_t contourID = s.flags & CONTOUR_ID_MASK;
        if (contourID != 0u && contourID <= contourCount)
        {
            const uint32_t pathID = contours[contourID - 1].z;
            const uint32_t pid = pathID <= pathCount ? pathID : 0u;
            const float strokeRadius = pathData[(size_t)pid * 16 + 6];
            const float featherRadius = pathData[(size_t)pid * 16 + 7];
            if (pid != 0u && featherRadius != 0.f && strokeRadius == 0.f)
                soften_feathered_cubic(P0, &P1, &P2, P3, featherRadius, featherTables + kFeatherTableSize);
        }
    }
    uint32_t parametricSegmentCount = s.counts & 0x3ffu;
    if ((s.flags & CULL_EXCESS_TESSELLATION_SEGMENTS_CONTOUR_FLAG) != 0u)
    {
        // An outer-cubic patch has a fixed span: Wang's formula again on the
        // path's matrix, the excess segments collapsed to T = 0
        // (tessellate.glsl:195-211).
        const uint32_t contourID = s.flags & CONTOUR_ID_MASK;
        const uint32_t pathID = contourID != 0u && contourID <= contourCount ? contours[contourID - 1].z : 0u;
        const float* m = pathData + (size_t)(pathID <= pathCount ? pathID : 0u) * 16;
        const F2 a{(-2.f * P1.x + P2.x) + P0.x, (-2.f * P1.y + P2.y) + P0.y};
        const F2 b{(-2.f * P2.x + P3.x) + P1.x, (-2.f * P2.y + P3.y) + P1.y};
        const F2 d0{m[0] * a.x + m[2] * a.y, m[1] * a.x + m[3] * a.y};
        const F2 d1{m[0] * b.x + m[2] * b.y, m[1] * b.x + m[3] * b.y};
        const float mm = fmaxf(d0.x * d0.x + d0.y * d0.y, d1.x * d1.x + d1.y * d1.y);
        const float n = fmaxf(ceilf(sqrtf(.75f * 4.f * sqrtf(mm))), 1.f);
        parametricSegmentCount = min((uint32_t)n, parametricSegmentCount);
    }
    const uint32_t polarSegmentCount = (s.counts >> 10) & 0x3ffu;
    const uint32_t joinSegmentCount = s.counts >> 20;
    const uint32_t totalVertexCount = parametricSegmentCount + polarSegmentCount + joinSegmentCount - 1u;
    const F2 joinTangent{s.jt[0], s.jt[1]};

    // Vertex stage.
    F2 t0 = f2sub(f2neq(P0, P1) ? P1 : f2neq(P1, P2) ? P2 : P3, P0);
    F2 t1 = f2sub(P3, f2neq(P3, P2) ? P2 : f2neq(P2, P1) ? P1 : P0);
    float radsPerPolarSegment = acosf(cosine_between(t0, t1)) / float(polarSegmentCount);
    float turn = det2(f2sub(P2, P0), f2sub(P3, P1));
    if (turn == 0.f)
        turn = det2(t0, t1);
    if (turn < 0.f)
        radsPerPolarSegment = -radsPerPolarSegment;
    uint32_t cf = s.flags;
    float radsPerJoinSegment = 0.f;
    if (joinSegmentCount > 1u)
    {
        float joinTheta = acosf(cosine_between(t1, joinTangent));
        float joinSpan = float(joinSegmentCount);
        if ((cf & (JOIN_TYPE_MASK | EMULATED_STROKE_CAP_CONTOUR_FLAG)) ==
            (ROUND_JOIN_CONTOUR_FLAG | EMULATED_STROKE_CAP_CONTOUR_FLAG))
            joinSpan -= 2.f;
        radsPerJoinSegment = joinTheta / joinSpan;
        if (det2(t1, joinTangent) < 0.f)
            radsPerJoinSegment = -radsPerJoinSegment;
    }
    if (x1 < x0)
        cf |= MIRRORED_CONTOUR_CONTOUR_FLAG;

    // Fragment stage.
    float px = float(x) + .5f;
    float vertexIdx = fmaxf(floorf(float(totalVertexCount) - fabsf(float(x1) - px)), 0.f);
    float parametric = float(parametricSegmentCount);
    float joinCount = float(joinSegmentCount);
    float rads = radsPerPolarSegment;
    F2 p0 = P0, p1 = P1, p2 = P2, p3 = P3;
    F2 ta = t0, tb = t1;
    float mergedSegmentCount = float(totalVertexCount) - joinCount;
    float mergedVertexID = vertexIdx;
    if (mergedVertexID <= mergedSegmentCount)
    {
        cf &= ~JOIN_TYPE_MASK;
    }
    else
    {
        p0 = p1 = p2 = p3;
        ta = t1;
        tb = joinTangent;
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
        theta = atan2v(isTan0 ? ta : tb);
    }
    else if ((cf & RETROFIT_TRI_STRIP_CONTOUR_FLAG) != 0u)
    {
        tessCoord = p0;
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
            F2 C = f2sub(p1, p0);
            F2 D = f2sub(p3, p0);
            F2 E = f2sub(p2, p1);
            F2 B = f2sub(E, C);
            F2 A{-3.f * E.x + D.x, -3.f * E.y + D.y};
            F2 B_{B.x * (parametric * 2.f), B.y * (parametric * 2.f)};
            float pp = parametric * parametric;
            F2 C_{C.x * pp, C.y * pp};
            float lastParametricVertexID = 0.f;
            float maxParametricVertexID = fminf(parametric - 1.f, mergedVertexID);
            F2 tan0norm = f2norm(ta);
            float negAbsRads = -fabsf(rads);
            float maxRotation0 = (1.f + mergedVertexID) * fabsf(rads);
            for (int p = 9; p >= 0; --p)
            {
                float testID = lastParametricVertexID + float(1 << p);
                if (testID <= maxParametricVertexID)
                {
                    F2 tt{testID * A.x + B_.x, testID * A.y + B_.y};
                    tt = F2{testID * tt.x + C_.x, testID * tt.y + C_.y};
                    float cosRotation = f2dot(f2norm(tt), tan0norm);
                    float maxRotation = fminf(testID * negAbsRads + maxRotation0, PI_);
                    if (cosRotation >= cosf(maxRotation))
                        lastParametricVertexID = testID;
                }
            }
            float parametricT = lastParametricVertexID / parametric;
            float lastPolarVertexID = mergedVertexID - lastParametricVertexID;
            float theta0 = acosf(clampf(tan0norm.x, -1.f, 1.f));
            theta0 = tan0norm.y >= 0.f ? theta0 : -theta0;
            theta = lastPolarVertexID * rads + theta0;
            F2 nrm{sinf(theta), -cosf(theta)};
            float a = f2dot(nrm, A), b_over_2 = f2dot(nrm, B), c = f2dot(nrm, C);
            float discr_over_4 = fmaxf(b_over_2 * b_over_2 - a * c, 0.f);
            float q = sqrtf(discr_over_4);
            if (b_over_2 > 0.f)
                q = -q;
            q -= b_over_2;
            float _5qa = -.5f * q * a;
            bool useQA = fabsf(q * q + _5qa) < fabsf(a * c + _5qa);
            float rs = useQA ? q : c, rt = useQA ? a : q;
            polarT = rt != 0.f ? rs / rt : 0.f;
            polarT = clampf(polarT, 0.f, 1.f);
            if (lastPolarVertexID == 0.f)
                polarT = 0.f;
            T = fmaxf(parametricT, polarT);
        }
        F2 ab = f2mix(p0, p1, T);
        F2 bc = f2mix(p1, p2, T);
        F2 cd = f2mix(p2, p3, T);
        F2 abc = f2mix(ab, bc, T);
        F2 bcd = f2mix(bc, cd, T);
        tessCoord = f2mix(abc, bcd, T);
        if (T != polarT)
            theta = atan2v(f2sub(bcd, abc));
    }
    TessTexel o;
    o.x = tessCoord.x;
    o.y = tessCoord.y;
    if ((cf & JOIN_TYPE_MASK) == FEATHER_JOIN_CONTOUR_FLAG)
        o.z = ((uint32_t)mergedSegmentCount << 16) | (uint32_t)mergedVertexID;
    else
        o.z = f2u(theta - TWO_PI * floorf(theta / TWO_PI));
    o.w = cf;
    tex[(size_t)row * kTessTextureWidth + (uint32_t)x] = o;
}

// ---------------------------------------------------------------------------
// Segment expansion: the draw vertex shader (draw_path_common.glsl:275-789)
// evaluated at both ends of every forward tessellation segment.
// ---------------------------------------------------------------------------
struct PathDataV
{
    float m[6];
    float strokeRadius, featherRadius;
    uint32_t z;
    float at[3];
    uint32_t cr[4];
};
struct ContourV
{
    float mx, my;
    uint32_t pathID, vertexIndex0;
};

__device__ __forceinline__ float signf_(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ float manhattan_pixel_width(const float M[4], float nx, float ny)
{
    float vx = M[0] * nx + M[2] * ny;
    float vy = M[1] * nx + M[3] * ny;
    return (fabsf(vx) + fabsf(vy)) * (1.f / (vx * vx + vy * vy));
}
// Vertex positions on the 1/256-px grid.  A non-finite position snaps to
// kSnapInvalid: the triangles using it are dropped, as a GPU rasteriser drops
// primitives with NaN or infinite positions.  Finite positions are exact up
// to 2^30 subpixels (4 194 304 px) from the origin and clamped beyond.
constexpr int32_t kSnapInvalid = INT32_MIN;
constexpr float kSnapLimit = 1073741824.f; // 2^30
__device__ __forceinline__ int32_t snap(float v)
{
    const float s = v * 256.f;
    if (!isfinite(s))
        return kSnapInvalid;
    return __float2int_rn(fminf(fmaxf(s, -kSnapLimit), kSnapLimit));
}

// The guard box (subpixels): the raster's integer predicates are exact for
// coordinates in [-2^23, 2^23), and a canvas spans at most [0, 2^22).  A
// primitive with a vertex outside the box is replaced by the pieces of it
// that the clamp onto the box maps it to: for points strictly inside the
// box, the clamp preserves every winding number (the segment from a point to
// its clamp stays outside the open box), and pieces whose border or stroke
// triangles lie inside the box keep them.
constexpr int32_t kGuardLo = -(1 << 22);
constexpr int32_t kGuardHi = (1 << 23) - 1;
__device__ __forceinline__ bool in_guard(int32_t v) { return v >= kGuardLo && v <= kGuardHi; }
__device__ __forceinline__ int32_t clamp_guard(int32_t v) { return min(max(v, kGuardLo), kGuardHi); }

struct VsOut
{
    float x, y, cx, cy;
};

// Evaluates one patch vertex: tessellation texel `k` (with flags `cf`) and
// outset `outset` (+1/-1).  Fill coverage is implied (1 inner, 0 outer).
__device__ VsOut draw_vertex(const TessTexel* __restrict__ tex, int texels, int k, uint32_t cf,
                             float outset, const PathDataV& pd)
{
    constexpr float PI_ = 3.14159265359f;
    constexpr float TWO_PI = 6.28318530718f;
    const float AA_RADIUS = .5f;
    TessTexel tv = tex[k];
    const float M[4] = {pd.m[0], pd.m[1], pd.m[2], pd.m[3]};
    float theta = u2f(tv.z);
    float nx = sinf(theta), ny = -cosf(theta);
    float offx, offy, covx = 0.f, covy = 0.f;
    float strokeRadius = pd.strokeRadius;
    float featherRadius = pd.featherRadius;
    if (featherRadius != 0.f)
    {
        // never harder than 1.5 standard deviations across 1/2 px (:478-485)
        const float mx = M[0] * nx + M[2] * ny, my = M[1] * nx + M[3] * ny;
        featherRadius = fmaxf(featherRadius, 1.f / sqrtf(mx * mx + my * my));
    }
    if (strokeRadius != 0.f)
    {
        float det = M[0] * M[3] - M[2] * M[1];
        outset *= signf_(det);
        if (cf & LEFT_JOIN_CONTOUR_FLAG)
            outset = fminf(outset, 0.f);
        if (cf & RIGHT_JOIN_CONTOUR_FLAG)
            outset = fmaxf(outset, 0.f);
        // a feathered stroke's AA ramp is its feather (draw_path_common.glsl:499-509)
        float aaRadius = featherRadius != 0.f ? featherRadius : manhattan_pixel_width(M, nx, ny) * AA_RADIUS;
        float globalCoverage = 1.f;
        if (aaRadius > strokeRadius && featherRadius == 0.f)
        {
            globalCoverage = strokeRadius / aaRadius;
            strokeRadius = aaRadius;
        }
        float vox = nx * (strokeRadius + aaRadius), voy = ny * (strokeRadius + aaRadius);
        float x = outset * (strokeRadius + aaRadius);
        covx = (1.f / (aaRadius * 2.f)) * (x + strokeRadius) + .5f;
        covy = (1.f / (aaRadius * 2.f)) * (-x + strokeRadius) + .5f;
        uint32_t joinType = cf & JOIN_TYPE_MASK;
        if (joinType > ROUND_JOIN_CONTOUR_FLAG)
        {
            int peekDir = 2;
            if ((cf & JOIN_TANGENT_0_CONTOUR_FLAG) == 0u)
                peekDir = -peekDir;
            if ((cf & MIRRORED_CONTOUR_CONTOUR_FLAG) != 0u)
                peekDir = -peekDir;
            float otherTheta = u2f(tex[min(max(k + peekDir, 0), texels - 1)].z);
            float joinAngle = fabsf(otherTheta - theta);
            if (joinAngle > PI_)
                joinAngle = TWO_PI - joinAngle;
            bool isTan0 = (cf & JOIN_TANGENT_0_CONTOUR_FLAG) != 0u;
            bool isLeftJoin = (cf & LEFT_JOIN_CONTOUR_FLAG) != 0u;
            float bisectTheta = joinAngle * (isTan0 == isLeftJoin ? -.5f : .5f) + theta;
            float bx = sinf(bisectTheta), by = -cosf(bisectTheta);
            float bisectPixelWidth = manhattan_pixel_width(M, bx, by);
            float miterRatio = cosf(joinAngle * .5f);
            float clipRadius;
            if ((joinType == MITER_CLIP_JOIN_CONTOUR_FLAG) ||
                (joinType == MITER_REVERT_JOIN_CONTOUR_FLAG && miterRatio >= .25f))
            {
                float miterInverseLimit = (cf & EMULATED_STROKE_CAP_CONTOUR_FLAG) != 0u ? 1.f : .25f;
                clipRadius = strokeRadius * (1.f / fmaxf(miterRatio, miterInverseLimit));
            }
            else
            {
                clipRadius = strokeRadius * miterRatio + bisectPixelWidth * .5f;
            }
            float clipAARadius = clipRadius + bisectPixelWidth * AA_RADIUS;
            if ((cf & JOIN_TANGENT_INNER_CONTOUR_FLAG) != 0u)
            {
                float strokeAARadius = strokeRadius + aaRadius;
                float slop = aaRadius * .125f;
                if (strokeAARadius <= clipAARadius * miterRatio + slop)
                {
                    float miterAARadius = strokeAARadius * (1.f / miterRatio);
                    vox = bx * miterAARadius;
                    voy = by * miterAARadius;
                }
                else
                {
                    float bax = bx * clipAARadius, bay = by * clipAARadius;
                    float k0 = vox * vox + voy * voy, k1 = bax * bax + bay * bay;
                    float a = vox, b = voy, c = bax, d = bay;
                    float idet = 1.f / (a * d - c * b);
                    float nvx = (k0 * d - k1 * b) * idet;
                    float nvy = (-k0 * c + k1 * a) * idet;
                    vox = nvx;
                    voy = nvy;
                }
            }
            float ptx = fabsf(outset) * vox, pty = fabsf(outset) * voy;
            float clipDistance =
                (clipAARadius - (ptx * bx + pty * by)) / (bisectPixelWidth * (AA_RADIUS * 2.f));
            if ((cf & LEFT_JOIN_CONTOUR_FLAG) != 0u)
                covy = clipDistance;
            else
                covx = clipDistance;
        }
        covx *= globalCoverage;
        covy *= globalCoverage;
        covy = fmaxf(covy, 1e-4f);
        if (featherRadius != 0.f)
            covx = -2.f - covx; // FEATHER_COVERAGE_BIAS marks a feathered stroke
        float px_ = outset * vox, py_ = outset * voy;
        offx = M[0] * px_ + M[2] * py_;
        offy = M[1] * px_ + M[3] * py_;
    }
    else
    {
        float det = M[0] * M[3] - M[2] * M[1];
        float idet = 1.f / det;
        float vx = outset * nx, vy = outset * ny;
        float rx = (vx * M[3] + vy * -M[1]) * idet;
        float ry = (vx * -M[2] + vy * M[0]) * idet;
        offx = signf_(rx) * AA_RADIUS;
        offy = signf_(ry) * AA_RADIUS;
    }
    VsOut o;
    o.x = (M[0] * tv.x + M[2] * tv.y) + offx + pd.m[4];
    o.y = (M[1] * tv.x + M[3] * tv.y) + offy + pd.m[5];
    o.cx = covx;
    o.cy = covy;
    return o;
}

// One end of a feathered fill's segment: the vertex shader's feather branches
// (draw_path_common.glsl:373-485, 661-744) for the end's edge vertex (outset
// 0) and outer vertex (outset +1) of a midpoint-fan-center-AA patch.  A
// feather-join texel steps its own angle from the texels on either side of
// the join; its spokes carry the corner's local coordinates.
struct FeatherEnd
{
    float ex, ey, tx, ty; // edge and outer positions (pixels)
    float c[4];           // the outer vertex's coverage vector
    float y0E;            // y0 of the edge vertex's vector (cotTheta shared)
    bool join;
};

__device__ __forceinline__ void pack_feathered_coverages(float cornerTheta, float spokeX, float spokeY,
                                                         float outset, float* out)
{
    // draw_path_common.glsl:107-150
    constexpr float PI_OVER_2 = 1.57079632679f;
    const float lx = (1.f - spokeX * fabsf(outset)) * .5f, ly = (1.f - spokeY * fabsf(outset)) * .5f;
    float cotTheta, y0;
    if (fabsf(cornerTheta - PI_OVER_2) < 1.f / 1e3f)
    {
        cotTheta = 0.f;
        y0 = 0.f;
    }
    else
    {
        const float tanTheta = tanf(cornerTheta);
        cotTheta = signf_(PI_OVER_2 - cornerTheta) / fmaxf(fabsf(tanTheta), 1.f / 1e6f);
        y0 = cotTheta >= 0.f ? ly - (1.f - lx) * tanTheta : ly + lx * tanTheta;
    }
    out[0] = fmaxf(lx, 0.f) + .25f;
    out[1] = -ly - 2.f;
    out[2] = cotTheta;
    out[3] = y0;
}

__device__ FeatherEnd feather_end(const TessTexel* __restrict__ tex, int texels, int k, uint32_t cf,
                                  const PathDataV& pd, uint32_t vertexIndex0)
{
    constexpr float PI_ = 3.14159265359f;
    constexpr float TWO_PI = 6.28318530718f;
    constexpr float PI_OVER_2 = 1.57079632679f;
    const TessTexel tv = tex[k];
    const float M[4] = {pd.m[0], pd.m[1], pd.m[2], pd.m[3]};
    FeatherEnd r;
    r.join = (cf & JOIN_TYPE_MASK) == FEATHER_JOIN_CONTOUR_FLAG;
    float theta, edge0 = 0.f, corner = 0.f, outset = 1.f, outsetE = 0.f;
    if (r.join)
    {
        float joinVertexID = (float)(tv.z & 0xffffu);
        float joinSegmentCount = (float)(tv.z >> 16);
        int ex = (int)(-joinVertexID - 1.f), ey = (int)(joinSegmentCount - joinVertexID + 1.f);
        if (cf & MIRRORED_CONTOUR_CONTOUR_FLAG)
        {
            ex = -ex;
            ey = -ey;
        }
        const TessTexel before = tex[min(max(k + ex, 0), texels - 1)];
        TessTexel after = tex[min(max(k + ey, 0), texels - 1)];
        if ((after.w & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu)) !=
            (before.w & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu)))
            after = tex[min((int)vertexIndex0, texels - 1)];
        edge0 = u2f(before.z);
        const float edge1 = u2f(after.z);
        corner = edge1 - edge0;
        if (fabsf(corner) > PI_)
            corner -= TWO_PI * signf_(corner);
        const float nonHelper = joinSegmentCount + 1.f - 3.f; // FEATHER_JOIN_HELPER_VERTEX_COUNT
        const float fwd = fminf(fmaxf(rintf(fabsf(corner) / PI_ * nonHelper), 1.f), nonHelper - 1.f);
        const float bwd = nonHelper - fwd;
        if (joinVertexID <= bwd)
        {
            // a backward segment: the angle outside the join
            corner = -(PI_ * signf_(corner) - corner);
            joinSegmentCount = bwd;
            if (joinVertexID == bwd)
            {
                outset = -outset;
                outsetE = -outsetE;
            }
        }
        else if (joinVertexID == bwd + 1.f)
        {
            // the throwaway vertex between the backward and forward segments
            joinVertexID = 0.f;
            joinSegmentCount = 0.f;
            outset = 0.f;
            outsetE = 0.f;
        }
        else
        {
            joinVertexID -= bwd + 2.f;
            joinSegmentCount = fwd;
        }
        theta = joinVertexID == joinSegmentCount ? edge1 : edge0 + corner * (joinVertexID / joinSegmentCount);
    }
    else
    {
        theta = u2f(tv.z);
    }
    const float nx = sinf(theta), ny = -cosf(theta);
    float featherRadius = pd.featherRadius;
    {
        const float mx = M[0] * nx + M[2] * ny, my = M[1] * nx + M[3] * ny;
        featherRadius = fmaxf(featherRadius, 1.f / sqrtf(mx * mx + my * my));
    }
    if (r.join)
    {
        if (corner < 0.f)
        {
            edge0 += corner;
            corner = -corner;
        }
        float spoke = theta - edge0;
        const float sm = spoke + PI_OVER_2;
        spoke = (sm - TWO_PI * floorf(sm / TWO_PI)) - PI_OVER_2; // GLSL mod
        spoke = fminf(fmaxf(spoke, 0.f), corner);
        if (spoke > corner * .5f)
            spoke = corner - spoke;
        const float sx = sinf(spoke), sy = cosf(spoke);
        float cE[4];
        pack_feathered_coverages(corner, sx, sy, outsetE, cE);
        pack_feathered_coverages(corner, sx, sy, outset, r.c);
        r.y0E = cE[3];
    }
    else
    {
        r.c[0] = 0.f;
        r.c[1] = -2.f;
        r.c[2] = 1e6f;
        r.c[3] = 0.f;
        r.y0E = .5f;
    }
    const float bx = M[0] * tv.x + M[2] * tv.y, by = M[1] * tv.x + M[3] * tv.y;
    const float oE = outsetE * featherRadius, oT = outset * featherRadius;
    const float px0 = oE * nx, py0 = oE * ny, px1 = oT * nx, py1 = oT * ny;
    r.ex = bx + (M[0] * px0 + M[2] * py0) + pd.m[4];
    r.ey = by + (M[1] * px0 + M[3] * py0) + pd.m[5];
    r.tx = bx + (M[0] * px1 + M[2] * py1) + pd.m[4];
    r.ty = by + (M[1] * px1 + M[3] * py1) + pd.m[5];
    return r;
}

__device__ __forceinline__ int64_t orient64(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t qx,
                                            int32_t qy)
{
    return (int64_t)(bx - ax) * (int64_t)(qy - ay) - (int64_t)(by - ay) * (int64_t)(qx - ax);
}

// The bounding box of a primitive's points (a feathered segment's inner
// points 2E - T included).
__device__ __forceinline__ void prim_extent(const Prim& p, int32_t* minx, int32_t* miny, int32_t* maxx,
                                            int32_t* maxy)
{
    *minx = min(min(p.p[0], p.p[2]), min(p.p[4], p.p[6]));
    *maxx = max(max(p.p[0], p.p[2]), max(p.p[4], p.p[6]));
    *miny = min(min(p.p[1], p.p[3]), min(p.p[5], p.p[7]));
    *maxy = max(max(p.p[1], p.p[3]), max(p.p[5], p.p[7]));
    if (p.flags & kPrimFeather)
    {
        // 2E - T for both ends: E is the midpoint of T and B, so the box of
        // {E, T, B} is the box of {T, B} widened to 2E - T.
        const int32_t b0x = 2 * p.p[0] - p.p[2], b0y = 2 * p.p[1] - p.p[3];
        const int32_t b1x = 2 * p.p[4] - p.p[6], b1y = 2 * p.p[5] - p.p[7];
        *minx = min(*minx, min(b0x, b1x));
        *maxx = max(*maxx, max(b0x, b1x));
        *miny = min(*miny, min(b0y, b1y));
        *maxy = max(*maxy, max(b0y, b1y));
    }
}

// A forward texel's two outset vertices, snapped, with stroke coverage.
struct VtxPair
{
    int32_t p[4]; // outset -1 (x, y), outset +1 (x, y)
    float cov[4];
};

__device__ __forceinline__ uint32_t find_draw(const DrawTess* __restrict__ draws, uint32_t drawCount, uint32_t g)
{
    uint32_t lo = 0, hi = drawCount;
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (draws[mid].primBase <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// The draw of primitive g (the last draw whose primBase <= g) for a block of
// consecutive primitives: one binary search for the block's first primitive,
// then a short forward walk per thread (a block spans few draws).  Every
// thread of the block must call it.
__device__ __forceinline__ uint32_t find_draw_block(const DrawTess* __restrict__ draws, uint32_t drawCount,
                                                    uint32_t g)
{
    __shared__ uint32_t sFirst;
    if (threadIdx.x == 0)
        sFirst = find_draw(draws, drawCount, blockIdx.x * blockDim.x);
    __syncthreads();
    uint32_t k = sFirst;
    while (k + 1 < drawCount && draws[k + 1].primBase <= g)
        ++k;
    return k;
}

// Vertex stage once per forward texel (draw_path_common.glsl:275-789); the
// segment between texel i and next(i) reuses both ends.
__global__ void vertex_kernel(const TessTexel* __restrict__ tex, int texels, const PathDataV* __restrict__ paths,
                              const DrawTess* __restrict__ draws, uint32_t drawCount, uint32_t totalPrims,
                              VtxPair* __restrict__ verts)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t di = find_draw_block(draws, drawCount, g);
    if (g >= totalPrims)
        return;
    const DrawTess d = draws[di];
    const int i = min((int)(d.forwardBase + (g - d.primBase)), texels - 2);
    const PathDataV pd = paths[d.pathID];
    if (pd.featherRadius != 0.f && pd.strokeRadius == 0.f)
        return; // feathered segments evaluate their ends in expand_kernel
    const uint32_t cfi = tex[i].w;
    const VsOut a0 = draw_vertex(tex, texels, i, cfi, -1.f, pd);
    const VsOut a1 = draw_vertex(tex, texels, i, cfi, +1.f, pd);
    VtxPair v;
    v.p[0] = snap(a0.x);
    v.p[1] = snap(a0.y);
    v.p[2] = snap(a1.x);
    v.p[3] = snap(a1.y);
    v.cov[0] = a0.cx;
    v.cov[1] = a0.cy;
    v.cov[2] = a1.cx;
    v.cov[3] = a1.cy;
    verts[g] = v;
}

// The pieces a primitive with vertices outside the guard box is replaced by
// (written to out[0..n) when W; the count either way).
//
// Fill: the inner edge I0 -> I1 maps under the clamp onto the box to a chain
// of straight pieces (the clamp is affine between the parameters where the
// edge crosses the box's four lines); the points of the chain are rounded
// once, so consecutive pieces, and the neighbouring primitives, meet exactly.
// A piece inside the box keeps the matching part of the border
// parallelogram (outer points at the same parameters; a fill's border
// coverage is the affine function that is 1 on the inner and 0 on the outer
// line, the same on any part of it); a piece on the box's boundary lies
// 16384 px or more from every canvas and keeps the inner edge only.
//
// Stroke: the quad is cut at the parameters where either side crosses a box
// line; parts inside the box are exact sub-quads with interpolated coverage,
// parts wholly beyond one box line are dropped, and the rest (strokes wider
// than the box margin) are clamped onto the box.
__device__ __forceinline__ int32_t round_guard(double v)
{
    return clamp_guard((int32_t)fmin(fmax(rint(v), (double)kGuardLo), (double)kGuardHi));
}
__device__ __forceinline__ void add_crossings(double a, double b, double* ts, int* n)
{
    // parameters in (0, 1) where a + t (b - a) meets a box line
    if (a == b)
        return;
    const double lines[2] = {(double)kGuardLo, (double)kGuardHi};
#pragma unroll
    for (int k = 0; k < 2; ++k)
    {
        const double t = (lines[k] - a) / (b - a);
        if (t > 0. && t < 1.)
            ts[(*n)++] = t;
    }
}
__device__ __forceinline__ void sort_params(double* ts, int n)
{
    for (int i = 1; i < n; ++i)
    {
        const double v = ts[i];
        int j = i - 1;
        while (j >= 0 && ts[j] > v)
        {
            ts[j + 1] = ts[j];
            --j;
        }
        ts[j + 1] = v;
    }
}
template <bool W> __device__ uint32_t clip_pieces(const Prim& p, uint32_t flags, Prim* __restrict__ out)
{
    double ts[10];
    int nt = 0;
    ts[nt++] = 0.;
    uint32_t n = 0;
    const uint32_t keep = flags & ~(kPrimEmpty | kPrimFar);
    if (flags & kPrimFill)
    {
        const double ax = p.p[0], ay = p.p[1], bx = p.p[4], by = p.p[5];
        add_crossings(ax, bx, ts, &nt);
        add_crossings(ay, by, ts, &nt);
        ts[nt++] = 1.;
        sort_params(ts, nt);
        int32_t px = clamp_guard(p.p[0]), py = clamp_guard(p.p[1]);
        for (int k = 0; k + 1 < nt; ++k)
        {
            const double t0 = ts[k], t1 = ts[k + 1];
            if (!(t1 > t0))
                continue;
            const bool last = k + 2 == nt;
            const int32_t qx = last ? clamp_guard(p.p[4]) : round_guard(ax + t1 * (bx - ax));
            const int32_t qy = last ? clamp_guard(p.p[5]) : round_guard(ay + t1 * (by - ay));
            const double tm = .5 * (t0 + t1);
            const double mx = ax + tm * (bx - ax), my = ay + tm * (by - ay);
            const bool inside = mx > (double)kGuardLo && mx < (double)kGuardHi && my > (double)kGuardLo &&
                                my < (double)kGuardHi;
            if (px != qx || py != qy || inside)
            {
                if (W)
                {
                    Prim q;
                    q.p[0] = px;
                    q.p[1] = py;
                    q.p[4] = qx;
                    q.p[5] = qy;
                    if (inside)
                    {
                        const double ox = p.p[2], oy = p.p[3], ux = p.p[6], uy = p.p[7];
                        q.p[2] = round_guard(ox + t0 * (ux - ox));
                        q.p[3] = round_guard(oy + t0 * (uy - oy));
                        q.p[6] = round_guard(ox + t1 * (ux - ox));
                        q.p[7] = round_guard(oy + t1 * (uy - oy));
                        q.flags = keep;
                    }
                    else
                    {
                        q.p[2] = px;
                        q.p[3] = py;
                        q.p[6] = qx;
                        q.p[7] = qy;
                        q.flags = keep | kPrimNoT1 | kPrimNoT2;
                    }
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        q.cov[c] = p.cov[c];
                    q.path = p.path;
                    q.pad[0] = q.pad[1] = 0;
                    out[n] = q;
                }
                ++n;
            }
            px = qx;
            py = qy;
        }
        return n;
    }
    // stroke: sides P0- -> P1- (vertices 0, 2) and P0+ -> P1+ (1, 3)
    add_crossings(p.p[0], p.p[4], ts, &nt);
    add_crossings(p.p[1], p.p[5], ts, &nt);
    add_crossings(p.p[2], p.p[6], ts, &nt);
    add_crossings(p.p[3], p.p[7], ts, &nt);
    ts[nt++] = 1.;
    sort_params(ts, nt);
    for (int k = 0; k + 1 < nt; ++k)
    {
        const double t0 = ts[k], t1 = ts[k + 1];
        if (!(t1 > t0))
            continue;
        double c[8]; // sub-quad corners: P0-, P0+, P1-, P1+ at t0 / t1
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
            for (int a = 0; a < 2; ++a)
            {
                const double from = p.p[2 * v + a], to = p.p[4 + 2 * v + a];
                c[2 * v + a] = from + t0 * (to - from);
                c[4 + 2 * v + a] = from + t1 * (to - from);
            }
        bool allIn = true;
        bool beyond[4] = {true, true, true, true}; // all x < lo, x > hi, y < lo, y > hi
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const double x = c[2 * v], y = c[2 * v + 1];
            allIn &= x >= (double)kGuardLo && x <= (double)kGuardHi && y >= (double)kGuardLo &&
                     y <= (double)kGuardHi;
            beyond[0] &= x < (double)kGuardLo;
            beyond[1] &= x > (double)kGuardHi;
            beyond[2] &= y < (double)kGuardLo;
            beyond[3] &= y > (double)kGuardHi;
        }
        if (!allIn && (beyond[0] || beyond[1] || beyond[2] || beyond[3]))
            continue;
        if (W)
        {
            Prim q;
#pragma unroll
            for (int v = 0; v < 8; ++v)
                q.p[v] = round_guard(c[v]);
            // exact ends where the cut is at the primitive's own ends
            if (t0 == 0.)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    q.p[v] = clamp_guard(p.p[v]);
            if (t1 == 1.)
#pragma unroll
                for (int v = 4; v < 8; ++v)
                    q.p[v] = clamp_guard(p.p[v]);
#pragma unroll
            for (int a = 0; a < 4; ++a)
            {
                const float from = p.cov[a], to = p.cov[4 + a];
                q.cov[a] = (float)((double)from + t0 * (double)(to - from));
                q.cov[4 + a] = (float)((double)from + t1 * (double)(to - from));
            }
            q.flags = keep;
            q.path = p.path;
            q.pad[0] = q.pad[1] = 0;
            out[n] = q;
        }
        ++n;
    }
    return n;
}

// A fill primitive's non-finite inner vertices move to the contour's
// midpoint (the fan triangles through them are dropped).
__device__ __forceinline__ void replace_inner_vertices(Prim& out, int32_t mx, int32_t my)
{
    if (out.p[0] == kSnapInvalid || out.p[1] == kSnapInvalid)
    {
        out.p[0] = mx;
        out.p[1] = my;
    }
    if (out.p[4] == kSnapInvalid || out.p[5] == kSnapInvalid)
    {
        out.p[4] = mx;
        out.p[5] = my;
    }
}

// Drops the triangles through non-finite vertices: T1 uses vertices 0, 1, 2
// and T2 vertices 2, 1, 3 (fills: I0, O0, I1 and I1, O0, O1; strokes: P0-,
// P0+, P1- and P1-, P0+, P1+).  The non-finite vertices get a finite
// vertex's place for the bounds.  True when nothing of the primitive is left.
__device__ __forceinline__ bool drop_nonfinite(Prim& out, uint32_t* flags)
{
    bool bad[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        bad[k] = out.p[2 * k] == kSnapInvalid || out.p[2 * k + 1] == kSnapInvalid;
    if (bad[0] || bad[1] || bad[2])
        *flags |= kPrimNoT1;
    if (bad[1] || bad[2] || bad[3])
        *flags |= kPrimNoT2;
    int good = -1;
#pragma unroll
    for (int k = 3; k >= 0; --k)
        if (!bad[k])
            good = k;
    if (good < 0)
        return true;
    const int32_t gx = out.p[2 * good], gy = out.p[2 * good + 1];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (bad[k])
        {
            out.p[2 * k] = gx;
            out.p[2 * k + 1] = gy;
        }
    return false;
}

// Records a primitive leaving the guard box with its piece count.
__device__ __forceinline__ void count_far(const Prim& out, uint32_t flags, uint32_t g, uint32_t di,
                                       uint32_t* __restrict__ farList, uint32_t* __restrict__ farCounts,
                                       uint32_t* __restrict__ drawPieces)
{
    const uint32_t n = clip_pieces<false>(out, flags, nullptr);
    if (n != 0)
    {
        farList[atomicAdd(&farCounts[0], 1u)] = g;
        atomicAdd(&farCounts[1], n);
        atomicAdd(&drawPieces[di], n);
    }
}

// Writes the guard-box pieces of the far primitives after the flush's
// primitives, each draw's pieces in its own range.
__global__ void far_pieces_kernel(Prim* __restrict__ prims, const uint32_t* __restrict__ farList,
                                  const uint32_t* __restrict__ farCounts, const DrawTess* __restrict__ draws,
                                  uint32_t drawCount, const uint32_t* __restrict__ drawPieceOffset,
                                  uint32_t* __restrict__ pieceCursor, uint32_t totalPrims)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= farCounts[0])
        return;
    const uint32_t g = farList[i];
    const Prim p = prims[g];
    const uint32_t di = find_draw(draws, drawCount, g);
    const uint32_t n = clip_pieces<false>(p, p.flags, nullptr);
    const uint32_t at = totalPrims + drawPieceOffset[di] + atomicAdd(&pieceCursor[di], n);
    clip_pieces<true>(p, p.flags, prims + at);
}

// The draw list binning reads when a flush has far primitives: the draws,
// then one more per draw owning its pieces.
__global__ void piece_draws_kernel(const DrawTess* __restrict__ draws, uint32_t drawCount,
                                   const uint32_t* __restrict__ drawPieces,
                                   const uint32_t* __restrict__ drawPieceOffset, uint32_t totalPrims,
                                   DrawTess* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= drawCount)
        return;
    DrawTess d = draws[i];
    out[i] = d;
    d.forwardCount = drawPieces[i];
    d.primBase = totalPrims + drawPieceOffset[i];
    out[drawCount + i] = d;
}

__global__ void expand_kernel(const TessTexel* __restrict__ tex, int texels, const VtxPair* __restrict__ verts,
                              const ContourV* __restrict__ contours, uint32_t contourCount,
                              const PathDataV* __restrict__ paths, const DrawTess* __restrict__ draws,
                              uint32_t drawCount, uint32_t totalPrims, Prim* __restrict__ prims,
                              PathCells* __restrict__ cells, uint32_t* __restrict__ badContours,
                              uint32_t* __restrict__ specialList, uint32_t* __restrict__ specialCount)
{
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t di = find_draw_block(draws, drawCount, g);
    if (g >= totalPrims)
        return;
    const DrawTess d = draws[di];
    // A reverse-only draw (a feathered fill through the atlas under a
    // left-handed matrix) holds its contours back to front: read it from the
    // end, so each segment is its forward one.
    const bool reversed = d.dirs == 1u;
    const int i = reversed ? max((int)(d.forwardBase + d.forwardCount - 1u - (g - d.primBase)), 1)
                           : min((int)(d.forwardBase + (g - d.primBase)), texels - 2);
    Prim out;
    out.path = d.pathID;
    out.pad[0] = out.pad[1] = 0;
    const TessTexel ti = tex[i];
    const uint32_t cfi = ti.w;
    // A texel naming no ContourData record (ID 0 or beyond contourCount) is
    // counted and the flush fails; the clamp only keeps this read in bounds.
    const uint32_t rawContourID = cfi & CONTOUR_ID_MASK;
    if (rawContourID == 0u || rawContourID > contourCount)
        atomicAdd(badContours, 1u);
    const uint32_t contourID = min(max(rawContourID, 1u), max(contourCount, 1u));
    const ContourV cd = contours[contourID - 1];
    const PathDataV pd = paths[d.pathID];
    // next(i): the next vertex of the contour, wrapping closed contours to
    // vertexIndex0 (draw_path_common.glsl:336-388).
    const int step = reversed ? -1 : 1;
    int j = i + step;
    uint32_t cfj = tex[j].w;
    auto sameContour = [cfi](uint32_t w) {
        return (w & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu)) == (cfi & (MIRRORED_CONTOUR_CONTOUR_FLAG | 0xffffu));
    };
    // An interior-triangulated fill's outline is its outer-cubic patches; the
    // triangle strips retrofitted into its patches (the glue between a
    // chopped cubic's chord and its pieces', draw.cpp:2510-2545) lie inside
    // and are not part of it: their texels give no segment, and a segment
    // reaching one continues to the next outline texel of the contour.
    bool wrap;
    if ((cfi & RETROFIT_TRI_STRIP_CONTOUR_FLAG) != 0u)
    {
        j = i;
        cfj = cfi;
        wrap = false;
    }
    else
    {
        if ((cfj & RETROFIT_TRI_STRIP_CONTOUR_FLAG) != 0u && sameContour(cfj))
        {
            int k = j;
            while (k + step >= 0 && k + step < texels && sameContour(tex[k].w) &&
                   (tex[k].w & RETROFIT_TRI_STRIP_CONTOUR_FLAG) != 0u)
                k += step;
            j = k;
            cfj = tex[j].w;
        }
        wrap = !sameContour(cfj) || (cfj & RETROFIT_TRI_STRIP_CONTOUR_FLAG) != 0u;
    }
    if (wrap)
    {
        bool isClosed = pd.strokeRadius == 0.f || cd.mx != 0.f;
        j = isClosed ? min((int)cd.vertexIndex0, texels - 1) : i;
        cfj = (tex[j].w & ~MIRRORED_CONTOUR_CONTOUR_FLAG) | (cfi & MIRRORED_CONTOUR_CONTOUR_FLAG);
    }
    const bool stroke = pd.strokeRadius != 0.f;
    const bool feathered = !stroke && pd.featherRadius != 0.f;
    if (feathered)
    {
        // [E0, T0, E1, T1]; the inner points are 2E - T
        const FeatherEnd a = feather_end(tex, texels, i, cfi, pd, cd.vertexIndex0);
        const FeatherEnd b = feather_end(tex, texels, j, cfj, pd, cd.vertexIndex0);
        out.p[0] = snap(a.ex);
        out.p[1] = snap(a.ey);
        out.p[2] = snap(a.tx);
        out.p[3] = snap(a.ty);
        out.p[4] = snap(b.ex);
        out.p[5] = snap(b.ey);
        out.p[6] = snap(b.tx);
        out.p[7] = snap(b.ty);
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            out.cov[k] = a.c[k];
            out.cov[4 + k] = b.c[k];
        }
        out.pad[0] = f2u(a.y0E);
        out.pad[1] = f2u(b.y0E);
    }
    else
    {
        // fill: [I0, O0, I1, O1]; stroke: [P0-, P0+, P1-, P1+]
        const VtxPair va = verts[g];
        VtxPair vb;
        const int64_t rel = (int64_t)j - (int64_t)d.forwardBase;
        if (rel >= 0 && rel < (int64_t)d.forwardCount && cfj == tex[j].w)
            vb = verts[d.primBase + (uint32_t)rel];
        else
        {
            // next(i) outside this draw's forward texels, or with other flags
            const VsOut b0 = draw_vertex(tex, texels, j, cfj, -1.f, pd);
            const VsOut b1 = draw_vertex(tex, texels, j, cfj, +1.f, pd);
            vb.p[0] = snap(b0.x);
            vb.p[1] = snap(b0.y);
            vb.p[2] = snap(b1.x);
            vb.p[3] = snap(b1.y);
            vb.cov[0] = b0.cx;
            vb.cov[1] = b0.cy;
            vb.cov[2] = b1.cx;
            vb.cov[3] = b1.cy;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            out.p[k] = va.p[k];
            out.p[4 + k] = vb.p[k];
            out.cov[k] = va.cov[k];
            out.cov[4 + k] = vb.cov[k];
        }
    }
    uint32_t flags = stroke ? kPrimStroke : kPrimFill;
    if (!stroke && ((cfi & NEGATE_PATH_FILL_COVERAGE_FLAG) != 0u))
        flags |= kPrimNegate;
    bool empty = false;
    // A fill contour without a finite midpoint has no fan: its inner edges
    // draw nothing.
    if (!stroke)
    {
        const float fx = (pd.m[0] * cd.mx + pd.m[2] * cd.my) + 0.f + pd.m[4];
        const float fy = (pd.m[1] * cd.mx + pd.m[3] * cd.my) + 0.f + pd.m[5];
        if (!isfinite(fx * 256.f) || !isfinite(fy * 256.f))
            flags |= kPrimNoEdge;
    }
    int32_t lo = out.p[0], hi = out.p[0];
#pragma unroll
    for (int k = 1; k < 8; ++k)
    {
        lo = min(lo, out.p[k]);
        hi = max(hi, out.p[k]);
    }
    if (feathered && in_guard(lo) && in_guard(hi))
    {
        // the inner points 2E - T must lie in the box as well
#pragma unroll
        for (int k = 0; k < 2; ++k)
        {
            const int32_t bxk = 2 * out.p[4 * k] - out.p[4 * k + 2], byk = 2 * out.p[4 * k + 1] - out.p[4 * k + 3];
            lo = min(lo, min(bxk, byk));
            hi = max(hi, max(bxk, byk));
        }
    }
    if (stroke && pd.featherRadius != 0.f)
        flags |= kPrimFeatherStroke;
    if (feathered)
    {
        const bool j0 = (cfi & JOIN_TYPE_MASK) == FEATHER_JOIN_CONTOUR_FLAG;
        const bool j1 = (cfj & JOIN_TYPE_MASK) == FEATHER_JOIN_CONTOUR_FLAG;
        flags |= kPrimFeather | (j0 ? kPrimJoin0 : 0u) | (j1 ? kPrimJoin1 : 0u);
    }
    if (!in_guard(lo) || !in_guard(hi))
    {
        // A vertex outside the guard box or without a finite position
        // (kSnapInvalid lies below the box): special_prims_kernel settles it.
        // A feathered segment there is settled as a fill's border
        // [E0, T0, E1, T1] without its feather.
        flags &= ~(kPrimFeather | kPrimJoin0 | kPrimJoin1);
        flags |= kPrimSpecial;
        empty = true;
        specialList[atomicAdd(specialCount, 1u)] = g;
    }
    else if (stroke)
    {
        // Triangles (P0-, P0+, P1-) and (P1-, P0+, P1+); clockwise only.
        int64_t A1 = orient64(out.p[0], out.p[1], out.p[2], out.p[3], out.p[4], out.p[5]);
        int64_t A2 = orient64(out.p[4], out.p[5], out.p[2], out.p[3], out.p[6], out.p[7]);
        empty = A1 <= 0 && A2 <= 0;
    }
    else
    {
        // (a feathered segment's inner triangles mirror its outer ones)
        int64_t A1 = orient64(out.p[0], out.p[1], out.p[2], out.p[3], out.p[4], out.p[5]);
        int64_t A2 = orient64(out.p[4], out.p[5], out.p[2], out.p[3], out.p[6], out.p[7]);
        const bool noEdge = (flags & kPrimNoEdge) || (out.p[0] == out.p[4] && out.p[1] == out.p[5]);
        empty = noEdge && A1 == 0 && A2 == 0;
    }
    if (empty)
        flags |= kPrimEmpty;
    out.flags = flags;
    prims[g] = out;
    // Unclamped union of the path's tile bounds; cells_layout_kernel clips it
    // to the canvas, so geometry off the canvas still widens the grid toward
    // it.  Lanes of one path (prims of a path are contiguous) reduce first so
    // a wave issues one set of atomics per path it touches.
    int32_t tx0 = 0x7fffffff, ty0 = 0x7fffffff, tx1 = -0x7fffffff, ty1 = -0x7fffffff;
    if (!empty)
    {
        int32_t minx, maxx, miny, maxy;
        prim_extent(out, &minx, &miny, &maxx, &maxy);
        tx0 = minx >> (kSubpixelLog2 + kTileLog2);
        ty0 = miny >> (kSubpixelLog2 + kTileLog2);
        tx1 = (maxx >> (kSubpixelLog2 + kTileLog2)) + 1;
        ty1 = (maxy >> (kSubpixelLog2 + kTileLog2)) + 1;
    }
    const int lane = threadIdx.x & 63;
    const uint32_t first = (uint32_t)__shfl((int)d.pathID, 0);
    const bool same = d.pathID == first;
    const uint64_t sameMask = __ballot(same);
    if (sameMask == ~0ull)
    {
        // the whole wave (all 64 lanes active) belongs to one path
#pragma unroll
        for (int k = 32; k > 0; k >>= 1)
        {
            tx0 = min(tx0, __shfl_xor(tx0, k));
            ty0 = min(ty0, __shfl_xor(ty0, k));
            tx1 = max(tx1, __shfl_xor(tx1, k));
            ty1 = max(ty1, __shfl_xor(ty1, k));
        }
        if (lane == __builtin_ctzll(sameMask) && tx0 < tx1)
        {
            PathCells* pc = &cells[d.pathID];
            atomicMin(&pc->tx0, tx0);
            atomicMin(&pc->ty0, ty0);
            atomicMax(&pc->tx1, tx1);
            atomicMax(&pc->ty1, ty1);
        }
    }
    else if (!empty)
    {
        PathCells* pc = &cells[d.pathID];
        atomicMin(&pc->tx0, tx0);
        atomicMin(&pc->ty0, ty0);
        atomicMax(&pc->tx1, tx1);
        atomicMax(&pc->ty1, ty1);
    }
}

// The primitives expand_kernel set aside (a vertex outside the guard box or
// without a finite position), in a grid-stride loop over their list: the
// triangles through non-finite vertices are dropped (a fill's inner vertex
// moves to the contour's midpoint, its fan triangles being dropped), and a
// primitive still leaving the box is counted for its pieces.  Its path's
// bounds take the primitive clamped onto the box.
__global__ void special_prims_kernel(const TessTexel* __restrict__ tex, int texels,
                                     const ContourV* __restrict__ contours, uint32_t contourCount,
                                     const PathDataV* __restrict__ paths, const DrawTess* __restrict__ draws,
                                     uint32_t drawCount, Prim* __restrict__ prims, PathCells* __restrict__ cells,
                                     const uint32_t* __restrict__ specialList,
                                     const uint32_t* __restrict__ specialCount, uint32_t* __restrict__ farList,
                                     uint32_t* __restrict__ farCounts, uint32_t* __restrict__ drawPieces)
{
    const uint32_t n = *specialCount;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    {
        const uint32_t g = specialList[k];
        Prim out = prims[g];
        uint32_t flags = out.flags & ~(kPrimEmpty | kPrimSpecial);
        const uint32_t di = find_draw(draws, drawCount, g);
        const DrawTess d = draws[di];
        const bool stroke = (flags & kPrimStroke) != 0u;
        bool anyBad = false;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            anyBad |= out.p[c] == kSnapInvalid;
        if (anyBad && !stroke && !(flags & kPrimNoEdge))
        {
            const int i = d.dirs == 1u ? max((int)(d.forwardBase + d.forwardCount - 1u - (g - d.primBase)), 1)
                                       : min((int)(d.forwardBase + (g - d.primBase)), texels - 2);
            const uint32_t cid = min(max(tex[i].w & CONTOUR_ID_MASK, 1u), max(contourCount, 1u));
            const ContourV cd = contours[cid - 1];
            const PathDataV pd = paths[d.pathID];
            const float fx = (pd.m[0] * cd.mx + pd.m[2] * cd.my) + 0.f + pd.m[4];
            const float fy = (pd.m[1] * cd.mx + pd.m[3] * cd.my) + 0.f + pd.m[5];
            replace_inner_vertices(out, snap(fx), snap(fy));
        }
        bool empty = anyBad && drop_nonfinite(out, &flags);
        if (!empty)
        {
            const int64_t A1 =
                (flags & kPrimNoT1) ? 0 : orient64(out.p[0], out.p[1], out.p[2], out.p[3], out.p[4], out.p[5]);
            const int64_t A2 =
                (flags & kPrimNoT2) ? 0 : orient64(out.p[4], out.p[5], out.p[2], out.p[3], out.p[6], out.p[7]);
            if (stroke)
                empty = A1 <= 0 && A2 <= 0;
            else
                empty = ((flags & kPrimNoEdge) || (out.p[0] == out.p[4] && out.p[1] == out.p[5])) && A1 == 0 &&
                        A2 == 0;
        }
        if (empty)
        {
            out.flags = flags | kPrimEmpty;
            prims[g] = out;
            continue;
        }
        int32_t lo = out.p[0], hi = out.p[0];
#pragma unroll
        for (int c = 1; c < 8; ++c)
        {
            lo = min(lo, out.p[c]);
            hi = max(hi, out.p[c]);
        }
        if (!in_guard(lo) || !in_guard(hi))
        {
            // binned as its pieces, written once their count is known
            flags |= kPrimEmpty | kPrimFar;
            count_far(out, flags, g, di, farList, farCounts, drawPieces);
        }
        out.flags = flags;
        prims[g] = out;
        int32_t q[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            q[c] = clamp_guard(out.p[c]);
        PathCells* pc = &cells[d.pathID];
        atomicMin(&pc->tx0, min(min(q[0], q[2]), min(q[4], q[6])) >> (kSubpixelLog2 + kTileLog2));
        atomicMin(&pc->ty0, min(min(q[1], q[3]), min(q[5], q[7])) >> (kSubpixelLog2 + kTileLog2));
        atomicMax(&pc->tx1, (max(max(q[0], q[2]), max(q[4], q[6])) >> (kSubpixelLog2 + kTileLog2)) + 1);
        atomicMax(&pc->ty1, (max(max(q[1], q[3]), max(q[5], q[7])) >> (kSubpixelLog2 + kTileLog2)) + 1);
    }
}

// ---------------------------------------------------------------------------
// Per-path cell grids.
// ---------------------------------------------------------------------------
__global__ void cells_init_kernel(PathCells* __restrict__ cells, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    PathCells c;
    c.tx0 = c.ty0 = 0x7fffffff;
    c.tx1 = c.ty1 = -0x7fffffff;
    c.base = 0;
    c.width = 0;
    c.pad[0] = c.pad[1] = 0;
    cells[i] = c;
}

// Single block: sizes each path's grid and assigns bases (exclusive scan).
// The canvas of each path (0xffffffff for paths outside every canvas): one
// block per canvas.
__global__ void path_canvas_kernel(const CanvasDev* __restrict__ canvases, uint32_t* __restrict__ pathCanvas)
{
    const CanvasDev c = canvases[blockIdx.x];
    for (uint32_t p = c.firstPath + threadIdx.x; p < c.pathEnd; p += blockDim.x)
        pathCanvas[p] = blockIdx.x;
}

__global__ void cells_layout_kernel(PathCells* __restrict__ cells, uint32_t n,
                                    const uint32_t* __restrict__ pathCanvas,
                                    const CanvasDev* __restrict__ canvases, uint32_t* __restrict__ totalCells)
{
    __shared__ uint32_t partial[1024];
    __shared__ uint64_t carry; // 64-bit: a total beyond 2^32-1 saturates below
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += blockDim.x)
    {
        uint32_t i = base + threadIdx.x;
        uint32_t area = 0;
        PathCells c;
        if (i < n)
        {
            c = cells[i];
            const uint32_t cv = pathCanvas[i];
            const int tilesX = cv != 0xffffffffu ? canvases[cv].tilesX : 0;
            const int tilesY = cv != 0xffffffffu ? canvases[cv].tilesY : 0;
            c.tx0 = max(c.tx0, 0);
            c.ty0 = max(c.ty0, 0);
            c.tx1 = min(c.tx1, tilesX);
            c.ty1 = min(c.ty1, tilesY);
            if (c.tx0 < c.tx1 && c.ty0 < c.ty1)
                area = (uint32_t)((c.tx1 - c.tx0) * (c.ty1 - c.ty0));
            else
            {
                c.tx0 = c.ty0 = 0;
                c.tx1 = c.ty1 = 0;
            }
        }
        partial[threadIdx.x] = area;
        __syncthreads();
        for (uint32_t off = 1; off < blockDim.x; off <<= 1)
        {
            uint32_t v = threadIdx.x >= off ? partial[threadIdx.x - off] : 0;
            __syncthreads();
            partial[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < n)
        {
            c.base = (uint32_t)carry + partial[threadIdx.x] - area;
            c.width = (uint32_t)(c.tx1 - c.tx0);
            cells[i] = c;
        }
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1)
            carry += partial[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *totalCells = (uint32_t)min(carry, (uint64_t)0xffffffffu);
}

// ---------------------------------------------------------------------------
// Binning.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void prim_tile_range(const Prim& p, const PathCells& c, int* tx0, int* ty0,
                                                int* tx1, int* ty1)
{
    int32_t minx, miny, maxx, maxy;
    prim_extent(p, &minx, &miny, &maxx, &maxy);
    *tx0 = max(minx >> (kSubpixelLog2 + kTileLog2), c.tx0);
    *ty0 = max(miny >> (kSubpixelLog2 + kTileLog2), c.ty0);
    *tx1 = min((maxx >> (kSubpixelLog2 + kTileLog2)) + 1, c.tx1);
    *ty1 = min((maxy >> (kSubpixelLog2 + kTileLog2)) + 1, c.ty1);
}

// The extent along v of the part of a primitive inside the slab lo <= u <= hi
// (u, v: its four vertices' coordinates across and along the slab, each of
// magnitude below 2^24 so exact in float), from its two triangles' five edges
// clipped to the slab, widened by the float error; false when no edge meets
// the slab.  The part of a polygon inside a slab always has boundary inside
// it, so the clipped edges bound it.
__device__ __forceinline__ bool slab_extent(const float (&u)[4], const float (&v)[4], float lo, float hi,
                                            float* vlo, float* vhi)
{
    constexpr int ea[5] = {0, 1, 2, 1, 3}, eb[5] = {1, 2, 0, 3, 2};
    float rlo = INFINITY, rhi = -INFINITY, eps = 1.f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
    {
        float u0 = u[ea[k]], v0 = v[ea[k]], u1 = u[eb[k]], v1 = v[eb[k]];
        if (u0 > u1)
        {
            const float tu = u0, tv = v0;
            u0 = u1;
            v0 = v1;
            u1 = tu;
            v1 = tv;
        }
        if (u1 < lo || u0 > hi)
            continue;
        float va = v0, vb = v1;
        if (u0 < lo || u1 > hi)
        {
            // u1 > u0 here.  The slope is within 2^-22 relative, and each
            // clipped v within |dv| 2^-20 + |v| 2^-22 of the true one.
            const float dv = v1 - v0;
            const float m = dv * __builtin_amdgcn_rcpf(u1 - u0);
            if (u0 < lo)
                va = v0 + (lo - u0) * m;
            if (u1 > hi)
                vb = v0 + (hi - u0) * m;
            eps = fmaxf(eps, fabsf(dv) * (1.f / (1 << 19)) + (fabsf(va) + fabsf(vb)) * (1.f / (1 << 21)));
        }
        rlo = fminf(rlo, fminf(va, vb));
        rhi = fmaxf(rhi, fmaxf(va, vb));
    }
    *vlo = rlo - eps;
    *vhi = rhi + eps;
    return rlo <= rhi;
}

// "q right of edge a->b" for the ray-left crossing rule, with the sample point
// perturbed by (+e, +e^2): s*D(q) >= 0 where s = +1 for upward edges.
__device__ __forceinline__ bool right_of(int32_t ax, int32_t ay, int32_t bx, int32_t by, int64_t qx,
                                         int64_t qy)
{
    int64_t D = (int64_t)(bx - ax) * (qy - ay) - (int64_t)(by - ay) * (qx - ax);
    return by < ay ? D >= 0 : D <= 0;
}

// The cells of a path's grid that primitive p can touch (row-major,
// grid-local index), passed to visit(cell).  The raster reads a tile's
// primitives at the points of [X0, X0 + 3968] x [Y0, Y0 + 3968] (subpixels:
// sample points, sample-line crossings right of X0, the left edge's
// crossings), so a primitive spanning several tile rows and columns visits,
// in each tile row, the tiles whose region widened by 64 meets the part of
// the primitive in that row's widened slab.
template <typename F>
__device__ __forceinline__ void for_prim_cells(const Prim& p, const PathCells& c, F visit)
{
    constexpr int kTileSub = 1 << (kSubpixelLog2 + kTileLog2);
    constexpr float kNear = 64.f, kFar = (float)(kTileSub - (1 << kSubpixelLog2) / 2) + 64.f;
    int tx0, ty0, tx1, ty1;
    prim_tile_range(p, c, &tx0, &ty0, &tx1, &ty1);
    float vx[4], vy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        vx[k] = (float)p.p[2 * k];
        vy[k] = (float)p.p[2 * k + 1];
    }
    // A feathered segment is two quads, [E0, T0, E1, T1] and its mirror
    // [E0, B0, E1, B1] (B = 2E - T): its part in a slab is bounded by both.
    const bool feather = (p.flags & kPrimFeather) != 0u;
    float bx[4], by[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const int e = k & 2; // the edge point of this corner's end
        bx[k] = (k & 1) ? (float)(2 * p.p[2 * e] - p.p[2 * k]) : vx[k];
        by[k] = (k & 1) ? (float)(2 * p.p[2 * e + 1] - p.p[2 * k + 1]) : vy[k];
    }
    const bool trim = ty1 - ty0 > 1 && tx1 - tx0 > 1;
    for (int ty = ty0; ty < ty1; ++ty)
    {
        int a = tx0, b = tx1 - 1;
        if (trim)
        {
            const float Y0 = (float)(ty * kTileSub);
            float xlo, xhi;
            bool any = slab_extent(vy, vx, Y0 - kNear, Y0 + kFar, &xlo, &xhi);
            if (feather)
            {
                float xlo2, xhi2;
                if (slab_extent(by, bx, Y0 - kNear, Y0 + kFar, &xlo2, &xhi2))
                {
                    xlo = any ? fminf(xlo, xlo2) : xlo2;
                    xhi = any ? fmaxf(xhi, xhi2) : xhi2;
                    any = true;
                }
            }
            if (!any)
                continue;
            // tile tx's widened region spans [tx*4096 - 64, tx*4096 + 3968 + 64]
            a = max(a, (int)ceilf(fmaxf((xlo - kFar) * (1.f / kTileSub), -1.f)));
            b = min(b, (int)floorf(fminf((xhi + kNear) * (1.f / kTileSub), (float)c.tx1)));
        }
        for (int tx = a; tx <= b; ++tx)
            visit((uint32_t)(ty - c.ty0) * c.width + (uint32_t)(tx - c.tx0));
    }
}

// Corner crossings of a fill primitive's inner edge I0->I1 on tile-row lines:
// visit(cell, +-1) at the first corner of each crossed row right of the edge.
template <typename F>
__device__ __forceinline__ void for_prim_crossings(const Prim& p, const PathCells& c, F visit)
{
    if (!(p.flags & kPrimFill) || (p.flags & kPrimNoEdge))
        return;
    const int32_t ax = p.p[0], ay = p.p[1], bx = p.p[4], by = p.p[5];
    if (ay == by)
        return;
    // NEGATE_PATH_FILL_COVERAGE negates the whole fan, backdrop included.
    const int s = (by < ay ? 1 : -1) * ((p.flags & kPrimNegate) ? -1 : 1);
    const int32_t ymin = min(ay, by), ymax = max(ay, by);
    const int unit = 1 << (kSubpixelLog2 + kTileLog2);
    const int tyA = max(c.ty0, (ymin + unit - 1) >> (kSubpixelLog2 + kTileLog2));
    for (int ty = tyA; ty < c.ty1; ++ty)
    {
        const int64_t Y = (int64_t)ty * unit;
        if (!(ymin <= Y && Y < ymax))
            break;
        // Smallest column c with corner (c*unit, Y) right of the edge.
        const double xe = ax + (double)(bx - ax) * (double)(Y - ay) / (double)(by - ay);
        int64_t cc = (int64_t)floor(xe / unit) + 1;
        while (right_of(ax, ay, bx, by, (cc - 1) * unit, Y))
            --cc;
        while (!right_of(ax, ay, bx, by, cc * unit, Y))
            ++cc;
        const int col = (int)max<int64_t>(cc, (int64_t)c.tx0);
        if (col < c.tx1)
            visit((uint32_t)(ty - c.ty0) * c.width + (uint32_t)(col - c.tx0), s);
    }
}

// Binning per draw: one 256-thread workgroup takes one draw's primitives and
// accumulates its (primitive, tile) counts and corner crossings over its
// path's cell grid in LDS, then adds them to the global grids once per cell.
// Grids larger than kBinLdsCells use global atomics directly.
constexpr uint32_t kBinLdsCells = 8192;

__device__ __forceinline__ uint32_t grid_cells(const PathCells& c)
{
    return c.tx0 < c.tx1 && c.ty0 < c.ty1 ? c.width * (uint32_t)(c.ty1 - c.ty0) : 0u;
}

__global__ __launch_bounds__(256) void bin_count_draw_kernel(const Prim* __restrict__ prims,
                                                             const DrawTess* __restrict__ draws,
                                                             const PathCells* __restrict__ cells,
                                                             uint32_t* __restrict__ cellCount,
                                                             int32_t* __restrict__ cellBackdrop,
                                                             uint32_t* __restrict__ pathDraws)
{
    // One word per cell: the primitive count in the low 16 bits and the
    // corner-crossing sum, two's complement, in the high 16.  A draw with
    // fewer than 32768 primitives keeps the count below 2^15 (no carry) and
    // the sum within +-2^15 (each primitive adds at most +-1 to a cell), so
    // both halves are exact; larger draws use global atomics.  32 KB of LDS
    // keeps five workgroups per CU.
    __shared__ uint32_t sCell[kBinLdsCells];
    const DrawTess d = draws[blockIdx.x];
    if (d.pathID == 0 || d.forwardCount == 0)
        return;
    const PathCells c = cells[d.pathID];
    const uint32_t nCells = grid_cells(c);
    if (nCells == 0)
        return;
    if (threadIdx.x == 0)
        atomicAdd(&pathDraws[d.pathID], 1u); // draws binning into this path's grid
    const bool local = nCells <= kBinLdsCells && d.forwardCount < 32768u;
    if (local)
    {
        for (uint32_t i = threadIdx.x; i < nCells; i += blockDim.x)
            sCell[i] = 0;
        __syncthreads();
    }
    uint32_t* gCount = cellCount + c.base;
    int32_t* gBack = cellBackdrop + c.base;
    for (uint32_t k = threadIdx.x; k < d.forwardCount; k += blockDim.x)
    {
        const Prim p = prims[d.primBase + k];
        if (p.flags & kPrimEmpty)
            continue;
        if (local)
        {
            for_prim_cells(p, c, [&](uint32_t cell) { atomicAdd(&sCell[cell], 1u); });
            for_prim_crossings(p, c, [&](uint32_t cell, int v) { atomicAdd(&sCell[cell], (uint32_t)v << 16); });
        }
        else
        {
            for_prim_cells(p, c, [&](uint32_t cell) { atomicAdd(&gCount[cell], 1u); });
            for_prim_crossings(p, c, [&](uint32_t cell, int v) { atomicAdd(&gBack[cell], v); });
        }
    }
    if (local)
    {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nCells; i += blockDim.x)
        {
            const uint32_t w = sCell[i];
            if ((w & 0xffffu) != 0)
                atomicAdd(&gCount[i], w & 0xffffu);
            if ((w >> 16) != 0)
                atomicAdd(&gBack[i], (int32_t)w >> 16);
        }
    }
}

// Writes each primitive index into its cells' bins.  A path binned by this
// draw alone owns its cells' ranges, so LDS cursors from 0 hand out slots
// after the cell offsets; a path shared by several draws counts in LDS,
// reserves once per cell on the global cursors, then hands out slots.
__global__ __launch_bounds__(256) void bin_scatter_draw_kernel(const Prim* __restrict__ prims,
                                                               const DrawTess* __restrict__ draws,
                                                               const PathCells* __restrict__ cells,
                                                               const uint32_t* __restrict__ cellOffset,
                                                               uint32_t* __restrict__ cellCursor,
                                                               const uint32_t* __restrict__ pathDraws,
                                                               uint32_t* __restrict__ binned)
{
    __shared__ uint32_t sCur[kBinLdsCells];
    const DrawTess d = draws[blockIdx.x];
    if (d.pathID == 0 || d.forwardCount == 0)
        return;
    const PathCells c = cells[d.pathID];
    const uint32_t nCells = grid_cells(c);
    if (nCells == 0)
        return;
    const uint32_t* gOffset = cellOffset + c.base;
    uint32_t* gCursor = cellCursor + c.base;
    if (nCells > kBinLdsCells)
    {
        for (uint32_t k = threadIdx.x; k < d.forwardCount; k += blockDim.x)
        {
            const uint32_t g = d.primBase + k;
            const Prim p = prims[g];
            if (p.flags & kPrimEmpty)
                continue;
            for_prim_cells(p, c, [&](uint32_t cell) { binned[gOffset[cell] + atomicAdd(&gCursor[cell], 1u)] = g; });
        }
        return;
    }
    if (pathDraws[d.pathID] == 1u)
    {
        // the cursors start at the cells' bin offsets (read once, coalesced)
        for (uint32_t i = threadIdx.x; i < nCells; i += blockDim.x)
            sCur[i] = gOffset[i];
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < d.forwardCount; k += blockDim.x)
        {
            const uint32_t g = d.primBase + k;
            const Prim p = prims[g];
            if (!(p.flags & kPrimEmpty))
                for_prim_cells(p, c, [&](uint32_t cell) { binned[atomicAdd(&sCur[cell], 1u)] = g; });
        }
        return;
    }
    for (uint32_t i = threadIdx.x; i < nCells; i += blockDim.x)
        sCur[i] = 0;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < d.forwardCount; k += blockDim.x)
    {
        const Prim p = prims[d.primBase + k];
        if (!(p.flags & kPrimEmpty))
            for_prim_cells(p, c, [&](uint32_t cell) { atomicAdd(&sCur[cell], 1u); });
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nCells; i += blockDim.x)
        if (sCur[i] != 0)
            sCur[i] = gOffset[i] + atomicAdd(&gCursor[i], sCur[i]);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < d.forwardCount; k += blockDim.x)
    {
        const uint32_t g = d.primBase + k;
        const Prim p = prims[g];
        if (!(p.flags & kPrimEmpty))
            for_prim_cells(p, c, [&](uint32_t cell) { binned[atomicAdd(&sCur[cell], 1u)] = g; });
    }
}

// Each cell's first bin and corner winding side by side, so the raster reads
// one 8-byte entry (and the next entry's offset for the count) per
// (path, tile) instead of two arrays.
// Prefix-sum of corner crossings along each row of each path's grid: one
// wave64 per (path, row), 64 cells per step with a carried running sum.  Each
// cell's result goes out with its bin offset as the raster's packed
// (offset, corner winding) entry; entry totalCells is the end sentinel.
__global__ __launch_bounds__(256) void backdrop_prefix_kernel(const PathCells* __restrict__ cells,
                                                              uint32_t pathCount, int maxRows,
                                                              const int32_t* __restrict__ cellBackdrop,
                                                              const uint32_t* __restrict__ cellOffset,
                                                              uint32_t totalCells, uint2* __restrict__ packed)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        packed[totalCells] = make_uint2(cellOffset[totalCells], 0u);
    if (maxRows <= 0)
        return;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const uint32_t path = wave / (uint32_t)maxRows;
    const int row = (int)(wave % (uint32_t)maxRows);
    if (path >= pathCount)
        return;
    const PathCells c = cells[path];
    if (row >= c.ty1 - c.ty0)
        return;
    const uint32_t rowBase = c.base + (uint32_t)row * c.width;
    const int32_t* r = cellBackdrop + rowBase;
    int32_t carry = 0;
    for (uint32_t x0 = 0; x0 < c.width; x0 += 64)
    {
        const uint32_t x = x0 + (uint32_t)lane;
        int32_t v = x < c.width ? r[x] : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1)
        {
            const int32_t t = __shfl_up(v, d);
            if (lane >= d)
                v += t;
        }
        if (x < c.width)
            packed[rowBase + x] = make_uint2(cellOffset[rowBase + x], (uint32_t)(carry + v));
        carry += __shfl(v, 63);
    }
}

// ---------------------------------------------------------------------------
// Tile raster.
// ---------------------------------------------------------------------------

// Half-plane test of edge a->b against a tile's pixels, as an integer
// threshold on  k(i, j) = -dy*(2i+1) + dx*(2j+1),  where the edge function at
// pixel (i, j) is D = D0 + 128*k and D0 is the edge function at the tile's
// top-left corner.
struct EdgeT
{
    int32_t A, B; // -dy, dx
    int64_t D0;
};
__device__ __forceinline__ EdgeT make_edge(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t X0,
                                           int32_t Y0)
{
    EdgeT e;
    int32_t dx = bx - ax, dy = by - ay;
    e.A = -dy;
    e.B = dx;
    e.D0 = (int64_t)dx * (int64_t)(Y0 - ay) - (int64_t)dy * (int64_t)(X0 - ax);
    return e;
}
// Width-generic forms of the edge helpers for the per-row work (instantiated
// for int64).
template <typename I> struct EdgeTI
{
    int32_t A, B; // -dy, dx
    I D0;
};
template <typename I>
__device__ __forceinline__ EdgeTI<I> make_edge_t(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t X0,
                                                 int32_t Y0)
{
    EdgeTI<I> e;
    const int32_t dx = bx - ax, dy = by - ay;
    e.A = -dy;
    e.B = dx;
    e.D0 = (I)dx * (I)(Y0 - ay) - (I)dy * (I)(X0 - ax);
    return e;
}
template <typename I> __device__ __forceinline__ int32_t clamp_thr_t(I t)
{
    const I lim = (I)1 << 30;
    return (int32_t)(t < -lim ? -lim : (t > lim ? lim : t));
}
template <> __device__ __forceinline__ int32_t clamp_thr_t<double>(double t)
{
    return (int32_t)fmin(fmax(t, -1073741824.0), 1073741824.0);
}
template <typename I> __device__ __forceinline__ I ceil_div128_t(I a) { return -((-a) >> 7); }
template <typename I> __device__ __forceinline__ I floor_div128_t(I a) { return a >> 7; }
// Edge functions in double: coordinates lie in [-2^23, 2^23) (1/256 px), so
// every product is below 2^49 and every sum below 2^50 -- exact, as are the
// divisions by 128 and their floor and ceiling.
template <> __device__ __forceinline__ double ceil_div128_t<double>(double a) { return ceil(a * (1.0 / 128.0)); }
template <> __device__ __forceinline__ double floor_div128_t<double>(double a) { return floor(a * (1.0 / 128.0)); }
template <typename I> __device__ __forceinline__ int32_t inside_threshold_t(const EdgeTI<I>& e, int o)
{
    const int pert = e.A != 0 ? (e.A > 0 ? 1 : -1) : (e.B > 0 ? 1 : (e.B < 0 ? -1 : 0));
    const I oD0 = o > 0 ? e.D0 : -e.D0;
    const I t = (o * pert > 0) ? ceil_div128_t<I>(-oD0) : floor_div128_t<I>(-oD0) + 1;
    return clamp_thr_t<I>(t);
}
template <typename I>
__device__ __forceinline__ I orient_t(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t qx, int32_t qy)
{
    return (I)(bx - ax) * (I)(qy - ay) - (I)(by - ay) * (I)(qx - ax);
}
__device__ __forceinline__ float dither_noise(int x, int y)
{
    // common.glsl:269-274, scale 1/256, bias -1/512 (gpu.cpp:822-824).
    float fx = float(x) + .5f, fy = float(y) + .5f;
    float v = 0.06711056f * fx + 0.00583715f * fy;
    float v1 = v - floorf(v);
    float w = 52.9829189f * v1;
    float v2 = w - floorf(w);
    return v2 * (1.f / 256.f) + (-.5f / 256.f);
}
// One 256-thread workgroup (four waves) per 16x16 tile; thread t owns pixel
// (t % 16, t / 16).  Paths are visited in submission order and resolved per
// pixel in that order.
//
// The paths that touch the tile (primitives binned to it, or a nonzero
// corner winding) are listed kWindow at a time and cut into batches of at most
// kBatchEntries paths, kAccArrays accumulator arrays and, unless the batch
// has a single path with primitives, kBatchPrims primitives.  Each path with
// primitives owns one LDS accumulator array, two when it has more than
// kPackedPrims primitives in the tile:
//   sLeft[a][row]     winding change along the tile's left edge, as a suffix
//                     delta over rows;
//   sAcc[a][row][i]   fills: +-1 where an inner edge crosses the row's sample
//                     line between pixels i-1 and i, times 2^24, plus the
//                     signed AA-border coverage in 1/65536 fixed point (a
//                     two-array path keeps the crossings in sAcc[a] and the
//                     border in sAcc[a + 1]); strokes: max coverage as float
//                     bits.
// With at most kPackedPrims primitives a pixel's crossing count stays within
// +-127 and its border sum within +-2^23, so the packed form is exact.
// Threads first take one primitive each (left-edge terms and the range of
// sample rows), then the (primitive, row) work items are compacted across the
// workgroup, fills before strokes.  Every thread then resolves its pixel for
// each path of the batch, turning the path's accumulators into coverage on
// the way (prefix of the left-edge deltas and of the row's crossings, fill
// rule).  The integer predicates are those of a per-pixel evaluation:
// crossing columns solve the same threshold inequality, and a triangle's
// pixels on a row are the candidates of a float span estimate (widened by its
// error bound) that pass the exact three-edge test.
constexpr int kWgThreads = 256;
constexpr int kAccArrays = 16;
constexpr int kBatchEntries = 64; // one wave's ballot
constexpr int kWindow = 128;      // paths listed per round (LDS: six workgroups per CU)
constexpr uint32_t kBatchPrims = kWgThreads;
constexpr uint32_t kPackedPrims = 127;
constexpr int kMaxItems = kBatchPrims * kTile;

struct TileGeom
{
    int32_t X0, Y0;
};

__device__ __forceinline__ void tri_row_candidates(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t cx,
                                                   int32_t cy, int32_t yq, int32_t X0, int* ilo, int* ihi)
{
    // Tile-relative x-range of the triangle on the row's sample line, as a
    // float estimate widened by its error bound (an fdividef of relative
    // error < 2^-21 scaled by the edge's width, plus rounding of the ends).
    float xl = INFINITY, xr = -INFINITY, eps = 8.f;
    const int32_t vx[3] = {ax - X0, bx - X0, cx - X0}, vy[3] = {ay, by, cy};
#pragma unroll
    for (int e = 0; e < 3; ++e)
    {
        const int32_t x0 = vx[e], y0 = vy[e], x1 = vx[(e + 1) % 3], y1 = vy[(e + 1) % 3];
        if ((y0 <= yq && yq <= y1) || (y1 <= yq && yq <= y0))
        {
            if (y0 == y1)
            {
                xl = fminf(xl, (float)min(x0, x1));
                xr = fmaxf(xr, (float)max(x0, x1));
            }
            else
            {
                const float dx = (float)(x1 - x0);
                const float t = __fdividef((float)(yq - y0), (float)(y1 - y0));
                const float x = (float)x0 + t * dx;
                xl = fminf(xl, x);
                xr = fmaxf(xr, x);
                eps = fmaxf(eps, fabsf(dx) * (1.f / (1 << 18)) + fabsf(x) * (1.f / (1 << 20)));
            }
        }
    }
    if (!(xl <= xr))
    {
        *ilo = 1;
        *ihi = 0;
        return;
    }
    // pixel i's sample x is 256 i + 128 (tile-relative): the pixels whose
    // samples fall in [xl - eps, xr + eps]
    const float lo = (xl - eps - 128.f) * (1.f / 256.f);
    const float hi = (xr + eps - 128.f) * (1.f / 256.f);
    *ilo = (int)fmaxf(ceilf(lo), 0.f);
    *ihi = (int)fminf(floorf(hi), (float)(kTile - 1));
}

// A border triangle's barycentric coverage at a pixel inside it, in
// 1/65536 fixed point, clamped to [0, 1] (it leaves that range only by
// rounding), so that a sum over 127 triangles stays below 2^23.
__device__ __forceinline__ int32_t border_fixed(float cv)
{
    return __float2int_rn(fminf(fmaxf(cv, 0.f), 1.f) * 65536.f);
}

// Exact "inside" test of an edge against pixel (i2, j2) with orientation o.
struct EdgeK
{
    int32_t A, B, thr; // o-premultiplied: inside iff A*i2 + B*j2 >= thr
};
template <typename I> __device__ __forceinline__ EdgeK edge_k(const EdgeTI<I>& e, int o)
{
    EdgeK k;
    k.A = o * e.A;
    k.B = o * e.B;
    k.thr = inside_threshold_t<I>(e, o);
    return k;
}

// Sample rows [*r0, *r1] of the tile that a primitive can touch inside the
// tile's columns.  The per-row work looks at sample points with x in
// [X0, X0 + 16 px] only (inner-edge crossings left of X0 are the left-edge
// winding's), so the rows are those where the primitive meets that slab: the
// y-extent of its five edges (those of both triangles, the inner edge among
// them) clipped to the slab widened by 64 subpixels, taken in float and
// widened by the float error, and intersected with the quad's exact y-extent.
// A part of a polygon inside a vertical slab always has boundary inside it, so
// the clipped edges bound it.
__device__ __forceinline__ void prim_rows(const Prim& p, const TileGeom& g, int* r0, int* r1)
{
    int32_t minx, miny, maxx, maxy;
    prim_extent(p, &minx, &miny, &maxx, &maxy);
    const int32_t lo = miny - g.Y0 - (1 << (kSubpixelLog2 - 1));
    const int32_t hi = maxy - g.Y0 - (1 << (kSubpixelLog2 - 1));
    int a0 = max(-((-lo) >> kSubpixelLog2), 0);   // ceil(lo / 256)
    int a1 = min(hi >> kSubpixelLog2, kTile - 1); // floor(hi / 256)
    const bool feather = (p.flags & kPrimFeather) != 0u;

    // tile-relative vertices: |coordinate| < 2^24, exact in float
    float vx[4], vy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        vx[k] = (float)(p.p[2 * k] - g.X0);
        vy[k] = (float)(p.p[2 * k + 1] - g.Y0);
    }
    float ylo, yhi;
    bool any = slab_extent(vx, vy, -64.f, (float)(kTile << kSubpixelLog2) + 64.f, &ylo, &yhi);
    if (feather)
    {
        // and the mirrored quad [E0, B0, E1, B1]
        float bx[4], by[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            bx[k] = (k & 1) ? 2.f * vx[k & 2] - vx[k] : vx[k];
            by[k] = (k & 1) ? 2.f * vy[k & 2] - vy[k] : vy[k];
        }
        float ylo2, yhi2;
        if (slab_extent(bx, by, -64.f, (float)(kTile << kSubpixelLog2) + 64.f, &ylo2, &yhi2))
        {
            ylo = any ? fminf(ylo, ylo2) : ylo2;
            yhi = any ? fmaxf(yhi, yhi2) : yhi2;
            any = true;
        }
    }
    if (!any)
    {
        *r0 = 1;
        *r1 = 0;
        return;
    }
    const float half = (float)(1 << (kSubpixelLog2 - 1)), inv = 1.f / (float)(1 << kSubpixelLog2);
    a0 = max(a0, (int)ceilf(fmaxf((ylo - half) * inv, -1.f)));
    a1 = min(a1, (int)floorf(fminf((yhi - half) * inv, (float)kTile)));
    *r0 = a0;
    *r1 = a1;
}

// Left-edge winding of a fill primitive's inner edge I0 -> I1: the change of
// the winding number from the tile's corner down to each row's sample line,
// which is 0 above some row and a constant +-1 from it on.
__device__ void fill_prim_left(const Prim& p, const TileGeom& g, int32_t* __restrict__ sLeft)
{
    const int32_t X0 = g.X0, Y0 = g.Y0;
    const int32_t xmin = min(p.p[0], p.p[4]), xmax = max(p.p[0], p.p[4]);
    if (!(xmin <= X0 && X0 < xmax) || (p.flags & kPrimNoEdge))
        return;
    const int neg = (p.flags & kPrimNegate) ? -1 : 1;
    if (p.p[1] != p.p[5])
    {
        const EdgeT e = make_edge(p.p[0], p.p[1], p.p[4], p.p[5], X0, Y0);
        const int s = p.p[5] < p.p[1] ? 1 : -1;
        // R(q) <=> s*D(q) >= 0 <=> s*k >= ceil(-s*D0/128)
        const int32_t T = clamp_thr(ceil_div128(s > 0 ? -e.D0 : e.D0));
        const bool Rcorner = (s > 0 ? e.D0 : -e.D0) >= 0;
        // the first row whose sample-line point is on the other side of the
        // edge than the corner (rows as a bit mask: no divergent exits)
        uint32_t flipped = 0;
#pragma unroll
        for (int j = 0; j < kTile; ++j)
            flipped |= (uint32_t)((s * (e.B * (2 * j + 1)) >= T) != Rcorner) << j;
        if (flipped != 0)
            atomicAdd(&sLeft[__builtin_ctz(flipped)], neg * s * (Rcorner ? -1 : 1));
    }
    else if (p.p[0] != p.p[4])
    {
        // A horizontal edge never meets a sample line, but where it crosses
        // the tile's left edge the winding below it steps by sign(dx).
        const int32_t h = p.p[1];
        if (Y0 < h)
        {
            const int32_t v = h - Y0 - (1 << (kSubpixelLog2 - 1));
            const int jstar = max(-((-v) >> kSubpixelLog2), 0); // first row with yq >= h
            if (jstar < kTile)
                atomicAdd(&sLeft[jstar], neg * (p.p[4] > p.p[0] ? 1 : -1));
        }
    }
}

// One sample row of a fill primitive: the inner edge's crossing column and
// the two AA border triangles.
template <typename I>
__device__ void fill_prim_row(const Prim& p, int j, const TileGeom& g, int32_t* __restrict__ sDelta,
                              int32_t* __restrict__ sBorder, int32_t deltaUnit)
{
    const int32_t X0 = g.X0, Y0 = g.Y0;
    const int32_t j2 = 2 * j + 1;
    const int32_t yq = Y0 + (j << kSubpixelLog2) + (1 << (kSubpixelLog2 - 1));
    const int neg = (p.flags & kPrimNegate) ? -1 : 1;
    const EdgeTI<I> e = make_edge_t<I>(p.p[0], p.p[1], p.p[4], p.p[5], X0, Y0);
    if (p.p[1] != p.p[5] && !(p.flags & kPrimNoEdge))
    {
        const int32_t ymin = min(p.p[1], p.p[5]), ymax = max(p.p[1], p.p[5]);
        if (ymin <= yq && yq < ymax)
        {
            const int s = p.p[5] < p.p[1] ? 1 : -1;
            const int32_t T = clamp_thr_t<I>(ceil_div128_t<I>(s > 0 ? -e.D0 : e.D0));
            const int32_t krow = e.B * j2;
            if (!(s * krow >= T))
            {
                // Smallest i with s*(A*(2i+1) + krow) >= T, where s*A = |dy|:
                // 2|dy| i >= T - |dy| - s*krow.  The quotient is exact in
                // double wherever it decides a column inside the tile.
                const int64_t sa2 = 2 * (int64_t)(s * e.A);
                const int64_t r = (int64_t)T - sa2 / 2 - (int64_t)s * krow;
                if (r <= sa2 * (kTile - 1))
                {
                    // ceil(r / sa2) in [0, kTile): a float estimate, then exact
                    int q = r <= 0 ? 0 : (int)ceilf(__fdividef((float)r, (float)sa2));
                    q = min(max(q, 0), kTile - 1);
                    while (q > 0 && sa2 * (q - 1) >= r)
                        --q;
                    while (sa2 * q < r)
                        ++q;
                    atomicAdd(&sDelta[j * kTile + q], neg * s * deltaUnit);
                }
            }
        }
    }
    // Border triangle T1 = (I0, O0, I1), coverage 1 - lambda(O0).
    const I A1 = (p.flags & kPrimNoT1) ? (I)0 : orient_t<I>(p.p[0], p.p[1], p.p[2], p.p[3], p.p[4], p.p[5]);
    if (A1 != 0)
    {
        int ilo, ihi;
        tri_row_candidates(p.p[0], p.p[1], p.p[2], p.p[3], p.p[4], p.p[5], yq, X0, &ilo, &ihi);
        if (ilo <= ihi)
        {
            const int o = A1 > 0 ? 1 : -1;
            const EdgeK k01 = edge_k<I>(make_edge_t<I>(p.p[0], p.p[1], p.p[2], p.p[3], X0, Y0), o);
            const EdgeK k12 = edge_k<I>(make_edge_t<I>(p.p[2], p.p[3], p.p[4], p.p[5], X0, Y0), o);
            const EdgeK k20 = edge_k<I>(make_edge_t<I>(p.p[4], p.p[5], p.p[0], p.p[1], X0, Y0), o);
            const float inv = __builtin_amdgcn_rcpf((float)A1);
            // edge values at pixel ilo, stepped exactly by 2A per pixel
            const int32_t i2 = 2 * ilo + 1;
            int32_t v01 = k01.A * i2 + k01.B * j2, v12 = k12.A * i2 + k12.B * j2, v20 = k20.A * i2 + k20.B * j2;
            int32_t ke = e.A * i2 + e.B * j2;
            const float D0 = (float)e.D0;
            for (int i = ilo; i <= ihi; ++i)
            {
                if (v01 >= k01.thr && v12 >= k12.thr && v20 >= k20.thr)
                {
                    const float De = D0 + 128.f * (float)ke;
                    atomicAdd(&sBorder[j * kTile + i], neg * o * border_fixed(1.f + De * inv));
                }
                v01 += 2 * k01.A;
                v12 += 2 * k12.A;
                v20 += 2 * k20.A;
                ke += 2 * e.A;
            }
        }
    }
    // Border triangle T2 = (I1, O0, O1), coverage lambda(I1).
    const I A2 = (p.flags & kPrimNoT2) ? (I)0 : orient_t<I>(p.p[4], p.p[5], p.p[2], p.p[3], p.p[6], p.p[7]);
    if (A2 != 0)
    {
        int ilo, ihi;
        tri_row_candidates(p.p[4], p.p[5], p.p[2], p.p[3], p.p[6], p.p[7], yq, X0, &ilo, &ihi);
        if (ilo <= ihi)
        {
            const int o = A2 > 0 ? 1 : -1;
            const EdgeK f01 = edge_k<I>(make_edge_t<I>(p.p[4], p.p[5], p.p[2], p.p[3], X0, Y0), o);
            const EdgeTI<I> f12e = make_edge_t<I>(p.p[2], p.p[3], p.p[6], p.p[7], X0, Y0);
            const EdgeK f12 = edge_k(f12e, o);
            const EdgeK f20 = edge_k<I>(make_edge_t<I>(p.p[6], p.p[7], p.p[4], p.p[5], X0, Y0), o);
            const float inv = __builtin_amdgcn_rcpf((float)A2);
            const int32_t i2 = 2 * ilo + 1;
            int32_t v01 = f01.A * i2 + f01.B * j2, v12 = f12.A * i2 + f12.B * j2, v20 = f20.A * i2 + f20.B * j2;
            int32_t k12 = f12e.A * i2 + f12e.B * j2;
            const float D0 = (float)f12e.D0;
            for (int i = ilo; i <= ihi; ++i)
            {
                if (v01 >= f01.thr && v12 >= f12.thr && v20 >= f20.thr)
                {
                    const float D12 = D0 + 128.f * (float)k12;
                    atomicAdd(&sBorder[j * kTile + i], neg * o * border_fixed(D12 * inv));
                }
                v01 += 2 * f01.A;
                v12 += 2 * f12.A;
                v20 += 2 * f20.A;
                k12 += 2 * f12e.A;
            }
        }
    }
}

// eval_feathered_fill (draw_path_common.glsl:153-240): the feather above y0
// plus a 4-sample midpoint integral of FEATHER(u) FEATHER'(t) across a
// corner, signed by coverages.x.
__device__ __forceinline__ float eval_feathered_fill(const float* c, const float* __restrict__ table)
{
    const float cotTheta = c[2];
    const float y0 = fmaxf(c[3], 0.f);
    float cov = cotTheta >= 0.f ? feather_table(table, y0) : 0.f;
    if (fabsf(cotTheta) < 1e3f)
    {
        const float x = fabsf(c[0]) - .25f;
        const float y = -c[1] - 2.f;
        const float dt = (y - y0) * 0.5984134206f;
        const float k[4] = {0.20888568955f, 0.62665706865f, 1.04442844776f, 1.46219982687f};
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            const float t = y0 + dt * k[i];
            const float u = t * -cotTheta + (y * cotTheta + x);
            const float t_ = t * 5.09593080173f + -2.54796540086f;
            sum += feather_table(table, u) * exp2f(-t_ * t_);
        }
        cov += sum * dt;
    }
    return cov * signf_(c[0]);
}

// One sample row of one triangle of a feathered segment: the coverage
// vectors interpolated at each pixel inside it, evaluated, and added with the
// triangle's orientation (the fixed-point border units of the fill rows).
__device__ __forceinline__ void feather_tri_row(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t cx,
                                                int32_t cy, const float* ca, const float* cb, const float* cc,
                                                int neg, int j, const TileGeom& g, int32_t* __restrict__ sBorder,
                                                const float* __restrict__ table)
{
    typedef double I;
    const I A = orient_t<I>(ax, ay, bx, by, cx, cy);
    if (A == 0)
        return;
    const int32_t X0 = g.X0, Y0 = g.Y0;
    const int32_t j2 = 2 * j + 1;
    const int32_t yq = Y0 + (j << kSubpixelLog2) + (1 << (kSubpixelLog2 - 1));
    int ilo, ihi;
    tri_row_candidates(ax, ay, bx, by, cx, cy, yq, X0, &ilo, &ihi);
    if (ilo > ihi)
        return;
    const int o = A > 0 ? 1 : -1;
    const EdgeTI<I> eab = make_edge_t<I>(ax, ay, bx, by, X0, Y0);
    const EdgeTI<I> ebc = make_edge_t<I>(bx, by, cx, cy, X0, Y0);
    const EdgeTI<I> eca = make_edge_t<I>(cx, cy, ax, ay, X0, Y0);
    const EdgeK kab = edge_k<I>(eab, o), kbc = edge_k<I>(ebc, o), kca = edge_k<I>(eca, o);
    const float inv = 1.f / (float)A;
    const int32_t i2 = 2 * ilo + 1;
    int32_t vab = kab.A * i2 + kab.B * j2, vbc = kbc.A * i2 + kbc.B * j2, vca = kca.A * i2 + kca.B * j2;
    int32_t wab = eab.A * i2 + eab.B * j2, wbc = ebc.A * i2 + ebc.B * j2, wca = eca.A * i2 + eca.B * j2;
    const float dab = (float)eab.D0, dbc = (float)ebc.D0, dca = (float)eca.D0;
    for (int i = ilo; i <= ihi; ++i)
    {
        if (vab >= kab.thr && vbc >= kbc.thr && vca >= kca.thr)
        {
            // barycentrics: the edge opposite each vertex over the area
            const float la = (dbc + 128.f * (float)wbc) * inv;
            const float lb = (dca + 128.f * (float)wca) * inv;
            const float lc = (dab + 128.f * (float)wab) * inv;
            float c4[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                c4[q] = la * ca[q] + lb * cb[q] + lc * cc[q];
            const float v = fminf(fmaxf(eval_feathered_fill(c4, table), -1.f), 1.f);
            atomicAdd(&sBorder[j * kTile + i], neg * o * __float2int_rn(v * 65536.f));
        }
        vab += 2 * kab.A;
        vbc += 2 * kbc.A;
        vca += 2 * kca.A;
        wab += 2 * eab.A;
        wbc += 2 * ebc.A;
        wca += 2 * eca.A;
    }
}

// One sample row of a feathered segment [E0, T0, E1, T1]: the edge's crossing
// column and its four border triangles (the inner points are 2E - T).  Kept
// out of line, reading the primitive itself, so the fill and stroke rows keep
// their registers.
__device__ __noinline__ void feather_prim_row(const Prim* __restrict__ prims, uint32_t idx, int j, int32_t X0,
                                              int32_t Y0, int32_t* __restrict__ sDelta,
                                              int32_t* __restrict__ sBorder, int32_t deltaUnit,
                                              const float* __restrict__ table)
{
    Prim p = prims[idx];
    const TileGeom g{X0, Y0};
    {
        Prim edge = p;
        edge.flags |= kPrimNoT1 | kPrimNoT2;
        fill_prim_row<double>(edge, j, g, sDelta, sBorder, deltaUnit);
    }
    const int neg = (p.flags & kPrimNegate) ? -1 : 1;
    float cE0[4] = {.5f, -2.f, 1e6f, .5f}, cT0[4] = {0.f, -2.f, 1e6f, 0.f};
    float cE1[4] = {.5f, -2.f, 1e6f, .5f}, cT1[4] = {0.f, -2.f, 1e6f, 0.f};
    if (p.flags & kPrimJoin0)
    {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cT0[q] = p.cov[q];
        cE0[0] = .75f;
        cE0[1] = -2.5f;
        cE0[2] = cT0[2];
        cE0[3] = u2f(p.pad[0]);
    }
    if (p.flags & kPrimJoin1)
    {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cT1[q] = p.cov[4 + q];
        cE1[0] = .75f;
        cE1[1] = -2.5f;
        cE1[2] = cT1[2];
        cE1[3] = u2f(p.pad[1]);
    }
    const int32_t e0x = p.p[0], e0y = p.p[1], t0x = p.p[2], t0y = p.p[3];
    const int32_t e1x = p.p[4], e1y = p.p[5], t1x = p.p[6], t1y = p.p[7];
    const int32_t b0x = 2 * e0x - t0x, b0y = 2 * e0y - t0y, b1x = 2 * e1x - t1x, b1y = 2 * e1y - t1y;
    feather_tri_row(e0x, e0y, t0x, t0y, e1x, e1y, cE0, cT0, cE1, neg, j, g, sBorder, table);
    feather_tri_row(e1x, e1y, t0x, t0y, t1x, t1y, cE1, cT0, cT1, neg, j, g, sBorder, table);
    feather_tri_row(e0x, e0y, b0x, b0y, e1x, e1y, cE0, cT0, cE1, neg, j, g, sBorder, table);
    feather_tri_row(e1x, e1y, b0x, b0y, b1x, b1y, cE1, cT0, cT1, neg, j, g, sBorder, table);
}

// find_stroke_coverage (draw_path.vert:500-508): min(x, y), or for a
// feathered stroke (x biased below -1.5) 1 - FEATHER(1 - x) - FEATHER(1 - y)
// (eval_feathered_stroke, draw_path_common.glsl:242-258).
template <bool FEATHER>
__device__ __forceinline__ float stroke_coverage(float cx, float cy, const float* __restrict__ table)
{
    if (FEATHER && cx < -1.5f)
        return 1.f - feather_table(table, 3.f + cx) - feather_table(table, 1.f - cy);
    return fminf(cx, cy);
}

template <typename I, bool FEATHER = false>
__device__ void stroke_prim_row(const Prim& p, int j, const TileGeom& g, uint32_t* __restrict__ sMax,
                                const float* __restrict__ table = nullptr)
{
    const int32_t X0 = g.X0, Y0 = g.Y0;
    const int32_t j2 = 2 * j + 1;
    const int32_t yq = Y0 + (j << kSubpixelLog2) + (1 << (kSubpixelLog2 - 1));
    // T1 = (P0-, P0+, P1-): vertices 0, 1, 2; T2 = (P1-, P0+, P1+): 2, 1, 3.
    // Only clockwise triangles draw.
    const I A1 = (p.flags & kPrimNoT1) ? (I)0 : orient_t<I>(p.p[0], p.p[1], p.p[2], p.p[3], p.p[4], p.p[5]);
    const I A2 = (p.flags & kPrimNoT2) ? (I)0 : orient_t<I>(p.p[4], p.p[5], p.p[2], p.p[3], p.p[6], p.p[7]);
    const EdgeTI<I> e12 = make_edge_t<I>(p.p[2], p.p[3], p.p[4], p.p[5], X0, Y0);
    if (A1 > 0)
    {
        int ilo, ihi;
        tri_row_candidates(p.p[0], p.p[1], p.p[2], p.p[3], p.p[4], p.p[5], yq, X0, &ilo, &ihi);
        if (ilo <= ihi)
        {
            const EdgeTI<I> e01 = make_edge_t<I>(p.p[0], p.p[1], p.p[2], p.p[3], X0, Y0);
            const EdgeTI<I> e20 = make_edge_t<I>(p.p[4], p.p[5], p.p[0], p.p[1], X0, Y0);
            const int32_t t01 = inside_threshold_t<I>(e01, 1), t12 = inside_threshold_t<I>(e12, 1),
                          t20 = inside_threshold_t<I>(e20, 1);
            const float inv = __builtin_amdgcn_rcpf((float)A1);
            // Edge values at pixel ilo, stepped exactly by 2A per pixel; the
            // interpolated coverages are linear in the column too and step by
            // float adds (within 16 ulp of a direct evaluation).
            const int32_t i2 = 2 * ilo + 1;
            int32_t k01 = e01.A * i2 + e01.B * j2, k12 = e12.A * i2 + e12.B * j2, k20 = e20.A * i2 + e20.B * j2;
            const float l0 = ((float)e12.D0 + 128.f * (float)k12) * inv;
            const float l1 = ((float)e20.D0 + 128.f * (float)k20) * inv;
            const float l2 = ((float)e01.D0 + 128.f * (float)k01) * inv;
            const float s0 = 256.f * (float)e12.A * inv, s1 = 256.f * (float)e20.A * inv,
                        s2 = 256.f * (float)e01.A * inv;
            float cx = l0 * p.cov[0] + l1 * p.cov[2] + l2 * p.cov[4];
            float cy = l0 * p.cov[1] + l1 * p.cov[3] + l2 * p.cov[5];
            const float dcx = s0 * p.cov[0] + s1 * p.cov[2] + s2 * p.cov[4];
            const float dcy = s0 * p.cov[1] + s1 * p.cov[3] + s2 * p.cov[5];
            for (int i = ilo; i <= ihi; ++i)
            {
                if (k01 >= t01 && k12 >= t12 && k20 >= t20)
                {
                    const float v = stroke_coverage<FEATHER>(cx, cy, table);
                    if (v > 0.f)
                        atomicMax(&sMax[j * kTile + i], __float_as_uint(v));
                }
                k01 += 2 * e01.A;
                k12 += 2 * e12.A;
                k20 += 2 * e20.A;
                cx += dcx;
                cy += dcy;
            }
        }
    }
    if (A2 > 0)
    {
        int ilo, ihi;
        tri_row_candidates(p.p[4], p.p[5], p.p[2], p.p[3], p.p[6], p.p[7], yq, X0, &ilo, &ihi);
        if (ilo <= ihi)
        {
            // T2 edges: P1- -> P0+ is the reverse of e12.
            const EdgeTI<I> e13 = make_edge_t<I>(p.p[2], p.p[3], p.p[6], p.p[7], X0, Y0);
            const EdgeTI<I> e32 = make_edge_t<I>(p.p[6], p.p[7], p.p[4], p.p[5], X0, Y0);
            const int32_t t21 = inside_threshold_t<I>(EdgeTI<I>{-e12.A, -e12.B, -e12.D0}, 1);
            const int32_t t13 = inside_threshold_t<I>(e13, 1), t32 = inside_threshold_t<I>(e32, 1);
            const float inv = __builtin_amdgcn_rcpf((float)A2);
            const int32_t i2 = 2 * ilo + 1;
            int32_t k12 = e12.A * i2 + e12.B * j2, k13 = e13.A * i2 + e13.B * j2, k32 = e32.A * i2 + e32.B * j2;
            const float l0 = ((float)e13.D0 + 128.f * (float)k13) * inv;  // P1-
            const float l1 = ((float)e32.D0 + 128.f * (float)k32) * inv;  // P0+
            const float l2 = -((float)e12.D0 + 128.f * (float)k12) * inv; // P1+
            const float s0 = 256.f * (float)e13.A * inv, s1 = 256.f * (float)e32.A * inv,
                        s2 = -256.f * (float)e12.A * inv;
            float cx = l0 * p.cov[4] + l1 * p.cov[2] + l2 * p.cov[6];
            float cy = l0 * p.cov[5] + l1 * p.cov[3] + l2 * p.cov[7];
            const float dcx = s0 * p.cov[4] + s1 * p.cov[2] + s2 * p.cov[6];
            const float dcy = s0 * p.cov[5] + s1 * p.cov[3] + s2 * p.cov[7];
            for (int i = ilo; i <= ihi; ++i)
            {
                if (-k12 >= t21 && k13 >= t13 && k32 >= t32)
                {
                    const float v = stroke_coverage<FEATHER>(cx, cy, table);
                    if (v > 0.f)
                        atomicMax(&sMax[j * kTile + i], __float_as_uint(v));
                }
                k12 += 2 * e12.A;
                k13 += 2 * e13.A;
                k32 += 2 * e32.A;
                cx += dcx;
                cy += dcy;
            }
        }
    }
}

// A feathered stroke's sample row, out of line like feather_prim_row.
__device__ __noinline__ void feather_stroke_prim_row(const Prim* __restrict__ prims, uint32_t idx, int j,
                                                     int32_t X0, int32_t Y0, uint32_t* __restrict__ sMax,
                                                     const float* __restrict__ table)
{
    const Prim p = prims[idx];
    stroke_prim_row<double, true>(p, j, TileGeom{X0, Y0}, sMax, table);
}

// Coverage of a winding count under the path's fill rule (draw_path.frag
// fill-rule resolution): clockwise clamps the signed count to [0, 1], nonzero
// takes min(|count|, 1), even-odd the distance to the nearest even integer
// (1 - |2 fract(|count| / 2) - 1|, the same value for every count this
// raster forms: integer windings plus multiples of 2^-16 of border coverage,
// for which both forms are exact).
__device__ __forceinline__ float fill_rule_coverage(float count, uint32_t flags)
{
    if (flags & 4u) // clockwise
        return fminf(fmaxf(count, 0.f), 1.f);
    if (flags & 2u) // even-odd
        return fminf(fabsf(fmaf(-2.f, rintf(count * .5f), count)), 1.f);
    return fminf(fabsf(count), 1.f);
}

// Image sampling with the reference's sampler state (pipeline_manager_vulkan.cpp:
// 11-46, 95-110): min/mag filter nearest or linear, mip mode nearest (level
// ceil(lod + 0.5) - 1, clamped), wrap clamp-to-edge / repeat / mirrored
// repeat per axis.  Texels are premultiplied RGBA8.
__device__ __forceinline__ int wrap_coord(int i, int n, uint32_t mode)
{
    if (mode == 1u) // repeat
    {
        const int m = i % n;
        return m < 0 ? m + n : m;
    }
    if (mode == 2u) // mirrored repeat
    {
        int m = i % (2 * n);
        if (m < 0)
            m += 2 * n;
        return m < n ? m : 2 * n - 1 - m;
    }
    return min(max(i, 0), n - 1);
}
__device__ void sample_image(const ImageDesc& img, float u, float v, float lod, uint32_t sampler, float* out)
{
    const uint32_t wrapX = sampler % 3u, wrapY = (sampler / 3u) % 3u, filter = sampler / 9u;
    int level = 0;
    if (lod > 0.f)
        level = min((int)ceilf(lod + .5f) - 1, (int)img.levels - 1);
    level = max(level, 0);
    const int w = max((int)img.width >> level, 1), h = max((int)img.height >> level, 1);
    const uint32_t* tex = img.texels + img.levelOffset[level];
    if (filter == 1u) // nearest
    {
        const int i = wrap_coord((int)floorf(u * (float)w), w, wrapX);
        const int j = wrap_coord((int)floorf(v * (float)h), h, wrapY);
        const uint32_t t = tex[(size_t)j * w + i];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            out[k] = float((t >> (8 * k)) & 0xffu) * (1.f / 255.f);
        return;
    }
    const float x = u * (float)w - .5f, y = v * (float)h - .5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int i0 = wrap_coord((int)fx0, w, wrapX), i1 = wrap_coord((int)fx0 + 1, w, wrapX);
    const int j0 = wrap_coord((int)fy0, h, wrapY), j1 = wrap_coord((int)fy0 + 1, h, wrapY);
    const uint32_t t00 = tex[(size_t)j0 * w + i0], t10 = tex[(size_t)j0 * w + i1];
    const uint32_t t01 = tex[(size_t)j1 * w + i0], t11 = tex[(size_t)j1 * w + i1];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const float c00 = float((t00 >> (8 * k)) & 0xffu) * (1.f / 255.f);
        const float c10 = float((t10 >> (8 * k)) & 0xffu) * (1.f / 255.f);
        const float c01 = float((t01 >> (8 * k)) & 0xffu) * (1.f / 255.f);
        const float c11 = float((t11 >> (8 * k)) & 0xffu) * (1.f / 255.f);
        const float top = (c10 - c00) * ax + c00;
        const float bot = (c11 - c01) * ax + c01;
        out[k] = (bot - top) * ay + top;
    }
}

// One texel of mip level L+1 from level L: the rounded mean of its 2x2
// footprint, clamped at odd edges.
__global__ void mip_kernel(uint32_t* __restrict__ texels, uint32_t srcOffset, int sw, int sh, uint32_t dstOffset,
                           int dw, int dh)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= dw || y >= dh)
        return;
    const uint32_t* src = texels + srcOffset;
    const int x0 = min(2 * x, sw - 1), x1 = min(2 * x + 1, sw - 1);
    const int y0 = min(2 * y, sh - 1), y1 = min(2 * y + 1, sh - 1);
    const uint32_t a = src[(size_t)y0 * sw + x0], b = src[(size_t)y0 * sw + x1];
    const uint32_t c = src[(size_t)y1 * sw + x0], d = src[(size_t)y1 * sw + x1];
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const uint32_t sum = ((a >> (8 * k)) & 0xffu) + ((b >> (8 * k)) & 0xffu) + ((c >> (8 * k)) & 0xffu) +
                             ((d >> (8 * k)) & 0xffu);
        out |= ((sum + 2u) >> 2) << (8 * k);
    }
    texels[dstOffset + (size_t)y * dw + x] = out;
}

// Advanced blend coefficients from unpremultiplied src and premultiplied dst
// (advanced_blend.glsl:87-280; unmultiply_rgb, common.glsl:210-216), fp32.
constexpr float kEpsFp16NonDenorm = 6.2e-5f;
__device__ __forceinline__ float min3f(const float* c) { return fminf(c[0], fminf(c[1], c[2])); }
__device__ __forceinline__ float max3f(const float* c) { return fmaxf(c[0], fmaxf(c[1], c[2])); }
__device__ __forceinline__ float lum3(const float* c) { return c[0] * .30f + c[1] * .59f + c[2] * .11f; }
__device__ __forceinline__ float sign0(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ void set_lum_d(const float* base, const float* lumColor, float* out)
{
    const float lumTarget = lum3(lumColor);
    const float l = lum3(base);
    float biased[3] = {base[0] - l, base[1] - l, base[2] - l};
    const float sx = lumTarget / fmaxf(kEpsFp16NonDenorm, -min3f(biased));
    const float sy = (1.f - lumTarget) / fmaxf(kEpsFp16NonDenorm, max3f(biased));
    const float satScale = fminf(1.f, fminf(sx, sy));
#pragma unroll
    for (int i = 0; i < 3; ++i)
        out[i] = biased[i] * satScale + lumTarget;
}
__device__ void set_lum_sat_d(const float* hue, const float* sat, const float* lumColor, float* out)
{
    const float satTarget = max3f(sat) - min3f(sat);
    const float mn = min3f(hue);
    float h[3] = {hue[0] - mn, hue[1] - mn, hue[2] - mn};
    const float scale = satTarget / fmaxf(kEpsFp16NonDenorm, max3f(h));
#pragma unroll
    for (int i = 0; i < 3; ++i)
        h[i] *= scale;
    set_lum_d(h, lumColor, out);
}
__device__ void advanced_blend_coeffs_d(const float* srcIn, const float* dp, uint32_t mode, float* out)
{
    float src[3] = {srcIn[0], srcIn[1], srcIn[2]};
    const float inva = dp[3] != 0.f ? 1.f / dp[3] : 0.f;
    const float dst[3] = {dp[0] * inva, dp[1] * inva, dp[2] * inva};
    switch (mode)
    {
        case 11u: // multiply
            for (int i = 0; i < 3; ++i)
                out[i] = src[i] * dst[i];
            break;
        case 1u: // screen
            for (int i = 0; i < 3; ++i)
                out[i] = src[i] + dst[i] - src[i] * dst[i];
            break;
        case 2u: // overlay
            for (int i = 0; i < 3; ++i)
            {
                const float sd = src[i] * dst[i];
                out[i] = 2.f * (dst[i] > .5f ? src[i] + dst[i] - sd - .5f : sd);
            }
            break;
        case 3u: // darken
            for (int i = 0; i < 3; ++i)
                out[i] = fminf(src[i], dst[i]);
            break;
        case 4u: // lighten
            for (int i = 0; i < 3; ++i)
                out[i] = fmaxf(src[i], dst[i]);
            break;
        case 5u: // color dodge
            for (int i = 0; i < 3; ++i)
            {
                const float d = fminf(fmaxf(dp[i], 0.f), dp[3]);
                const float denom = fminf(fmaxf(1.f - src[i], 0.f), 1.f) * dp[3];
                out[i] = denom == 0.f ? sign0(d) : fminf(1.f, d / denom);
            }
            break;
        case 6u: // color burn
            for (int i = 0; i < 3; ++i)
            {
                const float s = fminf(fmaxf(src[i], 0.f), 1.f);
                const float d = fminf(fmaxf(dp[i], 0.f), dp[3]);
                const float aa = dp[3] == 0.f ? 1.f : dp[3];
                const float numer = aa - d;
                out[i] = 1.f - (s == 0.f ? sign0(numer) : fminf(1.f, numer / (s * aa)));
            }
            break;
        case 7u: // hard light
            for (int i = 0; i < 3; ++i)
            {
                const float sd = src[i] * dst[i];
                out[i] = 2.f * (src[i] > .5f ? src[i] + dst[i] - sd - .5f : sd);
            }
            break;
        case 8u: // soft light
            for (int i = 0; i < 3; ++i)
            {
                float c;
                if (src[i] <= .5f)
                    c = 1.f - dst[i];
                else if (dst[i] <= .25f)
                    c = (16.f * dst[i] - 12.f) * dst[i] + 3.f;
                else
                    c = 1.f / sqrtf(dst[i]) - 1.f;
                out[i] = dst[i] + dst[i] * (2.f * src[i] - 1.f) * c;
            }
            break;
        case 9u: // difference
            for (int i = 0; i < 3; ++i)
                out[i] = fabsf(dst[i] - src[i]);
            break;
        case 10u: // exclusion
            for (int i = 0; i < 3; ++i)
                out[i] = src[i] + dst[i] - 2.f * src[i] * dst[i];
            break;
        case 12u: // hue
        case 13u: // saturation
        case 14u: // color
        case 15u: // luminosity
            for (int i = 0; i < 3; ++i)
                src[i] = fminf(fmaxf(src[i], 0.f), 1.f);
            if (mode == 12u)
                set_lum_sat_d(src, dst, dst, out);
            else if (mode == 13u)
                set_lum_sat_d(dst, src, dst, out);
            else if (mode == 14u)
                set_lum_d(src, dst, out);
            else
                set_lum_d(dst, src, out);
            break;
        default:
            for (int i = 0; i < 3; ++i)
                out[i] = src[i];
            break;
    }
}

// Four 0..255 channel values to RGBA8: round to nearest even, saturating.
__device__ __forceinline__ uint32_t pack_unorm8x4(const float* v)
{
    uint32_t p = __builtin_amdgcn_cvt_pk_u8_f32(v[0], 0u, 0u);
    p = __builtin_amdgcn_cvt_pk_u8_f32(v[1], 1u, p);
    p = __builtin_amdgcn_cvt_pk_u8_f32(v[2], 2u, p);
    return __builtin_amdgcn_cvt_pk_u8_f32(v[3], 3u, p);
}

// Gradient paint color, unpremultiplied, 0..255 scale (draw_path.vert:285-331,450-476).
__device__ __forceinline__ void sample_gradient255(const PathRaster& pr, int px, int py,
                                                   const uint32_t* __restrict__ ramps, uint32_t gradHeight,
                                                   float* src)
{
    const float fx = float(px) + .5f, fy = float(py) + .5f;
    const float pcx = pr.paintMatrix[0] * fx + pr.paintMatrix[2] * fy + pr.paintMatrix[4];
    const float pcy = pr.paintMatrix[1] * fx + pr.paintMatrix[3] * fy + pr.paintMatrix[5];
    float t = pr.paintType == LINEAR_GRADIENT_PAINT_TYPE ? pcx : sqrtf(pcx * pcx + pcy * pcy);
    t = fminf(fmaxf(t, 0.f), 1.f);
    const float u = pr.gradSpan > 1.f ? (1.f - 1.f / 512.f) * t + (.5f / 512.f) : (1.f / 512.f) * t + pr.gradSpan;
    const float tc = u * 512.f - .5f;
    const int i0 = (int)floorf(tc);
    const float fr = tc - (float)i0;
    const int ia = min(max(i0, 0), 511), ib = min(max(i0 + 1, 0), 511);
    const uint32_t* rowp = ramps + (size_t)min(pr.gradRow, gradHeight - 1) * kGradWidth;
    const uint32_t ta = rowp[ia], tb = rowp[ib];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const float va = float((ta >> (8 * k)) & 0xffu);
        const float vb = float((tb >> (8 * k)) & 0xffu);
        src[k] = (vb - va) * fr + va;
    }
}

// Advanced blend modes and image paints (draw_path.vert:431-545,
// draw_raster_order_path.frag:191-204), in 0..1 scale.
__device__ __noinline__ uint32_t resolve_general(const PathRaster* __restrict__ prp, float coverage, int px,
                                                 int py, float noise255, const uint32_t* __restrict__ ramps,
                                                 uint32_t gradHeight, const ImageDesc* __restrict__ images,
                                                 uint32_t colPacked)
{
    const PathRaster pr = *prp;
    float col[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        col[k] = float((colPacked >> (8 * k)) & 0xffu) * (1.f / 255.f);
    float src[4];
    const uint32_t blendMode = (pr.flags >> 8) & 0xfu;
    if (pr.paintType == SOLID_COLOR_PAINT_TYPE && blendMode == 0u)
    {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            src[k] = pr.color255[k] * (1.f / 255.f) * coverage;
    }
    else if (pr.paintType == SOLID_COLOR_PAINT_TYPE)
    {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            src[k] = pr.paintMatrix[k];
        src[3] *= coverage;
    }
    else
    {
        sample_gradient255(pr, px, py, ramps, gradHeight, src);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            src[k] *= 1.f / 255.f;
        src[3] *= coverage;
        if (blendMode == 0u)
        {
            src[0] *= src[3];
            src[1] *= src[3];
            src[2] *= src[3];
        }
    }
    if (pr.image != 0u)
    {
        // find_paint_color (draw_path.vert:479-493): modulate by the image
        // sample; with advanced blend the sample is unmultiplied first.
        const float fx = float(px) + .5f, fy = float(py) + .5f;
        const float u = pr.imageMatrix[0] * fx + pr.imageMatrix[2] * fy + pr.imageMatrix[4];
        const float v = pr.imageMatrix[1] * fx + pr.imageMatrix[3] * fy + pr.imageMatrix[5];
        float ic[4];
        sample_image(images[pr.image - 1u], u, v, pr.imageLod, pr.imageSampler, ic);
        if (blendMode != 0u)
        {
            const float inv = ic[3] != 0.f ? 1.f / ic[3] : 0.f;
            ic[0] *= inv;
            ic[1] *= inv;
            ic[2] *= inv;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            src[k] *= ic[k];
    }
    if (blendMode != 0u)
    {
        // mix the unpremultiplied paint with the blend coefficients by the
        // destination alpha, then premultiply
        float coeffs[3];
        advanced_blend_coeffs_d(src, col, blendMode, coeffs);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            src[k] = ((coeffs[k] - src[k]) * col[3] + src[k]) * src[3];
    }
    const float paintAlpha = src[3];
    float out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        out[k] = (src[k] + col[k] * (1.f - paintAlpha)) * 255.f;
    if (paintAlpha != 0.f)
    {
        out[0] += noise255;
        out[1] += noise255;
        out[2] += noise255;
    }
    return pack_unorm8x4(out);
}

// Clip, paint and src-over of one path into one pixel (draw_path.vert:431-545,
// draw_raster_order_path.frag:84-238), in 0..255 scale.  A zero coverage
// leaves the pixel unchanged, so callers skip it (except for clip updates).
__device__ __forceinline__ float clip_coverage(uint32_t clip)
{
    return __half2float(__ushort_as_half((unsigned short)(clip >> 16)));
}

// Clip rect coverage at a pixel center (common.glsl:376-400,
// draw_raster_order_path.frag:152-159): the least of the Manhattan distances,
// in AA widths, to the rect's four edges plus one half, rounded to half
// precision.  A singular inverse matrix gives its translation as the coverage.
__device__ __forceinline__ float clip_rect_coverage(const float* m, int px, int py)
{
    const float awx = fabsf(m[0]) + fabsf(m[2]), awy = fabsf(m[1]) + fabsf(m[3]);
    float d;
    if (awx != 0.f && awy != 0.f)
    {
        const float rx = 1.f / awx, ry = 1.f / awy;
        const float fx = float(px) + .5f, fy = float(py) + .5f;
        const float cx = m[0] * fx + m[2] * fy + m[4], cy = m[1] * fx + m[3] * fy + m[5];
        d = fminf(fminf(cx * rx + rx + .5f, cy * ry + ry + .5f), fminf(-cx * rx + rx + .5f, -cy * ry + ry + .5f));
    }
    else
        d = fminf(m[4], m[5]);
    return __half2float(__float2half_rn(d));
}

// The fields of a path's PathRaster that every resolve reads, 32 bytes (one
// scalar load); the rest is read from PathRaster where it is used.
struct PathPaint
{
    uint32_t kind;  // PathRaster flags | paintType << 16 | kPaintImage | kPaintPlain | kPaintNoOp
    uint32_t clips; // clipID | outerClipID << 16
    float color255[4];
    float alpha;
    float oneMinusAlpha; // 1 - alpha, the src-over factor at full coverage
};
static_assert(sizeof(PathPaint) == 32, "PathPaint layout");
constexpr uint32_t kPaintImage = 1u << 20;
// A solid src-over paint with no clip, clip rect or advanced blend mode: at
// full coverage it blends as dst * (1 - alpha) + color255 (kPaintPlain), or
// leaves the pixel as it is when alpha is 0 (kPaintNoOp).
constexpr uint32_t kPaintPlain = 1u << 21;
constexpr uint32_t kPaintNoOp = 1u << 22;

__device__ __forceinline__ void resolve_pixel(const PathPaint& pp, const PathRaster* __restrict__ prp,
                                              float coverage, int px, int py, float noise255,
                                              const uint32_t* __restrict__ ramps, uint32_t gradHeight,
                                              const ImageDesc* __restrict__ images, uint32_t& colPacked,
                                              uint32_t& clip)
{
    const uint32_t clipID = clip & 0xffffu;
    const uint32_t flags = pp.kind & 0xffffu;
    const uint32_t drawClip = pp.clips & 0xffffu, outerClip = pp.clips >> 16;
    if (flags & 8u) // clip update
    {
        if (outerClip != 0u)
            coverage = fminf(coverage, clipID == outerClip ? clip_coverage(clip) : 0.f);
        clip = ((uint32_t)__half_as_ushort(__float2half_rn(coverage)) << 16) | drawClip;
        return;
    }
    if (drawClip != 0u)
        coverage = clipID == drawClip ? fminf(clip_coverage(clip), coverage) : 0.f;
    if (flags & 16u)
        coverage = fminf(fmaxf(clip_rect_coverage(prp->clipRect, px, py), 0.f), coverage);
    if ((flags & 0xf00u) != 0u || (pp.kind & kPaintImage))
    {
        colPacked = resolve_general(prp, coverage, px, py, noise255, ramps, gradHeight, images, colPacked);
        return;
    }
    float src[4], a;
    if ((pp.kind >> 16 & 0xfu) == SOLID_COLOR_PAINT_TYPE)
    {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            src[k] = pp.color255[k] * coverage;
        a = pp.alpha * coverage;
    }
    else
    {
        sample_gradient255(*prp, px, py, ramps, gradHeight, src);
        src[3] *= coverage;
        a = src[3] * (1.f / 255.f);
        src[0] *= a;
        src[1] *= a;
        src[2] *= a;
    }
    // src-over, dst * (1 - a) + src, fused (as the reference's shaders are
    // compiled)
    const float ia = 1.f - a;
    const uint32_t c = colPacked;
    float out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        out[k] = fmaf(float((c >> (8 * k)) & 0xffu), ia, src[k]);
    // dither where the paint lands (adding +0 elsewhere changes nothing)
    const float n = a != 0.f ? noise255 : 0.f;
    out[0] += n;
    out[1] += n;
    out[2] += n;
    colPacked = pack_unorm8x4(out);
}

// resolve_pixel of a kPaintPlain paint at coverage 1: the same operations on
// the same values (src = color255 * 1, a = alpha * 1 != 0, so the dither
// applies).
__device__ __forceinline__ void resolve_plain_full(const PathPaint& pp, float noise255, uint32_t& colPacked)
{
    const uint32_t c = colPacked;
    // the factor in a vector register, so that each fused multiply-add takes
    // one scalar operand (the color) and needs no copy
    float ia = pp.oneMinusAlpha;
    asm volatile("" : "+v"(ia));
    float out[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        out[k] = fmaf(float((c >> (8 * k)) & 0xffu), ia, pp.color255[k]);
    out[0] += noise255;
    out[1] += noise255;
    out[2] += noise255;
    colPacked = pack_unorm8x4(out);
}

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Inclusive prefix sum within each DPP row of 16 lanes (row_shr:1, 2, 4, 8;
// lanes shifted in from outside the row read 0).  All 64 lanes must be active.
__device__ __forceinline__ int32_t row_scan16(int32_t x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);
    return x;
}

// Accumulator arrays a path with n primitives in the tile needs, packed with
// a slot count of one: slots | arrays << 16.
__device__ __forceinline__ uint32_t slot_arrays(uint32_t n)
{
    return n == 0 ? 0u : (1u | ((n > kPackedPrims ? 2u : 1u) << 16));
}

// ---- tile order: the raster's tiles heaviest first
//
// A tile's work is dominated by the paths it lists and the primitives binned
// to it, so its estimated cost is 4 units per listed path (a path with
// primitives here or a nonzero corner winding) plus one per primitive.  Tiles
// are grouped into kCostClasses classes, eight per octave of cost, and the
// order lists the classes from the heaviest down (the order within a class is
// whatever the atomics give: every tile is rendered independently, so the
// image does not depend on it).  Dispatching in this order starts the long
// tiles first, so they do not trail at the end of the launch.
constexpr int kCostClasses = 256;

__device__ __forceinline__ uint32_t cost_class(uint32_t cost)
{
    const uint32_t v = cost + 1u;
    const int e = 31 - __clz(v);
    const uint32_t f = e >= 3 ? (v >> (e - 3)) & 7u : (v << (3 - e)) & 7u;
    const uint32_t k = min((uint32_t)e * 8u + f, (uint32_t)kCostClasses - 1u);
    return (uint32_t)kCostClasses - 1u - k; // class 0 is the heaviest
}

// One wave per tile: its cost class into tileClass, counted in hist.
__global__ __launch_bounds__(256) void tile_cost_kernel(const PathCells* __restrict__ cells,
                                                        const uint2* __restrict__ cellPacked,
                                                        const CanvasDev* __restrict__ canvases, uint32_t canvasCount,
                                                        uint32_t tileTotal, uint32_t* __restrict__ tileClass,
                                                        uint32_t* __restrict__ hist)
{
    const uint32_t item = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (item >= tileTotal)
        return;
    uint32_t lo = 0, hi = canvasCount - 1;
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (canvases[mid].tileBase <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    const CanvasDev cv = canvases[lo];
    const int tile = (int)(item - cv.tileBase);
    const int tx = tile % cv.tilesX, ty = tile / cv.tilesX;
    uint32_t cost = 0;
    for (uint32_t p = cv.firstPath + (uint32_t)lane; p < cv.pathEnd; p += 64u)
    {
        const PathCells c = cells[p];
        if (tx >= c.tx0 && tx < c.tx1 && ty >= c.ty0 && ty < c.ty1)
        {
            const uint32_t cell = c.base + (uint32_t)(ty - c.ty0) * c.width + (uint32_t)(tx - c.tx0);
            const uint2 here = cellPacked[cell];
            const uint32_t n = cellPacked[cell + 1].x - here.x;
            cost += n != 0 ? 4u + n : (here.y != 0 ? 4u : 0u);
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1)
        cost += (uint32_t)__shfl_xor((int)cost, d);
    if (lane == 0)
    {
        const uint32_t cls = cost_class(cost);
        tileClass[item] = cls;
        atomicAdd(&hist[cls], 1u);
    }
}

// Places every tile in the order: its class's first position (the exclusive
// prefix of the histogram) plus its place among the tiles of that class.
// cursor[] starts at zero.
__global__ __launch_bounds__(256) void tile_order_kernel(const uint32_t* __restrict__ tileClass,
                                                         const uint32_t* __restrict__ hist,
                                                         uint32_t* __restrict__ cursor, uint32_t tileTotal,
                                                         uint32_t* __restrict__ order)
{
    static_assert(kCostClasses == 256, "one class per thread");
    __shared__ uint32_t sPre[kCostClasses];
    const int t = threadIdx.x;
    uint32_t v = hist[t];
    sPre[t] = v;
    __syncthreads();
    for (int d = 1; d < kCostClasses; d <<= 1)
    {
        const uint32_t a = t >= d ? sPre[t - d] : 0u;
        __syncthreads();
        sPre[t] += a;
        __syncthreads();
    }
    const uint32_t first = sPre[t] - v; // exclusive
    __syncthreads();
    sPre[t] = first;
    __syncthreads();
    const uint32_t item = blockIdx.x * 256u + (uint32_t)t;
    if (item < tileTotal)
    {
        const uint32_t cls = tileClass[item];
        order[sPre[cls] + atomicAdd(&cursor[cls], 1u)] = item;
    }
}

// FEATHER: the flush has feathered paths (their rows are out-of-line calls,
// whose call frames cost the other variant registers and scratch).
// PERSISTENT: a resident grid whose workgroups take tiles from a device
// queue (the counter at queueHead, zero at launch) until it runs dry; else one
// workgroup per tile.  Either way the k-th tile taken is tileOrder[k] (the
// flush's tiles heaviest first) or, without an order, tile k.  WAVES is the
// occupancy the kernel is compiled for (waves per SIMD).
template <bool FEATHER, bool PERSISTENT, int WAVES>
__global__ __launch_bounds__(kWgThreads, WAVES) void raster_kernel(
    const PathCells* __restrict__ cells, const PathRaster* __restrict__ raster,
    const PathPaint* __restrict__ pathPaint, const uint2* __restrict__ cellPacked, const uint32_t* __restrict__ binned, const Prim* __restrict__ prims,
    const uint32_t* __restrict__ ramps, uint32_t gradHeight, const ImageDesc* __restrict__ images,
    const CanvasDev* __restrict__ canvases, uint32_t canvasCount, uint32_t dither,
    const float* __restrict__ featherTable, const uint32_t* __restrict__ tileOrder, uint32_t tileTotal,
    uint32_t* __restrict__ queueHead)
{
    __shared__ __align__(16) int32_t sAcc[kAccArrays][kTile * kTile];
    __shared__ int32_t sLeft[kAccArrays][kTile];
    // The listed paths of the current window of 256 paths, in order, with
    // exclusive prefixes of their primitive counts and of slot_arrays().
    // sRel[q] = (path, primitive count n, corner winding C, slot/array
    // prefix) -- one 16-byte read per entry in the resolve; sRelB[q] = (bin
    // offset, primitive prefix).
    __shared__ uint4 sRel[kWindow];
    __shared__ uint2 sRelB[kWindow];
    __shared__ uint32_t sWave[3][kWgThreads / 64];
    __shared__ uint8_t sOwner[kMaxItems];
    __shared__ uint32_t sItemPrim[kWgThreads];
    __shared__ uint32_t sItemInfo[kWgThreads]; // array | wide << 4 | stroke << 5 | row0 << 8 | item0 << 16

    __shared__ uint32_t sTake;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    for (;;)
    {
        uint32_t take = blockIdx.x;
        if (PERSISTENT)
        {
            __syncthreads(); // the previous tile is done with sTake and the arrays
            if (t == 0)
                sTake = atomicAdd(queueHead, 1u);
            __syncthreads();
            take = sTake;
            if (take >= tileTotal)
                break; // the same value in every wave: the whole workgroup leaves
        }
        const uint32_t item = tileOrder != nullptr ? tileOrder[take] : take;
        // this tile's canvas: the last one whose first tile is at or before it
        uint32_t lo = 0, hi = canvasCount - 1;
        while (lo < hi)
        {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (canvases[mid].tileBase <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const CanvasDev cv = canvases[lo];
        const int width = cv.width, height = cv.height, tilesX = cv.tilesX;
        uint32_t* __restrict__ target = cv.pixels;
        const uint32_t loadPreserve = cv.loadPreserve, clearPremul = cv.clearPremul;
        const uint32_t pathBegin = cv.firstPath, pathEnd = cv.pathEnd;
        const int tile = (int)(item - cv.tileBase);
        const int tx = tile % tilesX, ty = tile / tilesX;
        const int px = tx * kTile + (t & (kTile - 1)), py = ty * kTile + (t >> kTileLog2);
        TileGeom g;
        g.X0 = tx << (kSubpixelLog2 + kTileLog2);
        g.Y0 = ty << (kSubpixelLog2 + kTileLog2);

        // The pixel's RGBA8 color (quantized after every path, as the reference's
        // color buffer is), and its clip buffer entry as clip ID (low 16 bits)
        // with fp16 coverage (high 16 bits), like the reference's clipBuffer.
        const bool inImage = px < width && py < height;
        uint32_t col = clearPremul;
        if (loadPreserve && inImage)
            col = target[(size_t)py * width + px];
        uint32_t clip = 0u;
        const float noise255 = dither ? dither_noise(px, py) * 255.f : 0.f;

        // clear accumulators for the first batch (the listing's barriers order
        // this before their use)
        static_assert(kAccArrays * kTile == kWgThreads, "one sLeft entry per thread");
        for (int q = t; q < kAccArrays * kTile * kTile / 4; q += kWgThreads)
            reinterpret_cast<int4*>(&sAcc[0][0])[q] = make_int4(0, 0, 0, 0);
        (&sLeft[0][0])[t] = 0;

        for (uint32_t base = pathBegin; base < pathEnd; base += kWindow)
        {
            // ---- list the window's paths that touch the tile
            const uint32_t p = base + (uint32_t)t;
            uint32_t off = 0, n = 0;
            int32_t C = 0;
            bool listed = false;
            if (t < kWindow && p < pathEnd)
            {
                const PathCells c = cells[p];
                if (tx >= c.tx0 && tx < c.tx1 && ty >= c.ty0 && ty < c.ty1)
                {
                    const uint32_t cell = c.base + (uint32_t)(ty - c.ty0) * c.width + (uint32_t)(tx - c.tx0);
                    const uint2 here = cellPacked[cell];
                    off = here.x;
                    n = cellPacked[cell + 1].x - off;
                    C = (int32_t)here.y;
                    // A path with no primitives here covers the tile with its corner
                    // winding's coverage; when that is zero it draws nothing, and a
                    // clip update's zero-coverage write reads the same as no write
                    // (every clip draw has a fresh ID), so it is not listed.
                    listed = n != 0;
                    if (n == 0 && C != 0)
                    {
                        const uint32_t flags = pathPaint[p].kind & 0xffffu;
                        listed = !(flags & 1u) && fill_rule_coverage((float)C, flags) != 0.f;
                    }
                }
            }
            const uint32_t vSA = slot_arrays(n);
            uint32_t iSA = vSA, iP = n;
    #pragma unroll
            for (int d = 1; d < 64; d <<= 1)
            {
                const uint32_t a0 = (uint32_t)__shfl_up((int)iSA, d), a1 = (uint32_t)__shfl_up((int)iP, d);
                if (lane >= d)
                {
                    iSA += a0;
                    iP += a1;
                }
            }
            const uint64_t bal = __ballot(listed);
            const uint32_t posInWave = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            if (lane == 63)
            {
                sWave[0][wave] = iSA;
                sWave[1][wave] = iP;
                sWave[2][wave] = (uint32_t)__popcll(bal);
            }
            __syncthreads();
            uint32_t preSA = 0, preP = 0, preL = 0, nRel = 0;
    #pragma unroll
            for (int w = 0; w < kWgThreads / 64; ++w)
            {
                if (w < wave)
                {
                    preSA += sWave[0][w];
                    preP += sWave[1][w];
                    preL += sWave[2][w];
                }
                nRel += sWave[2][w];
            }
            nRel = uniform_u32(nRel);
            if (listed)
            {
                const uint32_t q = preL + posInWave;
                sRel[q] = make_uint4(p, n, (uint32_t)C, preSA + iSA - vSA);
                sRelB[q] = make_uint2(off, preP + iP - n);
            }
            __syncthreads();

            for (uint32_t s = 0; s < nRel;)
            {
                // ---- the batch: entries s .. s+len-1, cut where a limit breaks
                const uint32_t baseP = uniform_u32(sRelB[s].y), baseSA = uniform_u32(sRel[s].w);
                uint32_t len;
                {
                    const uint32_t i = s + (uint32_t)lane;
                    bool stop = true;
                    if (i < nRel)
                    {
                        const uint32_t ni = sRel[i].y;
                        const uint32_t sa = sRel[i].w + slot_arrays(ni) - baseSA;
                        const uint32_t primsIncl = sRelB[i].y + ni - baseP;
                        stop = (sa >> 16) > (uint32_t)kAccArrays || (primsIncl > kBatchPrims && (sa & 0xffffu) > 1u);
                    }
                    const uint64_t m = __ballot(stop);
                    len = uniform_u32(m ? (uint32_t)__builtin_ctzll(m) : (uint32_t)kBatchEntries);
                }
                const uint32_t last = s + len - 1;
                const uint32_t nPrim = uniform_u32(sRelB[last].y + sRel[last].y - baseP);
                const uint32_t arraysUsed = uniform_u32((sRel[last].w + slot_arrays(sRel[last].y) - baseSA) >> 16);

                if (arraysUsed != 0)
                {
                    // (the arrays and sLeft are clear: set at the start and
                    // cleared again after the previous batch's resolve)
                    for (uint32_t c0 = 0; c0 < nPrim; c0 += kBatchPrims)
                    {
                        // ---- one primitive per thread
                        const uint32_t f = c0 + (uint32_t)t;
                        uint32_t rowsF = 0, rowsS = 0, info = 0, primIdx = 0;
                        if (f < nPrim)
                        {
                            // owning entry: the last one whose first primitive is at or
                            // before f (a batch has at most 64 entries: six fixed steps,
                            // selects rather than branches)
                            uint32_t lo = s;
    #pragma unroll
                            for (uint32_t step = kBatchEntries / 2; step > 0; step >>= 1)
                            {
                                const uint32_t cand = lo + step;
                                const uint32_t probe = sRelB[min(cand, last)].y - baseP;
                                lo = (cand <= last && probe <= f) ? cand : lo;
                            }
                            const uint4 r = sRel[lo];
                            const uint2 rb = sRelB[lo];
                            const uint32_t a = (r.w - baseSA) >> 16;
                            const bool wide = r.y > kPackedPrims;
                            const bool stroke = pathPaint[r.x].kind & 1u;
                            primIdx = binned[rb.x + (f - (rb.y - baseP))];
                            const Prim pm = prims[primIdx];
                            int r0, r1;
                            prim_rows(pm, g, &r0, &r1);
                            const uint32_t rows = (uint32_t)max(r1 - r0 + 1, 0);
                            if (stroke)
                                rowsS = rows;
                            else
                            {
                                rowsF = rows;
                                fill_prim_left(pm, g, sLeft[a]);
                            }
                            info = a | (wide ? 16u : 0u) | (stroke ? 32u : 0u) | ((uint32_t)min(r0, kTile - 1) << 8);
                        }
                        // exclusive prefix of the row counts, fills then strokes
                        const uint32_t v = rowsF | (rowsS << 16);
                        uint32_t inc = v;
    #pragma unroll
                        for (int d = 1; d < 64; d <<= 1)
                        {
                            const uint32_t a0 = (uint32_t)__shfl_up((int)inc, d);
                            if (lane >= d)
                                inc += a0;
                        }
                        if (lane == 63)
                            sWave[0][wave] = inc;
                        __syncthreads();
                        uint32_t pre = 0, tot = 0;
    #pragma unroll
                        for (int w = 0; w < kWgThreads / 64; ++w)
                        {
                            const uint32_t x = sWave[0][w];
                            if (w < wave)
                                pre += x;
                            tot += x;
                        }
                        const uint32_t ex = pre + inc - v;
                        const uint32_t totF = tot & 0xffffu, total = uniform_u32(totF + (tot >> 16));
                        const uint32_t item0 = rowsS != 0 ? totF + (ex >> 16) : (ex & 0xffffu);
                        for (uint32_t k = 0; k < rowsF + rowsS; ++k)
                            sOwner[item0 + k] = (uint8_t)t;
                        sItemPrim[t] = primIdx;
                        sItemInfo[t] = info | (item0 << 16);
                        __syncthreads();

                        // ---- (primitive, row) items
                        for (uint32_t it = (uint32_t)t; it < total; it += kWgThreads)
                        {
                            const uint32_t owner = sOwner[it];
                            const uint32_t inf = sItemInfo[owner];
                            const int a = (int)(inf & 15u);
                            const int row = (int)((inf >> 8) & 15u) + (int)(it - (inf >> 16));
                            const uint32_t pIdx = sItemPrim[owner];
                            const Prim pm = prims[pIdx];
                            if (inf & 32u)
                            {
                                if (FEATHER && (pm.flags & kPrimFeatherStroke))
                                    feather_stroke_prim_row(prims, pIdx, row, g.X0, g.Y0,
                                                            reinterpret_cast<uint32_t*>(sAcc[a]), featherTable);
                                else
                                    stroke_prim_row<double>(pm, row, g, reinterpret_cast<uint32_t*>(sAcc[a]));
                            }
                            else if (FEATHER && (pm.flags & kPrimFeather))
                                feather_prim_row(prims, pIdx, row, g.X0, g.Y0, sAcc[a], sAcc[a + ((inf >> 4) & 1u)],
                                                 (inf & 16u) ? 1 : (1 << 24), featherTable);
                            else if (inf & 16u)
                                fill_prim_row<double>(pm, row, g, sAcc[a], sAcc[a + 1], 1);
                            else
                                fill_prim_row<double>(pm, row, g, sAcc[a], sAcc[a], 1 << 24);
                        }
                        __syncthreads();
                    }
                    // left-edge deltas -> prefix over rows, in place: thread t
                    // holds array t / 16, row t % 16, and a DPP row is one array
                    (&sLeft[0][0])[t] = row_scan16((&sLeft[0][0])[t]);
                    __syncthreads();
                }

                // ---- resolve the batch in submission order.  Every thread takes
                // the same entry, so a path's coverage comes from its accumulators
                // here: the winding at a pixel is the corner winding, plus the
                // left-edge prefix at its row, plus the crossings of its row up to
                // its column -- a 16-lane scan, a DPP row being a row of the tile.
                // Each pixel's cells are cleared as they are read (a two-array
                // path's border array too).
                for (uint32_t e = 0; e < len; ++e)
                {
                    const uint32_t i = s + e;
                    const uint4 rel = sRel[i];
                    const uint32_t path = uniform_u32(rel.x);
                    const uint32_t ni = uniform_u32(rel.y);
                    const int32_t C = (int32_t)uniform_u32(rel.z);
                    const PathPaint pr = pathPaint[path];
                    const uint32_t prFlags = pr.kind & 0xffffu;
                    if (ni == 0)
                    {
                        // No primitives here: the listing keeps such a path only
                        // where its corner winding gives coverage, and an integer
                        // winding's coverage is 1 under every fill rule.
                        if (pr.kind & kPaintPlain)
                            resolve_plain_full(pr, noise255, col);
                        else if (!(pr.kind & kPaintNoOp))
                            resolve_pixel(pr, raster + path, 1.f, px, py, noise255, ramps, gradHeight, images, col,
                                          clip);
                        continue;
                    }
                    float cov;
                    {
                        const uint32_t a = uniform_u32((rel.w - baseSA) >> 16);
                        const int32_t v = sAcc[a][t];
                        sAcc[a][t] = 0;
                        if (prFlags & 1u)
                            cov = fminf(__int_as_float(v), 1.f);
                        else
                        {
                            int32_t d, b;
                            if (ni > kPackedPrims)
                            {
                                d = v;
                                b = sAcc[a + 1][t];
                                sAcc[a + 1][t] = 0;
                            }
                            else
                            {
                                // |border| < 2^23: it is the low 24 bits, sign-extended
                                d = (v + (1 << 23)) >> 24;
                                b = (int32_t)((uint32_t)v << 8) >> 8;
                            }
                            const int32_t left = sLeft[a][t >> kTileLog2];
                            const int32_t w = C + left + row_scan16(d);
                            cov = fill_rule_coverage(fmaf((float)b, 1.f / 65536.f, (float)w), prFlags); // exact scaling
                        }
                    }
                    if ((cov != 0.f && !(pr.kind & kPaintNoOp)) || (prFlags & 8u))
                        resolve_pixel(pr, raster + path, cov, px, py, noise255, ramps, gradHeight, images, col, clip);
                }
                __syncthreads(); // every wave has read sLeft
                (&sLeft[0][0])[t] = 0;
                __syncthreads(); // the next batch accumulates into the arrays and sLeft
                s += len;
            }
        }
        if (inImage)
            target[(size_t)py * width + px] = col;
        if (!PERSISTENT)
            break;
    }
}

// Unpacks PathData/PaintData/PaintAux into PathRaster (gpu.cpp:859-1063,
// draw_path.vert:431-545).
__global__ void path_raster_kernel(const PathDataV* __restrict__ paths, const uint2* __restrict__ paints,
                                   const float* __restrict__ aux, const uint32_t* __restrict__ pathImage,
                                   uint32_t pathCount, uint32_t gradHeight, PathRaster* __restrict__ out,
                                   PathPaint* __restrict__ outPaint)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > pathCount)
        return;
    PathRaster r;
    const PathDataV pd = paths[i];
    const uint2 paint = paints[i];
    const float* ax = aux + (size_t)i * 32;
    uint32_t paintType = paint.x & 0xfu;
    uint32_t flags = 0;
    if (pd.strokeRadius != 0.f)
        flags |= 1u;
    else if (paint.x & PAINT_FLAG_EVEN_ODD_FILL)
        flags |= 2u;
    else if (!(paint.x & PAINT_FLAG_NON_ZERO_FILL))
        flags |= 4u;
    r.paintType = paintType;
    r.outerClipID = 0;
    if (paintType == CLIP_UPDATE_PAINT_TYPE)
    {
        flags |= 8u;
        r.clipID = paint.y >> 16;
        r.outerClipID = paint.x >> 16;
    }
    else
    {
        r.clipID = paint.x >> 16;
    }
    if (paintType != CLIP_UPDATE_PAINT_TYPE)
    {
        flags |= ((paint.x >> 4) & 0xfu) << 8; // shader blend mode
        if (paint.x & PAINT_FLAG_HAS_CLIP_RECT)
            flags |= 16u;
    }
    r.flags = flags;
    const float a = float(paint.y >> 24) * (1.f / 255.f);
    r.color255[0] = float(paint.y & 0xffu) * a;
    r.color255[1] = float((paint.y >> 8) & 0xffu) * a;
    r.color255[2] = float((paint.y >> 16) & 0xffu) * a;
    r.color255[3] = float(paint.y >> 24);
    r.alpha = a;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        r.paintMatrix[k] = ax[k];
    if (paintType == SOLID_COLOR_PAINT_TYPE)
    {
        // Solid paints carry no paint matrix; keep the unpremultiplied color
        // there for advanced blend modes.
#pragma unroll
        for (int k = 0; k < 4; ++k)
            r.paintMatrix[k] = float((paint.y >> (8 * k)) & 0xffu) * (1.f / 255.f);
    }
    float span0 = ax[6], span1 = ax[7];
    r.gradSpan = span0 > .9f ? 2.f : span1;
    r.gradRow = (uint32_t)(u2f(paint.y) * (float)gradHeight);
    // Image paints (gpu.cpp:1006-1034): image matrix and a constant LOD.
    const uint32_t img = pathImage[i];
    r.image = img & 0xffffu;
    r.imageSampler = img >> 16;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        r.imageMatrix[k] = ax[16 + k];
    r.imageLod = ax[22];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        r.clipRect[k] = ax[8 + k];
    r.pad = 0;
    out[i] = r;
    PathPaint pp;
    pp.kind = r.flags | (r.paintType << 16) | (r.image != 0u ? kPaintImage : 0u);
    pp.clips = (r.clipID & 0xffffu) | (r.outerClipID << 16);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        pp.color255[k] = r.color255[k];
    pp.alpha = r.alpha;
    pp.oneMinusAlpha = 1.f - r.alpha;
    const bool plain = (flags & (8u | 16u | 0xf00u)) == 0u && (r.clipID & 0xffffu) == 0u &&
                       paintType == SOLID_COLOR_PAINT_TYPE && r.image == 0u;
    if (plain)
        pp.kind |= r.alpha != 0.f ? kPaintPlain : kPaintNoOp;
    outPaint[i] = pp;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of u32 (three phases, 4096 items per block).
// ---------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr int kScanBlock = 256;
constexpr int kScanTile = kScanItems * kScanBlock;
// A scan total that does not fit 32 bits reads as this value.
constexpr uint32_t kScanSaturated = 0xffffffffu;

__global__ void scan_reduce_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                   uint32_t* __restrict__ partials)
{
    // Sums in 64 bits; a block sum beyond 2^32-1 saturates, so the total
    // saturates too and the host rejects the flush (kScanSaturated).
    __shared__ uint64_t red[kScanBlock];
    uint32_t base = blockIdx.x * kScanTile;
    uint64_t s = 0;
    for (int k = 0; k < kScanItems; ++k)
    {
        uint32_t i = base + k * kScanBlock + threadIdx.x;
        if (i < n)
            s += in[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kScanBlock / 2; off > 0; off >>= 1)
    {
        if (threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[blockIdx.x] = (uint32_t)min(red[0], (uint64_t)kScanSaturated);
}

__global__ void scan_partials_kernel(uint32_t* __restrict__ partials, uint32_t n,
                                     uint32_t* __restrict__ total)
{
    __shared__ uint32_t buf[1024];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += blockDim.x)
    {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < n ? partials[i] : 0;
        if (v == kScanSaturated)
            carry = kScanSaturated; // benign race: every writer stores the same value
        buf[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < blockDim.x; off <<= 1)
        {
            uint32_t t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n)
            partials[i] = (uint32_t)carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1)
            carry += buf[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *total = (uint32_t)min(carry, (uint64_t)kScanSaturated);
}

__global__ void scan_apply_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                  const uint32_t* __restrict__ partials, const uint32_t* __restrict__ total,
                                  uint32_t* __restrict__ out)
{
    __shared__ uint32_t buf[kScanTile];
    __shared__ uint32_t wsum[kScanBlock];
    uint32_t base = blockIdx.x * kScanTile;
    // Each thread owns kScanItems consecutive items.
    uint32_t local[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k)
    {
        uint32_t i = base + threadIdx.x * kScanItems + k;
        uint32_t v = i < n ? in[i] : 0;
        local[k] = s;
        s += v;
    }
    wsum[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < kScanBlock; off <<= 1)
    {
        uint32_t t = threadIdx.x >= off ? wsum[threadIdx.x - off] : 0;
        __syncthreads();
        wsum[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t prefix = partials[blockIdx.x] + wsum[threadIdx.x] - s;
    for (int k = 0; k < kScanItems; ++k)
    {
        uint32_t i = base + threadIdx.x * kScanItems + k;
        if (i < n)
            out[i] = prefix + local[k];
    }
    (void)buf;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        out[n] = *total;
}
} // namespace rivehip
