// CPU ORACLE (test infrastructure only): feathering, restated from the
// reference -- the Gaussian-integral function tables and their lookup, the
// curvature softening of feathered fills on the CPU and in the tessellation
// shader, and the feathered coverage evaluation of the path shaders.
// Compiled with -ffp-contract=off.  See oracle.hpp.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>

namespace oracle
{
namespace
{
constexpr float kPI = 3.14159265f;
constexpr uint32_t kTableSize = GAUSSIAN_TABLE_SIZE;
constexpr float kStddevs = 3.f; // GAUSSIAN_INTEGRAL_TEXTURE_STDDEVS (constants.glsl:21)

inline uint32_t fbits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }

// eval_normal_distribution (gpu.cpp:1984-1989): the exponent is formed in
// double (the literal -.5 is a double) and handed to expf.
float eval_normal_distribution(float x, float mu, float inverseSigma)
{
    const float ONE_OVER_SQRT_2_PI = 0.398942280401433f;
    float y = (x - mu) * inverseSigma;
    return expf((float)(-.5 * (double)y * (double)y)) * inverseSigma * ONE_OVER_SQRT_2_PI;
}
} // namespace

// gpu::cast_f32_to_f16 (gpu.cpp:1959-1980): the dropped mantissa bits are
// rounded half up by adding 0x1000 first; exponents beyond fp16 saturate.
uint16_t cast_f32_to_f16(float x)
{
    uint32_t b = fbits(x) + 0x00001000u;
    uint32_t e = (b & 0x7F800000u) >> 23;
    uint32_t m = b & 0x007FFFFFu;
    uint32_t r = (b & 0x80000000u) >> 16;
    if (e > 112u)
        r |= (((e - 112u) << 10) & 0x7C00u) | m >> 13;
    if (e < 113u && e > 101u)
        r |= (((0x007FF000u + m) >> (125u - e)) + 1u) >> 1;
    if (e > 143u)
        r |= 0x7FFFu;
    return (uint16_t)r;
}

float cast_f16_to_f32(uint16_t h)
{
    // Exact widening (gpu.cpp:1944-1957).
    const uint32_t s = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0)
        v = ldexpf((float)m, -24);
    else if (e == 31)
        v = m ? std::numeric_limits<float>::quiet_NaN() : INFINITY;
    else
        v = ldexpf((float)(m | 0x400u), (int)e - 25);
    uint32_t u = fbits(v) | s;
    memcpy(&v, &u, 4);
    return v;
}

// generate_gausian_integral_table (gpu.cpp:1991-2063): a running integral of
// the normal distribution (mu = 256, sigma = 512/6, 7 samples per entry),
// shifted so the middle is 1/2, made monotonic and clamped to [0, 1], then
// eased into 0 and 1 over the first and last 8 entries.
void gaussian_integral_table(float table[GAUSSIAN_TABLE_SIZE])
{
    float sigma = kTableSize / (kStddevs * 2);
    float inverseSigma = 1 / sigma;
    float mu = kTableSize * .5f;
    float integral = 0;
    for (uint32_t i = 0; i < kTableSize; ++i)
    {
        const int SAMPLES = 7;
        float barCenterX = (float)i;
        for (int sample = 0; sample < SAMPLES; ++sample)
        {
            float dx = (float)(sample - (SAMPLES >> 1)) / SAMPLES;
            integral += eval_normal_distribution(barCenterX + dx, mu, inverseSigma) / SAMPLES;
        }
        table[i] = integral;
    }
    float shift = (float)(.5 - (table[kTableSize / 2 - 1] + table[kTableSize / 2]) / 2);
    table[0] = fminf(fmaxf(0, table[0] + shift), 1);
    for (uint32_t i = 1; i < kTableSize; ++i)
        table[i] = fminf(fmaxf(table[i - 1], table[i] + shift), 1);
    const uint32_t EASE = 8;
    for (uint32_t i = 0; i < kTableSize; ++i)
    {
        if (i < EASE)
            table[i] *= (float)i / (float)EASE;
        if (i > (kTableSize - EASE) - 1)
        {
            float diffToOne = (float)(1.0 - table[i]);
            float fi = (float)(i - (kTableSize - EASE) + 1) / (float)EASE;
            table[i] = table[i] + (diffToOne * fi);
        }
    }
}

// generate_inverse_gausian_integral_table (gpu.cpp:2126-2181): the same
// integral at 32 samples per entry, inverted by interpolating each sample
// interval that crosses a cell center.
void inverse_gaussian_integral_table(float table[GAUSSIAN_TABLE_SIZE])
{
    const uint32_t MULTIPLIER = 32;
    float sigma = kTableSize / (kStddevs * 2);
    float inverseSigma = 1 / sigma;
    float mu = kTableSize * .5f;
    const uint32_t samples = kTableSize * MULTIPLIER;
    float integral = 0;
    for (uint32_t i = 0; i < (samples + 1) / 2; ++i)
    {
        float barCenterX = (float)i / MULTIPLIER;
        integral += eval_normal_distribution(barCenterX, mu, inverseSigma) / MULTIPLIER;
    }
    integral = (float)(.5 - integral);
    float lastInverseX = std::numeric_limits<float>::quiet_NaN(), lastInverseY = 0;
    for (uint32_t i = 0; i < kTableSize; ++i)
        table[i] = 0;
    table[0] = 0;
    table[kTableSize - 1] = 1;
    for (uint32_t i = 0; i < samples; ++i)
    {
        float barCenterX = (float)i / MULTIPLIER;
        integral += eval_normal_distribution(barCenterX, mu, inverseSigma) / MULTIPLIER;
        float inverseX = fminf(fmaxf(0, integral), 1) * kTableSize;
        float inverseY = (i + .5f) / samples;
        uint32_t cell = (uint32_t)inverseX;
        float cellCenterX = cell + .5f;
        if (cellCenterX == mu)
            table[cell] = .5f;
        else if (lastInverseX <= cellCenterX && inverseX >= cellCenterX)
        {
            float t = (cellCenterX - lastInverseX) / (inverseX - lastInverseX);
            table[cell] = lastInverseY * (1 - t) + inverseY * t; // lerp (math_types.hpp:247-250)
        }
        lastInverseX = inverseX;
        lastInverseY = inverseY;
    }
}

namespace
{
struct Tables
{
    uint16_t f16[2][GAUSSIAN_TABLE_SIZE];
};
const Tables& tables()
{
    static Tables t;
    static std::once_flag once;
    std::call_once(once, [] {
        float f[GAUSSIAN_TABLE_SIZE];
        gaussian_integral_table(f);
        for (uint32_t i = 0; i < kTableSize; ++i)
            t.f16[0][i] = cast_f32_to_f16(f[i]);
        inverse_gaussian_integral_table(f);
        for (uint32_t i = 0; i < kTableSize; ++i)
            t.f16[1][i] = cast_f32_to_f16(f[i]);
    });
    return t;
}
} // namespace

const uint16_t* feather_table_f16() { return tables().f16[0]; }
const uint16_t* inverse_feather_table_f16() { return tables().f16[1]; }

// A 1D texture of fp16 texels sampled with linear filtering and clamp-to-edge
// at normalized coordinate x (the FEATHER / INVERSE_FEATHER macros,
// common.glsl:53-68), as the reference's own test looks tables up
// (gpu_namespace_test.cpp:68-77).
float table_lookup(const uint16_t* table, float x)
{
    x = fminf(fmaxf(0, x), 1);
    float sampleRegionLeft = x * kTableSize - .5f;
    int rightIdx = (int)fminf(sampleRegionLeft + 1, kTableSize - 1);
    int leftIdx = std::max(rightIdx - 1, 0);
    float t = fminf(fmaxf(0, sampleRegionLeft - leftIdx), 1);
    float a = cast_f16_to_f32(table[leftIdx]), b = cast_f16_to_f32(table[rightIdx]);
    return a * (1 - t) + b * t; // lerp (math_types.hpp:247-250)
}
float feather(float x) { return table_lookup(feather_table_f16(), x); }
float inverse_feather(float x) { return table_lookup(inverse_feather_table_f16(), x); }

// ---------------- the path shaders' feather evaluation --------------------

// pack_feathered_fill_coverages (draw_path_common.glsl:107-150).
void pack_feathered_fill_coverages(float cornerTheta, V2 spokeNorm, float outset, float out[4])
{
    V2 cornerLocalCoord{(1.f - spokeNorm.x * fabsf(outset)) * .5f, (1.f - spokeNorm.y * fabsf(outset)) * .5f};
    float cotTheta, y0;
    if (fabsf(cornerTheta - kPI_OVER_2) < 1.f / HORIZONTAL_COTANGENT_THRESHOLD)
    {
        cotTheta = 0;
        y0 = 0;
    }
    else
    {
        float tanTheta = tanf(cornerTheta);
        float s = kPI_OVER_2 - cornerTheta;
        float sgn = s > 0 ? 1.f : (s < 0 ? -1.f : 0.f);
        cotTheta = sgn / std::max(fabsf(tanTheta), 1.f / HORIZONTAL_COTANGENT_VALUE);
        y0 = cotTheta >= 0 ? cornerLocalCoord.y - (1.f - cornerLocalCoord.x) * tanTheta
                           : cornerLocalCoord.y + cornerLocalCoord.x * tanTheta;
    }
    out[0] = std::max(cornerLocalCoord.x, 0.f) + FEATHER_X_COORD_BIAS;
    out[1] = -cornerLocalCoord.y + FEATHER_COVERAGE_BIAS;
    out[2] = cotTheta;
    out[3] = y0;
}

// eval_feathered_fill (draw_path_common.glsl:153-240): the feather above y0
// plus a 4-sample midpoint integral of FEATHER(u) * FEATHER'(t) across the
// corner, signed by coverages.x.
float eval_feathered_fill(const float c[4])
{
    float cotTheta = c[2];
    float y0 = std::max(c[3], 0.f);
    float featherCoverage = cotTheta >= 0 ? feather(y0) : 0.f;
    if (fabsf(cotTheta) < HORIZONTAL_COTANGENT_THRESHOLD)
    {
        float x = fabsf(c[0]) - FEATHER_X_COORD_BIAS;
        float y = -c[1] + FEATHER_COVERAGE_BIAS;
        float dt = (y - y0) * 0.5984134206f;
        const float k[4] = {0.20888568955f, 0.62665706865f, 1.04442844776f, 1.46219982687f};
        float sum = 0;
        for (int i = 0; i < 4; ++i)
        {
            float t = y0 + dt * k[i];
            float u = t * -cotTheta + (y * cotTheta + x);
            float t_ = t * 5.09593080173f + -2.54796540086f;
            sum += feather(u) * exp2f(-t_ * t_);
        }
        featherCoverage += sum * dt;
    }
    float sx = c[0] > 0 ? 1.f : (c[0] < 0 ? -1.f : 0.f);
    return featherCoverage * sx;
}

// eval_feathered_stroke (draw_path_common.glsl:242-258).
float eval_feathered_stroke(const float c[4])
{
    float featherCoverage = 1.f;
    float leftOutsideCoverage = (1.f - FEATHER_COVERAGE_BIAS) + c[0];
    featherCoverage -= feather(leftOutsideCoverage);
    float rightOutsideCoverage = 1.f - c[1];
    featherCoverage -= feather(rightOutsideCoverage);
    return featherCoverage;
}

// ---------------- curvature softening in the tessellation shader ----------

// clamped_divide (bezier_utils.glsl:158-169).
float clamped_divide(float a, float b)
{
    a = b < 0 ? -a : a;
    b = fabsf(b);
    return a > 0 ? (a < b ? a / b : 1.f) : 0.f;
}

// find_cubic_max_height (bezier_utils.glsl:171-246, the Newton-Raphson
// branch): the height above the baseline p0->p3 and where it peaks.
float find_cubic_max_height(V2 p0, V2 p1, V2 p2, V2 p3, float* outT)
{
    V2 base = p3 - p0;
    float lengthBase = sqrtf(dot(base, base));
    if (lengthBase == 0)
    {
        *outT = .5f;
        return 0;
    }
    V2 norm{-base.y / lengthBase, base.x / lengthBase};
    float h2 = dot(norm, p2 - p0);
    float h1 = dot(norm, p1 - p0);
    float dh = h1 - h2;
    float _3A = 3.f * dh;
    float B = -h1 - dh;
    float C = h1;
    float t = .5f;
    for (int i = 0; i < 3; ++i)
    {
        float _3At = _3A * t;
        t = clamped_divide(_3At * t - C, 2.f * (_3At + B));
    }
    *outT = t;
    return fabsf(t * (t * (t * _3A + 3.f * B) + 3.f * C));
}

namespace
{
// cosine_between_vectors (bezier_utils.glsl:24-31).
float cosine_between_vectors(V2 a, V2 b)
{
    float ab_cosTheta = dot(a, b);
    float ab_pow2 = dot(a, a) * dot(b, b);
    return ab_pow2 == 0 ? 1.f : fminf(fmaxf(ab_cosTheta * (1.f / sqrtf(ab_pow2)), -1.f), 1.f);
}
// find_cubic_tangents (bezier_utils.glsl:64-70).
void glsl_cubic_tangents(V2 p0, V2 p1, V2 p2, V2 p3, V2 t[2])
{
    t[0] = (p0 != p1 ? p1 : p1 != p2 ? p2 : p3) - p0;
    t[1] = p3 - (p3 != p2 ? p2 : p2 != p1 ? p1 : p0);
}
} // namespace

// measure_cubic_local_curvature (bezier_utils.glsl:72-156): the rotation, in
// radians, of the tangent across a spread of "desiredSpread" (in local units,
// along the tangent) centered at T.
float measure_cubic_local_curvature(V2 p0, V2 p1, V2 p2, V2 p3, float T, float desiredSpread)
{
    V2 C = p1 - p0, D = p2 - p1, E = p3 - p0;
    V2 B = D - C;
    V2 A = D * -3.f + E;
    V2 tangent = ((A * T) + B * 2.f) * T + C;
    tangent = tangent * 3.f;
    float lengthTan = sqrtf(dot(tangent, tangent));
    if (lengthTan == 0)
        return 0;
    tangent = tangent * (1.f / lengthTan);
    float A_ = 2.f * dot(A, tangent);
    float C_ = 3.f * (A_ * T + 4.f * dot(B, tangent)) * T + 6.f * dot(C, tangent);
    float maxDT = fminf(T, 1.f - T);
    float maxSpread = (A_ * maxDT * maxDT + C_) * maxDT;
    float targetSpread = fminf(desiredSpread, maxSpread * .9999f);
    float dt;
    if (A_ == 0)
        dt = targetSpread / C_;
    else
    {
        float r = 1.f / A_;
        float b = C_ * r, c = -targetSpread * r;
        float Q = (-1.f / 3.f) * b, R = .5f * c;
        float discr = R * R - Q * Q * Q;
        if (discr < 0)
        {
            float sqrtQ = sqrtf(Q);
            float theta = acosf(R / (sqrtQ * sqrtQ * sqrtQ));
            dt = -2.f * sqrtQ * cosf(theta * (1.f / 3.f) + (-kPI * 2.f / 3.f));
        }
        else
        {
            float Ar = powf(fabsf(R) + sqrtf(discr), 1.f / 3.f);
            if (R < 0)
                Ar = -Ar;
            dt = Ar != 0 ? Ar + Q / Ar : 0.f;
        }
    }
    dt = fabsf(dt);
    float t0 = T - dt, t1 = T + dt;
    V2 tan0d = (A * t0 + B * 2.f) * t0 + C;
    V2 tan1d = (A * t1 + B * 2.f) * t1 + C;
    V2 tangents[2];
    glsl_cubic_tangents(p0, p1, p2, p3, tangents);
    V2 tan0 = t0 < 1e-3f ? tangents[0] : tan0d;
    V2 tan1 = t1 > 1.f - 1e-3f ? tangents[1] : tan1d;
    return acosf(cosine_between_vectors(tan0, tan1));
}

// The softening of a feathered fill's cubic in the tessellation shader
// (tessellate.glsl:146-193): the curve's height is reduced so the feather's
// center opacity dims with the curvature across one standard deviation.
void soften_feathered_cubic(V2* p0, V2* p1, V2* p2, V2* p3, float featherRadius)
{
    float maxHeightT;
    float height = find_cubic_max_height(*p0, *p1, *p2, *p3, &maxHeightT);
    float oneStddev = featherRadius * (1.f / kStddevs);
    float curvature = measure_cubic_local_curvature(*p0, *p1, *p2, *p3, maxHeightT, oneStddev);
    float dimming = 1.f - curvature * (1.f / kPI);
    V2 d = *p3 - *p0;
    float stddevsPow2 = dot(d, d) / (oneStddev * oneStddev);
    float dimmingByStddevs = (stddevsPow2 - 1.f) * .5f;
    dimming = fminf(dimming, dimmingByStddevs);
    dimming = fminf(dimming, .99f);
    float desiredOpacityOnCenter = .5f * dimming;
    float x = inverse_feather(desiredOpacityOnCenter) * -2.f + 1.f;
    float softness = clamped_divide(x * featherRadius, height);
    // mix(p0, p3, {1/3, 2/3}), then mix toward them by softness
    V2 f1 = (*p3 - *p0) * (1.f / 3.f) + *p0;
    V2 f2 = (*p3 - *p0) * (2.f / 3.f) + *p0;
    *p1 = (f1 - *p1) * softness + *p1;
    *p2 = (f2 - *p2) * softness + *p2;
}

// ---------------- curvature softening on the CPU --------------------------

namespace
{
// chop_cubic_at_uniform_rotation (rive_render_path.cpp:233-321).
void chop_cubic_at_uniform_rotation(std::vector<uint8_t>* verbs, std::vector<V2>* pts, const V2 p[4],
                                    const V2 tangents[2], int numChops, const float rot[4])
{
    // CubicCoeffs (bezier_utils.hpp:36-56)
    V2 C = p[1] - p[0], D = p[2] - p[1], E = p[3] - p[0];
    V2 B = D - C;
    V2 A = D * -3.f + E;
    V2 tangent = tangents[0];
    V2 rem[4] = {p[0], p[1], p[2], p[3]}; // the cubic still to be chopped
    float remainingT = 0;
    V2 chops[10];
    for (int i = 0; i < numChops; i += 4)
    {
        int numChopsRemaining = std::min(numChops - i, 4);
        float tx[4] = {0, 0, 0, 0}, ty[4] = {0, 0, 0, 0};
        for (int j = 0; j < numChopsRemaining; ++j)
        {
            // m = rotation * tangent.xxyy; tangent = m.xy + m.zw
            V2 nt{rot[0] * tangent.x + rot[2] * tangent.y, rot[1] * tangent.x + rot[3] * tangent.y};
            tangent = nt;
            tx[j] = tangent.x;
            ty[j] = tangent.y;
        }
        float roots[4];
        for (int j = 0; j < 4; ++j)
        {
            float a = A.x * ty[j] - A.y * tx[j];
            float b_over_2 = B.x * ty[j] - B.y * tx[j];
            float c = C.x * ty[j] - C.y * tx[j];
            float discr_over_4 = b_over_2 * b_over_2 - a * c;
            float q = sqrtf(discr_over_4);
            q = -b_over_2 - copysignf(q, b_over_2);
            roots[j] = c / q;
        }
        float t[4];
        int numT = 0;
        float maxT = remainingT;
        for (int j = 0; j < numChopsRemaining; ++j)
        {
            const float MIN_SPACING = 1e-4f;
            if (roots[j] > maxT + MIN_SPACING && roots[j] < 1 - MIN_SPACING)
                t[numT++] = maxT = roots[j];
        }
        for (int j = 0; j < numT; j += 2)
        {
            auto localize = [&](float tv) {
                return fminf(fmaxf((tv - remainingT) / (1 - remainingT), 0.f), 1.f);
            };
            if (j + 1 < numT)
            {
                float l0 = localize(t[j]), l1 = localize(t[j + 1]);
                chop_cubic_at2(rem, chops, l0, l1);
                verbs->push_back(kCubic);
                pts->insert(pts->end(), {chops[1], chops[2], chops[3]});
                verbs->push_back(kCubic);
                pts->insert(pts->end(), {chops[4], chops[5], chops[6]});
                for (int k = 0; k < 4; ++k)
                    rem[k] = chops[6 + k];
                remainingT = t[j + 1];
            }
            else
            {
                chop_cubic_at(rem, chops, localize(t[j]));
                verbs->push_back(kCubic);
                pts->insert(pts->end(), {chops[1], chops[2], chops[3]});
                for (int k = 0; k < 4; ++k)
                    rem[k] = chops[3 + k];
                remainingT = t[j];
            }
        }
    }
    verbs->push_back(kCubic);
    pts->insert(pts->end(), {rem[1], rem[2], rem[3]});
}

// math::measure_angle_between_vectors (bezier_utils.cpp:161-170).
float measure_angle_between_vectors(V2 a, V2 b)
{
    float den = sqrtf(dot(a, a) * dot(b, b));
    float cosTheta = dot(a, b) / den; // ieee_float_divide
    cosTheta = std::max(std::min(1.f, cosTheta), -1.f);
    return acosf(cosTheta);
}
} // namespace

// RiveRenderPath::makeSoftenedCopyForFeathering (rive_render_path.cpp:323-449):
// cubics are chopped convex, cusps are crossed with short lines, and each
// piece is chopped wherever its tangent has turned by a fixed angle that
// depends on the feather's size on screen.
void soften_path_for_feathering(const uint8_t* verbs, uint32_t verbCount, const V2* P, uint32_t ptCount,
                                float featherAmount, float matrixMaxScale, std::vector<uint8_t>* outVerbs,
                                std::vector<V2>* outPts)
{
    (void)ptCount;
    outVerbs->clear();
    outPts->clear();
    const float r_ = featherAmount * (kStddevs / 2) * matrixMaxScale * .25f;
    // calc_polar_segments_per_radian<2> (bezier_utils.hpp:108-113)
    float cosTheta = 1.f - (1.f / 2) / r_;
    float polarSegmentsPerRadian = .5f / acosf(std::max(cosTheta, -1.f));
    float rotationBetweenJoins = 1 / polarSegmentsPerRadian;
    const float MIN_ANGLE = FEATHER_POLAR_SEGMENT_MIN_ANGLE;
    if (rotationBetweenJoins < MIN_ANGLE)
    {
        float d = (MIN_ANGLE - rotationBetweenJoins) * 5.f;
        rotationBetweenJoins = MIN_ANGLE + d * d * d;
    }
    rotationBetweenJoins = std::min(rotationBetweenJoins, kPI / 2);
    // Mat2D::fromRotation (mat2d.cpp:10-19): {c, s, -s, c}
    auto fromRotation = [](float rad, float m[4]) {
        float s = 0, c = 1;
        if (rad != 0)
        {
            s = (float)sin((double)rad);
            c = (float)cos((double)rad);
        }
        m[0] = c;
        m[1] = s;
        m[2] = -s;
        m[3] = c;
    };
    float rotation[4], reverseRotation[4];
    fromRotation(rotationBetweenJoins, rotation);
    fromRotation(-rotationBetweenJoins, reverseRotation);

    // RawPath (raw_path.cpp:59-106): a line or cubic after a close first
    // re-opens the contour at its move point; close only closes an open one.
    bool contourIsOpen = false;
    size_t lastMoveIdx = 0;
    auto injectImplicitMove = [&] {
        if (!contourIsOpen)
        {
            V2 m = outPts->empty() ? V2{0, 0} : (*outPts)[lastMoveIdx];
            contourIsOpen = true;
            lastMoveIdx = outPts->size();
            outVerbs->push_back(kMove);
            outPts->push_back(m);
        }
    };
    uint32_t pi = 0;
    for (uint32_t vi = 0; vi < verbCount; ++vi)
    {
        const uint8_t verb = verbs[vi];
        switch (verb)
        {
            case kMove:
                contourIsOpen = true;
                lastMoveIdx = outPts->size();
                outVerbs->push_back(kMove);
                outPts->push_back(P[pi]);
                pi += 1;
                break;
            case kLine:
                injectImplicitMove();
                outVerbs->push_back(kLine);
                outPts->push_back(P[pi]);
                pi += 1;
                break;
            case kCubic:
            {
                injectImplicitMove();
                const V2* pts = P + pi - 1;
                float T[4] = {0, 0, 0, 0};
                V2 chops[16];
                bool areCusps = false;
                int n = find_cubic_convex_180_chops(pts, T, &areCusps);
                if (areCusps)
                {
                    const float CUSP_PADDING = 1e-2f;
                    for (int i = n - 1; i >= 0; --i)
                    {
                        float minT = i == 0 ? 0.f : (T[i - 1] + T[i]) * .5f;
                        float maxT = i + 1 == n ? 1.f : (T[i + 1] + T[i]) * .5f;
                        T[i * 2 + 1] = fminf(T[i] + CUSP_PADDING, maxT);
                        T[i * 2 + 0] = fmaxf(T[i] - CUSP_PADDING, minT);
                    }
                    n *= 2;
                }
                chop_cubic_at_n(pts, chops, T, n);
                const V2* p = chops;
                for (int i = 0; i <= n; ++i, p += 3)
                {
                    if (areCusps && (i & 1) == 1)
                    {
                        outVerbs->push_back(kLine);
                        outPts->push_back(p[3]);
                        continue;
                    }
                    V2 tangents[2] = {find_cubic_tan0(p), find_cubic_tan1(p)};
                    float rot = measure_angle_between_vectors(tangents[0], tangents[1]);
                    int numChops = (int)(rot / rotationBetweenJoins);
                    float turn = cross(p[2] - p[0], p[3] - p[1]);
                    if (turn == 0)
                        turn = cross(tangents[0], tangents[1]);
                    chop_cubic_at_uniform_rotation(outVerbs, outPts, p, tangents, numChops,
                                                   turn >= 0 ? rotation : reverseRotation);
                }
                pi += 3;
                break;
            }
            case kClose:
                if (contourIsOpen)
                {
                    outVerbs->push_back(kClose);
                    contourIsOpen = false;
                }
                break;
            default:
                break;
        }
    }
}
} // namespace oracle
