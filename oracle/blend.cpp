// CPU ORACLE (test infrastructure only): the advanced blend equations of the
// reference's shaders, restated in fp32.
//   advanced_blend.glsl:87-280 (HSL helpers, advanced_blend_coeffs)
//   common.glsl:210-216 (unmultiply_rgb)
//   gpu.cpp:717-753 (ConvertBlendModeToPLSBlendMode)
#include "oracle.hpp"

#include <algorithm>
#include <cmath>

namespace oracle
{
uint32_t shader_blend_mode(uint32_t m)
{
    switch (m)
    {
        case 14: return BLEND_MODE_SCREEN;
        case 15: return BLEND_MODE_OVERLAY;
        case 16: return BLEND_MODE_DARKEN;
        case 17: return BLEND_MODE_LIGHTEN;
        case 18: return BLEND_MODE_COLORDODGE;
        case 19: return BLEND_MODE_COLORBURN;
        case 20: return BLEND_MODE_HARDLIGHT;
        case 21: return BLEND_MODE_SOFTLIGHT;
        case 22: return BLEND_MODE_DIFFERENCE;
        case 23: return BLEND_MODE_EXCLUSION;
        case 24: return BLEND_MODE_MULTIPLY;
        case 25: return BLEND_MODE_HUE;
        case 26: return BLEND_MODE_SATURATION;
        case 27: return BLEND_MODE_COLOR;
        case 28: return BLEND_MODE_LUMINOSITY;
        default: return BLEND_SRC_OVER;
    }
}

namespace
{
constexpr float kEpsFp16NonDenorm = 6.2e-5f; // EPSILON_FP16_NON_DENORM (constants.glsl:233)
float min3(const float c[3]) { return std::min(c[0], std::min(c[1], c[2])); }
float max3(const float c[3]) { return std::max(c[0], std::max(c[1], c[2])); }
float lum(const float c[3]) { return c[0] * .30f + c[1] * .59f + c[2] * .11f; }
float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

// set_lum (advanced_blend.glsl:97-124)
void set_lum(const float base[3], const float lumColor[3], float out[3])
{
    const float lumTarget = lum(lumColor);
    float biased[3];
    const float l = lum(base);
    for (int i = 0; i < 3; ++i)
        biased[i] = base[i] - l;
    const float sx = lumTarget / std::max(kEpsFp16NonDenorm, -min3(biased));
    const float sy = (1.f - lumTarget) / std::max(kEpsFp16NonDenorm, max3(biased));
    const float satScale = std::min(1.f, std::min(sx, sy));
    for (int i = 0; i < 3; ++i)
        out[i] = biased[i] * satScale + lumTarget;
}

// set_lum_sat (advanced_blend.glsl:128-149)
void set_lum_sat(const float hue[3], const float sat[3], const float lumColor[3], float out[3])
{
    const float satTarget = max3(sat) - min3(sat);
    float h[3];
    const float mn = min3(hue);
    for (int i = 0; i < 3; ++i)
        h[i] = hue[i] - mn;
    const float satSource = max3(h);
    const float scale = satTarget / std::max(kEpsFp16NonDenorm, satSource);
    for (int i = 0; i < 3; ++i)
        h[i] *= scale;
    set_lum(h, lumColor, out);
}
} // namespace

void advanced_blend_coeffs(const float srcIn[3], const float dstPremulIn[4], uint32_t mode, float out[3])
{
    float src[3] = {srcIn[0], srcIn[1], srcIn[2]};
    float dp[4] = {dstPremulIn[0], dstPremulIn[1], dstPremulIn[2], dstPremulIn[3]};
    float dst[3];
    const float inva = dp[3] != 0.f ? 1.f / dp[3] : 0.f;
    for (int i = 0; i < 3; ++i)
        dst[i] = dp[i] * inva;
    switch (mode)
    {
        case BLEND_MODE_MULTIPLY:
            for (int i = 0; i < 3; ++i)
                out[i] = src[i] * dst[i];
            break;
        case BLEND_MODE_SCREEN:
            for (int i = 0; i < 3; ++i)
                out[i] = src[i] + dst[i] - src[i] * dst[i];
            break;
        case BLEND_MODE_OVERLAY:
            for (int i = 0; i < 3; ++i)
            {
                const float sd = src[i] * dst[i];
                out[i] = 2.f * (dst[i] > .5f ? src[i] + dst[i] - sd - .5f : sd);
            }
            break;
        case BLEND_MODE_DARKEN:
            for (int i = 0; i < 3; ++i)
                out[i] = std::min(src[i], dst[i]);
            break;
        case BLEND_MODE_LIGHTEN:
            for (int i = 0; i < 3; ++i)
                out[i] = std::max(src[i], dst[i]);
            break;
        case BLEND_MODE_COLORDODGE:
            for (int i = 0; i < 3; ++i)
            {
                const float d = clampf(dp[i], 0.f, dp[3]);
                const float denom = clampf(1.f - src[i], 0.f, 1.f) * dp[3];
                out[i] = denom == 0.f ? signf(d) : std::min(1.f, d / denom);
            }
            break;
        case BLEND_MODE_COLORBURN:
        {
            float a = dp[3];
            for (int i = 0; i < 3; ++i)
            {
                const float s = clampf(src[i], 0.f, 1.f);
                const float d = clampf(dp[i], 0.f, a);
                const float aa = a == 0.f ? 1.f : a;
                const float numer = aa - d;
                out[i] = 1.f - (s == 0.f ? signf(numer) : std::min(1.f, numer / (s * aa)));
            }
            break;
        }
        case BLEND_MODE_HARDLIGHT:
            for (int i = 0; i < 3; ++i)
            {
                const float sd = src[i] * dst[i];
                out[i] = 2.f * (src[i] > .5f ? src[i] + dst[i] - sd - .5f : sd);
            }
            break;
        case BLEND_MODE_SOFTLIGHT:
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
        case BLEND_MODE_DIFFERENCE:
            for (int i = 0; i < 3; ++i)
                out[i] = fabsf(dst[i] - src[i]);
            break;
        case BLEND_MODE_EXCLUSION:
            for (int i = 0; i < 3; ++i)
                out[i] = src[i] + dst[i] - 2.f * src[i] * dst[i];
            break;
        case BLEND_MODE_HUE:
        case BLEND_MODE_SATURATION:
        case BLEND_MODE_COLOR:
        case BLEND_MODE_LUMINOSITY:
        {
            for (int i = 0; i < 3; ++i)
                src[i] = clampf(src[i], 0.f, 1.f);
            if (mode == BLEND_MODE_HUE)
                set_lum_sat(src, dst, dst, out);
            else if (mode == BLEND_MODE_SATURATION)
                set_lum_sat(dst, src, dst, out);
            else if (mode == BLEND_MODE_COLOR)
                set_lum(src, dst, out);
            else
                set_lum(dst, src, out);
            break;
        }
        default:
            for (int i = 0; i < 3; ++i)
                out[i] = src[i];
            break;
    }
}
} // namespace oracle
