// Image decoding on the device: the per-pixel half of decoding an encoded
// image (RenderContext::decodeImage, render_context.cpp:237-262), after the
// host has undone the entropy coding (host/image_decode.cpp).
//
//  * jpeg_idct_kernel: one thread per 8x8 coefficient block: dequantization
//    and the inverse DCT.  Full-resolution components use the accurate
//    integer IDCT of IJG libjpeg (jidctint.c jpeg_idct_islow: the
//    Loeffler-Ligtenberg-Moschytz flow graph in 13-bit fixed point, two passes
//    with 2 extra bits between them).  Subsampled chroma is upsampled by the
//    transform itself, as libjpeg 7 and later do by default
//    (jdmaster.c: a 16-point output per 8 coefficients where the sampling
//    ratio allows): the block's cosine series evaluated at twice the sample
//    density, here in double precision and rounded.
//  * jpeg_color_kernel: one thread per output pixel: any remaining integer
//    upsampling by replication (jdsample.c int_upsample), then YCbCr -> RGB
//    with libjpeg 9's fixed-point tables (jdcolor.c build_ycc_rgb_table), to
//    RGBA8 with alpha 255 in the image's first level.
//  * premultiply_kernel: unpremultiplied RGBA8 to premultiplied, rounding as
//    the reference's Bitmap::pixelFormat does (simd::div255,
//    bitmap_decoder.cpp:66-90).
#pragma once

#include "device_common.hpp"

namespace rivehip
{
struct JpegPlane
{
    const int16_t* coef; // blocksW * blocksH blocks of 64, natural order
    const uint16_t* qt;  // 64 quantizers, natural order
    uint8_t* samples;    // (blocksH * 8 * scaleV) rows of `stride` bytes
    uint32_t blocksW, blocksH, stride, scaleH, scaleV;
};

// 8-bit sample from an IDCT output (the true sample minus 128), clamped.
__device__ __forceinline__ uint8_t idct_sample(int32_t v) { return (uint8_t)min(max(v + 128, 0), 255); }

// jpeg_idct_islow: CONST_BITS 13, PASS1_BITS 2.
__device__ void idct_islow(const int32_t* in, uint8_t* out, uint32_t stride)
{
    constexpr int kConstBits = 13, kPass1Bits = 2;
    constexpr int32_t F0_298631336 = 2446, F0_390180644 = 3196, F0_541196100 = 4433, F0_765366865 = 6270,
                      F0_899976223 = 7373, F1_175875602 = 9633, F1_501321110 = 12299, F1_847759065 = 15137,
                      F1_961570560 = 16069, F2_053119869 = 16819, F2_562915447 = 20995, F3_072711026 = 25172;
    int32_t ws[64];
    // columns
    for (int c = 0; c < 8; ++c)
    {
        const int32_t* col = in + c;
        if (col[8] == 0 && col[16] == 0 && col[24] == 0 && col[32] == 0 && col[40] == 0 && col[48] == 0 &&
            col[56] == 0)
        {
            const int32_t dc = col[0] * (1 << kPass1Bits);
            for (int r = 0; r < 8; ++r)
                ws[r * 8 + c] = dc;
            continue;
        }
        int32_t z2 = col[16], z3 = col[48];
        int32_t z1 = (z2 + z3) * F0_541196100;
        int32_t tmp2 = z1 + z3 * (-F1_847759065);
        int32_t tmp3 = z1 + z2 * F0_765366865;
        z2 = col[0];
        z3 = col[32];
        int32_t tmp0 = (z2 + z3) * (1 << kConstBits);
        int32_t tmp1 = (z2 - z3) * (1 << kConstBits);
        const int32_t t10 = tmp0 + tmp3, t13 = tmp0 - tmp3, t11 = tmp1 + tmp2, t12 = tmp1 - tmp2;
        tmp0 = col[56];
        tmp1 = col[40];
        tmp2 = col[24];
        tmp3 = col[8];
        z1 = tmp0 + tmp3;
        z2 = tmp1 + tmp2;
        z3 = tmp0 + tmp2;
        int32_t z4 = tmp1 + tmp3;
        const int32_t z5 = (z3 + z4) * F1_175875602;
        tmp0 *= F0_298631336;
        tmp1 *= F2_053119869;
        tmp2 *= F3_072711026;
        tmp3 *= F1_501321110;
        z1 *= -F0_899976223;
        z2 *= -F2_562915447;
        z3 *= -F1_961570560;
        z4 *= -F0_390180644;
        z3 += z5;
        z4 += z5;
        tmp0 += z1 + z3;
        tmp1 += z2 + z4;
        tmp2 += z2 + z3;
        tmp3 += z1 + z4;
        constexpr int s1 = kConstBits - kPass1Bits;
        constexpr int32_t r1 = 1 << (s1 - 1);
        ws[0 * 8 + c] = (t10 + tmp3 + r1) >> s1;
        ws[7 * 8 + c] = (t10 - tmp3 + r1) >> s1;
        ws[1 * 8 + c] = (t11 + tmp2 + r1) >> s1;
        ws[6 * 8 + c] = (t11 - tmp2 + r1) >> s1;
        ws[2 * 8 + c] = (t12 + tmp1 + r1) >> s1;
        ws[5 * 8 + c] = (t12 - tmp1 + r1) >> s1;
        ws[3 * 8 + c] = (t13 + tmp0 + r1) >> s1;
        ws[4 * 8 + c] = (t13 - tmp0 + r1) >> s1;
    }
    // rows
    constexpr int s2 = kConstBits + kPass1Bits + 3;
    constexpr int32_t r2 = 1 << (s2 - 1);
    for (int r = 0; r < 8; ++r)
    {
        const int32_t* w = ws + r * 8;
        uint8_t* o = out + (size_t)r * stride;
        if (w[1] == 0 && w[2] == 0 && w[3] == 0 && w[4] == 0 && w[5] == 0 && w[6] == 0 && w[7] == 0)
        {
            const uint8_t v = idct_sample((w[0] + (1 << (kPass1Bits + 2))) >> (kPass1Bits + 3));
            for (int c = 0; c < 8; ++c)
                o[c] = v;
            continue;
        }
        int32_t z2 = w[2], z3 = w[6];
        int32_t z1 = (z2 + z3) * F0_541196100;
        int32_t tmp2 = z1 + z3 * (-F1_847759065);
        int32_t tmp3 = z1 + z2 * F0_765366865;
        int32_t tmp0 = (w[0] + w[4]) * (1 << kConstBits);
        int32_t tmp1 = (w[0] - w[4]) * (1 << kConstBits);
        const int32_t t10 = tmp0 + tmp3, t13 = tmp0 - tmp3, t11 = tmp1 + tmp2, t12 = tmp1 - tmp2;
        tmp0 = w[7];
        tmp1 = w[5];
        tmp2 = w[3];
        tmp3 = w[1];
        z1 = tmp0 + tmp3;
        z2 = tmp1 + tmp2;
        z3 = tmp0 + tmp2;
        int32_t z4 = tmp1 + tmp3;
        const int32_t z5 = (z3 + z4) * F1_175875602;
        tmp0 *= F0_298631336;
        tmp1 *= F2_053119869;
        tmp2 *= F3_072711026;
        tmp3 *= F1_501321110;
        z1 *= -F0_899976223;
        z2 *= -F2_562915447;
        z3 *= -F1_961570560;
        z4 *= -F0_390180644;
        z3 += z5;
        z4 += z5;
        tmp0 += z1 + z3;
        tmp1 += z2 + z4;
        tmp2 += z2 + z3;
        tmp3 += z1 + z4;
        o[0] = idct_sample((t10 + tmp3 + r2) >> s2);
        o[7] = idct_sample((t10 - tmp3 + r2) >> s2);
        o[1] = idct_sample((t11 + tmp2 + r2) >> s2);
        o[6] = idct_sample((t11 - tmp2 + r2) >> s2);
        o[2] = idct_sample((t12 + tmp1 + r2) >> s2);
        o[5] = idct_sample((t12 - tmp1 + r2) >> s2);
        o[3] = idct_sample((t13 + tmp0 + r2) >> s2);
        o[4] = idct_sample((t13 - tmp0 + r2) >> s2);
    }
}

// C(u) cos((2x+1) u pi / 2n) for n = 8 (rows 0..7) and n = 16 (rows 8..23),
// x = row index within its block, u = column; uploaded once per device.
__constant__ double kIdctBasis[24][8];

// The block's cosine series at (8*scaleH) x (8*scaleV) sample positions.
__device__ void idct_scaled(const int32_t* in, uint8_t* out, uint32_t stride, int nh, int nv)
{
    const double(*bh)[8] = nh == 16 ? kIdctBasis + 8 : kIdctBasis;
    const double(*bv)[8] = nv == 16 ? kIdctBasis + 8 : kIdctBasis;
    double tmp[8][16];
    for (int v = 0; v < 8; ++v)
        for (int x = 0; x < nh; ++x)
        {
            double s = 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                s += bh[x][u] * (double)in[v * 8 + u];
            tmp[v][x] = s;
        }
    for (int y = 0; y < nv; ++y)
        for (int x = 0; x < nh; ++x)
        {
            double s = 0.0;
#pragma unroll
            for (int v = 0; v < 8; ++v)
                s += bv[y][v] * tmp[v][x];
            out[(size_t)y * stride + x] = (uint8_t)fmin(fmax(floor(0.25 * s + 128.5), 0.0), 255.0);
        }
}

__global__ __launch_bounds__(64) void jpeg_idct_kernel(JpegPlane p)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.blocksW * p.blocksH)
        return;
    const uint32_t bx = b % p.blocksW, by = b / p.blocksW;
    int32_t dq[64];
    const int16_t* c = p.coef + (size_t)b * 64;
#pragma unroll 8
    for (int k = 0; k < 64; ++k)
        dq[k] = (int32_t)c[k] * (int32_t)p.qt[k];
    uint8_t* out = p.samples + (size_t)by * 8 * p.scaleV * p.stride + (size_t)bx * 8 * p.scaleH;
    if (p.scaleH == 1 && p.scaleV == 1)
        idct_islow(dq, out, p.stride);
    else
        idct_scaled(dq, out, p.stride, 8 * (int)p.scaleH, 8 * (int)p.scaleV);
}

struct JpegColor
{
    const uint8_t* plane[3];
    uint32_t stride[3], repH[3], repV[3];
    uint32_t components, ycc, width, height;
    uint32_t* out; // width * height RGBA8 texels
};

__global__ __launch_bounds__(256) void jpeg_color_kernel(JpegColor c)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= c.width || y >= c.height)
        return;
    int s[3] = {0, 0, 0};
    for (uint32_t k = 0; k < c.components; ++k)
        s[k] = c.plane[k][(size_t)(y / c.repV[k]) * c.stride[k] + x / c.repH[k]];
    int r, g, b;
    if (c.components == 1)
        r = g = b = s[0];
    else if (!c.ycc)
        r = s[0], g = s[1], b = s[2];
    else
    {
        // FIX(v) = round(v * 2^16); DESCALE rounds half up.
        const int cb = s[1] - 128, cr = s[2] - 128;
        const int crR = (91881 * cr + 32768) >> 16;   // 1.402
        const int cbB = (116130 * cb + 32768) >> 16;  // 1.772
        const int g16 = -46802 * cr + (-22553 * cb + 32768); // 0.714136286, 0.344136286
        r = s[0] + crR;
        g = s[0] + (g16 >> 16);
        b = s[0] + cbB;
    }
    r = min(max(r, 0), 255);
    g = min(max(g, 0), 255);
    b = min(max(b, 0), 255);
    c.out[(size_t)y * c.width + x] = (uint32_t)r | (uint32_t)g << 8 | (uint32_t)b << 16 | 0xff000000u;
}

__device__ __forceinline__ uint32_t div255(uint32_t x)
{
    x += 128u;
    return (x + (x >> 8)) >> 8;
}

__global__ __launch_bounds__(256) void premultiply_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                          size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t p = in[i], a = p >> 24;
    if (a == 255u)
    {
        out[i] = p;
        return;
    }
    const uint32_t r = div255((p & 0xffu) * a), g = div255(((p >> 8) & 0xffu) * a),
                   b = div255(((p >> 16) & 0xffu) * a);
    out[i] = r | g << 8 | b << 16 | a << 24;
}
} // namespace rivehip
