// Encoded-image decoding on the host side: the container and entropy-coded
// parts of PNG and JPEG, which are sequential by nature.  The
// per-pixel work that follows (JPEG dequantization, inverse DCT, chroma
// upsampling and colour conversion; premultiplication; mip generation) runs
// on the device (image kernels in csrc/).
//
// Behaviour follows the reference's decoders (decoders/src/decode_png.cpp,
// decode_jpeg.cpp, bitmap_decoder_thirdparty.cpp), which wrap libpng 1.6 and
// IJG libjpeg 9f; the formats themselves are restated from their published
// specifications (PNG: ISO/IEC 15948, zlib/deflate: RFC 1950/1951; JPEG: ITU-T
// T.81).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace rivehip::image
{
enum Format : uint32_t
{
    kUnknown = 0,
    kPng = 1,
    kJpeg = 2,
    kWebp = 3,
};

// The format named by the leading bytes (the fingerprints of
// bitmap_decoder_thirdparty.cpp:36-70).
Format sniff(const uint8_t* bytes, size_t n);

// Width and height from the header alone.
bool probe(const uint8_t* bytes, size_t n, Format* format, uint32_t* width, uint32_t* height, std::string* err);

// A PNG as unpremultiplied RGBA8, after libpng's transformations in
// decode_png.cpp:87-116: palette and low-bit-depth gray expanded, tRNS turned
// into alpha, 16-bit samples stripped to their high byte, gray replicated to
// RGB, and alpha 255 where the image has none.
bool decode_png(const uint8_t* bytes, size_t n, uint32_t* width, uint32_t* height, std::vector<uint8_t>* rgba,
                std::string* err);

// A Huffman-coded 8-bit JPEG (baseline, extended or progressive) after
// entropy decoding of all its scans: the quantized coefficients of every 8x8 block, in
// natural (row-major) order, component by component, each component's blocks
// row-major over its MCU-padded grid.
struct JpegComponent
{
    uint8_t id, h, v, tq;    // component ID, sampling factors, quant table
    uint8_t scaleH, scaleV;  // samples per coefficient block side / 8 (1 or 2)
    uint8_t repH, repV;      // replication of those samples to the output grid
    uint32_t blocksW, blocksH;
    size_t coefOffset;       // first coefficient in Jpeg::coef
};

struct Jpeg
{
    uint32_t width = 0, height = 0;
    uint8_t hmax = 1, vmax = 1;
    bool ycc = true; // YCbCr (else RGB, or gray for one component)
    std::vector<JpegComponent> comps;
    uint16_t qt[4][64] = {}; // natural order
    size_t coefCount = 0;    // coefficients of all components' padded grids
    std::vector<int16_t> coef;
};

// Parses the header (frame, tables, component layout) and, when
// `coefficients` is set, decodes every scan into Jpeg::coef.
bool decode_jpeg(const uint8_t* bytes, size_t n, Jpeg* out, bool coefficients, std::string* err);
} // namespace rivehip::image
