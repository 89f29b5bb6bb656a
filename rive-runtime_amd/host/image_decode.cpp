// PNG and JPEG (sequential and progressive) decoding, host side (see
// image_decode.hpp).
#include "image_decode.hpp"

#include <algorithm>
#include <cstring>

namespace rivehip::image
{
namespace
{
uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
uint32_t be16(const uint8_t* p) { return (uint32_t)p[0] << 8 | p[1]; }

bool fail(std::string* err, const char* msg)
{
    if (err != nullptr)
        *err = msg;
    return false;
}

// ---------------------------------------------------------------- deflate
// RFC 1951.  Canonical Huffman codes are decoded through a table indexed by
// the next kFastBits bits of the (LSB-first) stream, falling back to a
// code-length walk for longer codes.
constexpr int kFastBits = 10;

struct BitReader
{
    const uint8_t* p;
    size_t n, pos = 0;
    uint64_t buf = 0;
    int cnt = 0;
    bool overrun = false;

    uint32_t peek(int k)
    {
        while (cnt < k)
        {
            uint64_t b = 0;
            if (pos < n)
                b = p[pos];
            else
                overrun = true;
            ++pos;
            buf |= b << cnt;
            cnt += 8;
        }
        return (uint32_t)(buf & ((1ull << k) - 1));
    }
    void drop(int k)
    {
        buf >>= k;
        cnt -= k;
    }
    uint32_t bits(int k)
    {
        if (k == 0)
            return 0;
        const uint32_t v = peek(k);
        drop(k);
        return v;
    }
    void align() { drop(cnt & 7); }
};

struct Huffman
{
    uint16_t counts[16] = {};
    uint16_t symbols[320] = {};
    uint16_t fast[1 << kFastBits] = {}; // symbol << 4 | length, 0 = slow path

    // Returns false for an over-subscribed set of lengths; an incomplete set
    // is allowed (as zlib does for single-code distance trees).
    bool build(const uint8_t* lengths, int n)
    {
        std::memset(counts, 0, sizeof(counts));
        std::memset(fast, 0, sizeof(fast));
        for (int i = 0; i < n; ++i)
            counts[lengths[i]]++;
        counts[0] = 0;
        int left = 1;
        for (int len = 1; len < 16; ++len)
        {
            left <<= 1;
            left -= counts[len];
            if (left < 0)
                return false;
        }
        uint16_t offs[16];
        offs[1] = 0;
        for (int len = 1; len < 15; ++len)
            offs[len + 1] = offs[len] + counts[len];
        for (int i = 0; i < n; ++i)
            if (lengths[i] != 0)
                symbols[offs[lengths[i]]++] = (uint16_t)i;
        // canonical codes, MSB-first, placed in the table bit-reversed
        int code = 0, index = 0;
        for (int len = 1; len <= kFastBits; ++len)
        {
            for (int k = 0; k < counts[len]; ++k, ++code, ++index)
            {
                int rev = 0;
                for (int b = 0; b < len; ++b)
                    rev |= ((code >> b) & 1) << (len - 1 - b);
                for (int fill = rev; fill < (1 << kFastBits); fill += 1 << len)
                    fast[fill] = (uint16_t)(symbols[index] << 4 | len);
            }
            code <<= 1;
        }
        return true;
    }

    int decode(BitReader& in) const
    {
        const uint16_t e = fast[in.peek(kFastBits)];
        if (e != 0)
        {
            in.drop(e & 15);
            return e >> 4;
        }
        int code = 0, first = 0, index = 0;
        for (int len = 1; len < 16; ++len)
        {
            code |= (int)in.bits(1);
            const int count = counts[len];
            if (code - count < first)
                return symbols[index + (code - first)];
            index += count;
            first += count;
            first <<= 1;
            code <<= 1;
        }
        return -1;
    }
};

const uint16_t kLengthBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                  31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLengthExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

// Inflates a zlib stream (RFC 1950) into `out`, which must come out exactly
// `expected` bytes long; the Adler-32 trailer is checked.
bool zlib_inflate(const uint8_t* src, size_t n, size_t expected, std::vector<uint8_t>* out, std::string* err)
{
    if (n < 6)
        return fail(err, "zlib stream too short");
    const uint32_t cmf = src[0], flg = src[1];
    if ((cmf & 15) != 8 || (cmf >> 4) > 7 || (cmf * 256 + flg) % 31 != 0)
        return fail(err, "bad zlib header");
    if (flg & 0x20)
        return fail(err, "zlib preset dictionary is not allowed in PNG");
    out->assign(expected, 0);
    uint8_t* o = out->data();
    size_t at = 0;
    BitReader in{src + 2, n - 2};
    Huffman lit, dist;
    bool last = false;
    while (!last)
    {
        last = in.bits(1) != 0;
        const uint32_t type = in.bits(2);
        if (type == 0)
        {
            in.align();
            const uint32_t len = in.bits(16), nlen = in.bits(16);
            if ((len ^ 0xffffu) != nlen)
                return fail(err, "bad stored block length");
            if (at + len > expected)
                return fail(err, "too much image data");
            for (uint32_t i = 0; i < len; ++i)
                o[at++] = (uint8_t)in.bits(8);
        }
        else if (type == 1 || type == 2)
        {
            uint8_t lengths[320];
            uint32_t nlit = 288, ndist = 30;
            if (type == 1)
            {
                for (int i = 0; i < 288; ++i)
                    lengths[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
                for (int i = 0; i < 30; ++i)
                    lengths[288 + i] = 5;
            }
            else
            {
                nlit = in.bits(5) + 257;
                ndist = in.bits(5) + 1;
                const uint32_t ncode = in.bits(4) + 4;
                static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
                uint8_t cl[19] = {};
                for (uint32_t i = 0; i < ncode; ++i)
                    cl[order[i]] = (uint8_t)in.bits(3);
                Huffman clh;
                if (!clh.build(cl, 19))
                    return fail(err, "bad code-length code");
                uint32_t i = 0;
                while (i < nlit + ndist)
                {
                    const int sym = clh.decode(in);
                    if (sym < 0)
                        return fail(err, "bad code length");
                    if (sym < 16)
                        lengths[i++] = (uint8_t)sym;
                    else
                    {
                        uint8_t value = 0;
                        uint32_t rep;
                        if (sym == 16)
                        {
                            if (i == 0)
                                return fail(err, "repeat with no previous length");
                            value = lengths[i - 1];
                            rep = 3 + in.bits(2);
                        }
                        else
                            rep = sym == 17 ? 3 + in.bits(3) : 11 + in.bits(7);
                        if (i + rep > nlit + ndist)
                            return fail(err, "code lengths overflow");
                        while (rep--)
                            lengths[i++] = value;
                    }
                    if (in.overrun)
                        return fail(err, "compressed data ends early");
                }
                if (lengths[256] == 0)
                    return fail(err, "no end-of-block code");
                if (nlit > 286 || ndist > 30)
                    return fail(err, "too many length or distance codes");
            }
            if (!lit.build(lengths, (int)nlit) || !dist.build(lengths + nlit, (int)ndist))
                return fail(err, "bad literal/length or distance code");
            for (;;)
            {
                const int sym = lit.decode(in);
                if (sym < 0 || in.overrun)
                    return fail(err, "bad literal/length code or data ends early");
                if (sym < 256)
                {
                    if (at >= expected)
                        return fail(err, "too much image data");
                    o[at++] = (uint8_t)sym;
                }
                else if (sym == 256)
                    break;
                else
                {
                    const int li = sym - 257;
                    if (li >= 29)
                        return fail(err, "bad length code");
                    const uint32_t len = kLengthBase[li] + in.bits(kLengthExtra[li]);
                    const int di = dist.decode(in);
                    if (di < 0 || di >= 30)
                        return fail(err, "bad distance code");
                    const uint32_t d = kDistBase[di] + in.bits(kDistExtra[di]);
                    if (d > at)
                        return fail(err, "distance before the start of the data");
                    if (at + len > expected)
                        return fail(err, "too much image data");
                    if (d >= len)
                        std::memcpy(o + at, o + at - d, len);
                    else
                        for (uint32_t k = 0; k < len; ++k)
                            o[at + k] = o[at + k - d]; // overlapping: repeats the last d bytes
                    at += len;
                }
            }
        }
        else
            return fail(err, "bad deflate block type");
        if (in.overrun)
            return fail(err, "compressed data ends early");
    }
    if (at != expected)
        return fail(err, "not enough image data");
    in.align();
    uint32_t adler = 0;
    for (int i = 0; i < 4; ++i)
        adler = adler << 8 | in.bits(8);
    if (in.overrun)
        return fail(err, "missing Adler-32");
    // Adler-32 in runs of 5552 bytes, the longest whose sums cannot overflow
    uint32_t a = 1, b = 0;
    for (size_t i = 0; i < expected;)
    {
        const size_t end = std::min(expected, i + 5552);
        for (; i < end; ++i)
        {
            a += o[i];
            b += a;
        }
        a %= 65521u;
        b %= 65521u;
    }
    if ((b << 16 | a) != adler)
        return fail(err, "Adler-32 mismatch");
    return true;
}

uint32_t crc32(const uint8_t* p, size_t n)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init)
    {
        for (uint32_t i = 0; i < 256; ++i)
        {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k)
                c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xffffffffu;
    for (size_t i = 0; i < n; ++i)
        c = table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c ^ 0xffffffffu;
}

// ------------------------------------------------------------------- PNG
const uint8_t kPngSignature[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
// libpng's default user limits (PNG_USER_WIDTH_MAX / PNG_USER_HEIGHT_MAX).
constexpr uint32_t kPngMaxSide = 1000000;

struct PngHeader
{
    uint32_t w = 0, h = 0;
    uint8_t depth = 0, ctype = 0, interlace = 0;
};

int png_channels(uint8_t ctype) { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }

bool png_header(const uint8_t* b, size_t n, PngHeader* h, std::string* err)
{
    if (n < 33 || std::memcmp(b, kPngSignature, 8) != 0)
        return fail(err, "not a PNG");
    if (be32(b + 8) != 13 || std::memcmp(b + 12, "IHDR", 4) != 0)
        return fail(err, "PNG does not start with IHDR");
    if (crc32(b + 12, 17) != be32(b + 29))
        return fail(err, "IHDR CRC mismatch");
    h->w = be32(b + 16);
    h->h = be32(b + 20);
    h->depth = b[24];
    h->ctype = b[25];
    h->interlace = b[28];
    if (h->w == 0 || h->h == 0 || h->w > kPngMaxSide || h->h > kPngMaxSide)
        return fail(err, "PNG dimensions out of range");
    const uint8_t d = h->depth;
    bool ok;
    switch (h->ctype)
    {
        case 0:
            ok = d == 1 || d == 2 || d == 4 || d == 8 || d == 16;
            break;
        case 3:
            ok = d == 1 || d == 2 || d == 4 || d == 8;
            break;
        case 2:
        case 4:
        case 6:
            ok = d == 8 || d == 16;
            break;
        default:
            ok = false;
    }
    if (!ok || b[26] != 0 || b[27] != 0 || h->interlace > 1)
        return fail(err, "bad PNG colour type, bit depth, compression, filter or interlace method");
    return true;
}

uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint8_t)(pa <= pb && pa <= pc ? a : pb <= pc ? b : c);
}

// Undoes the row filters of one (sub-)image in place; `rows` holds h rows of
// 1 + rowBytes bytes.
bool unfilter(uint8_t* rows, uint32_t h, size_t rowBytes, int bpp, std::string* err)
{
    uint8_t* prev = nullptr;
    for (uint32_t y = 0; y < h; ++y)
    {
        uint8_t* row = rows + (size_t)y * (rowBytes + 1);
        const uint8_t f = row[0];
        uint8_t* x = row + 1;
        const uint8_t* up = prev != nullptr ? prev + 1 : nullptr;
        switch (f)
        {
            case 0:
                break;
            case 1:
                for (size_t i = (size_t)bpp; i < rowBytes; ++i)
                    x[i] = (uint8_t)(x[i] + x[i - bpp]);
                break;
            case 2:
                if (up)
                    for (size_t i = 0; i < rowBytes; ++i)
                        x[i] = (uint8_t)(x[i] + up[i]);
                break;
            case 3:
                for (size_t i = 0; i < rowBytes; ++i)
                {
                    const int a = i >= (size_t)bpp ? x[i - bpp] : 0, b = up ? up[i] : 0;
                    x[i] = (uint8_t)(x[i] + ((a + b) >> 1));
                }
                break;
            case 4:
                for (size_t i = 0; i < rowBytes; ++i)
                {
                    const int a = i >= (size_t)bpp ? x[i - bpp] : 0, b = up ? up[i] : 0;
                    const int c = (i >= (size_t)bpp && up) ? up[i - bpp] : 0;
                    x[i] = (uint8_t)(x[i] + paeth(a, b, c));
                }
                break;
            default:
                return fail(err, "bad PNG row filter");
        }
        prev = row;
    }
    return true;
}

struct PngPass
{
    uint32_t x0, y0, dx, dy, w, h;
    size_t rowBytes;
};

std::vector<PngPass> png_passes(const PngHeader& hd)
{
    static const uint32_t adam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                         {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    std::vector<PngPass> passes;
    const size_t bitsPerPixel = (size_t)png_channels(hd.ctype) * hd.depth;
    if (hd.interlace == 0)
        passes.push_back({0, 0, 1, 1, hd.w, hd.h, (hd.w * bitsPerPixel + 7) / 8});
    else
        for (const auto& a : adam7)
        {
            const uint32_t w = hd.w > a[0] ? (hd.w - a[0] + a[2] - 1) / a[2] : 0;
            const uint32_t h = hd.h > a[1] ? (hd.h - a[1] + a[3] - 1) / a[3] : 0;
            if (w != 0 && h != 0)
                passes.push_back({a[0], a[1], a[2], a[3], w, h, (w * bitsPerPixel + 7) / 8});
        }
    return passes;
}
} // namespace

Format sniff(const uint8_t* b, size_t n)
{
    if (b == nullptr)
        return kUnknown;
    if (n >= 4 && b[0] == 0x89 && b[1] == 'P' && b[2] == 'N' && b[3] == 'G')
        return kPng;
    if (n >= 3 && b[0] == 0xff && b[1] == 0xd8 && b[2] == 0xff)
        return kJpeg;
    if (n >= 12 && std::memcmp(b, "RIFF", 4) == 0 && std::memcmp(b + 8, "WEBP", 4) == 0)
        return kWebp;
    return kUnknown;
}

bool decode_png(const uint8_t* b, size_t n, uint32_t* width, uint32_t* height, std::vector<uint8_t>* rgba,
                std::string* err)
{
    PngHeader hd;
    if (!png_header(b, n, &hd, err))
        return false;
    const int ch = png_channels(hd.ctype);
    std::vector<uint8_t> idat;
    uint8_t palette[256][4];
    uint32_t paletteCount = 0;
    bool hasTrns = false;
    uint16_t trnsKey[3] = {0, 0, 0};
    for (auto& e : palette)
        e[0] = e[1] = e[2] = 0, e[3] = 255;
    size_t pos = 33;
    bool sawIend = false;
    while (pos + 12 <= n)
    {
        const uint32_t len = be32(b + pos);
        if (len > 0x7fffffffu || pos + 12 + (size_t)len > n)
            return fail(err, "PNG chunk runs past the end of the data");
        const uint8_t* type = b + pos + 4;
        const uint8_t* data = b + pos + 8;
        const bool critical = (type[0] & 0x20) == 0;
        const bool crcOk = crc32(type, 4 + (size_t)len) == be32(data + len);
        if (!crcOk && critical)
            return fail(err, "PNG chunk CRC mismatch");
        pos += 12 + (size_t)len;
        if (!crcOk)
            continue; // a damaged ancillary chunk is dropped, as libpng does
        if (std::memcmp(type, "IDAT", 4) == 0)
            idat.insert(idat.end(), data, data + len);
        else if (std::memcmp(type, "PLTE", 4) == 0)
        {
            if (len % 3 != 0 || len == 0 || len > 768)
                return fail(err, "bad PLTE");
            paletteCount = len / 3;
            for (uint32_t i = 0; i < paletteCount; ++i)
                palette[i][0] = data[3 * i], palette[i][1] = data[3 * i + 1], palette[i][2] = data[3 * i + 2];
        }
        else if (std::memcmp(type, "tRNS", 4) == 0)
        {
            if (hd.ctype == 3)
            {
                for (uint32_t i = 0; i < len && i < 256; ++i)
                    palette[i][3] = data[i];
                hasTrns = len > 0;
            }
            else if (hd.ctype == 0 && len >= 2)
                trnsKey[0] = (uint16_t)be16(data), hasTrns = true;
            else if (hd.ctype == 2 && len >= 6)
                trnsKey[0] = (uint16_t)be16(data), trnsKey[1] = (uint16_t)be16(data + 2),
                trnsKey[2] = (uint16_t)be16(data + 4), hasTrns = true;
        }
        else if (std::memcmp(type, "IEND", 4) == 0)
        {
            sawIend = true;
            break;
        }
        else if (critical && std::memcmp(type, "IHDR", 4) != 0)
            return fail(err, "unknown critical PNG chunk");
    }
    (void)sawIend; // libpng reads the image without needing IEND
    if (hd.ctype == 3 && paletteCount == 0)
        return fail(err, "palette image without PLTE");
    const std::vector<PngPass> passes = png_passes(hd);
    size_t raw = 0;
    for (const PngPass& p : passes)
        raw += (size_t)p.h * (p.rowBytes + 1);
    // Deflate expands at most 1032:1 (a 258-byte match in two bits), so a
    // stream this short cannot hold the image: refuse before allocating it.
    if ((double)idat.size() * 1032.0 + 64.0 < (double)raw)
        return fail(err, "not enough image data for the PNG's dimensions");
    std::vector<uint8_t> filtered;
    if (!zlib_inflate(idat.data(), idat.size(), raw, &filtered, err))
        return false;
    const int bpp = std::max(1, ch * hd.depth / 8);
    rgba->assign((size_t)hd.w * hd.h * 4, 0);
    size_t at = 0;
    const int depth = hd.depth;
    for (const PngPass& p : passes)
    {
        uint8_t* rows = filtered.data() + at;
        if (!unfilter(rows, p.h, p.rowBytes, bpp, err))
            return false;
        for (uint32_t y = 0; y < p.h; ++y)
        {
            const uint8_t* row = rows + (size_t)y * (p.rowBytes + 1) + 1;
            for (uint32_t x = 0; x < p.w; ++x)
            {
                // samples at full precision, then to 8 bits
                uint32_t s[4] = {0, 0, 0, 0};
                for (int c = 0; c < ch; ++c)
                {
                    const size_t i = (size_t)x * ch + c;
                    if (depth == 16)
                        s[c] = be16(row + 2 * i);
                    else if (depth == 8)
                        s[c] = row[i];
                    else
                    {
                        const size_t bit = i * depth;
                        s[c] = (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1);
                    }
                }
                auto to8 = [&](uint32_t v) -> uint8_t {
                    if (depth == 16)
                        return (uint8_t)(v >> 8);
                    if (depth == 8)
                        return (uint8_t)v;
                    return (uint8_t)(v * (255u / ((1u << depth) - 1)));
                };
                uint8_t o[4];
                switch (hd.ctype)
                {
                    case 0:
                        o[0] = o[1] = o[2] = to8(s[0]);
                        o[3] = hasTrns && s[0] == trnsKey[0] ? 0 : 255;
                        break;
                    case 2:
                        o[0] = to8(s[0]), o[1] = to8(s[1]), o[2] = to8(s[2]);
                        o[3] = hasTrns && s[0] == trnsKey[0] && s[1] == trnsKey[1] && s[2] == trnsKey[2] ? 0 : 255;
                        break;
                    case 3:
                        if (s[0] >= paletteCount)
                            return fail(err, "palette index out of range");
                        std::memcpy(o, palette[s[0]], 4);
                        break;
                    case 4:
                        o[0] = o[1] = o[2] = to8(s[0]);
                        o[3] = to8(s[1]);
                        break;
                    default:
                        o[0] = to8(s[0]), o[1] = to8(s[1]), o[2] = to8(s[2]), o[3] = to8(s[3]);
                }
                const size_t px = (size_t)(p.y0 + y * p.dy) * hd.w + (p.x0 + x * p.dx);
                std::memcpy(rgba->data() + 4 * px, o, 4);
            }
        }
        at += (size_t)p.h * (p.rowBytes + 1);
    }
    *width = hd.w;
    *height = hd.h;
    return true;
}

// ------------------------------------------------------------------ JPEG
namespace
{
// Zigzag position k -> natural index, with 16 extra entries so that a
// corrupt run past the end lands on the last coefficient (as libjpeg's
// jpeg_natural_order does).
const uint8_t kNatural[64 + 16] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
    6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
    39, 46, 53, 60, 61, 54, 47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct JpegHuffman
{
    bool defined = false;
    uint8_t bits[17] = {};
    uint8_t vals[256] = {};
    int32_t maxcode[18];
    int32_t valptr[17];
    int32_t mincode[17];
    uint16_t fast[1 << 9]; // (value << 4 | length) for codes of up to 9 bits, else 0

    bool build()
    {
        int k = 0, code = 0;
        int total = 0;
        for (int l = 1; l <= 16; ++l)
            total += bits[l];
        if (total > 256)
            return false;
        std::memset(fast, 0, sizeof(fast));
        for (int l = 1; l <= 16; ++l)
        {
            valptr[l] = k;
            mincode[l] = code;
            for (int i = 0; i < bits[l]; ++i, ++k, ++code)
                if (l <= 9)
                    for (int f = code << (9 - l); f < (code + 1) << (9 - l); ++f)
                        fast[f] = (uint16_t)(vals[k] << 4 | l);
            maxcode[l] = bits[l] ? code - 1 : -1;
            if (code > (1 << l))
                return false;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        return true;
    }
};

// The entropy-coded segment reader: bytes with 0xFF00 stuffing, stopping at
// a marker.  Past it (or past the end) it supplies zero bits, and when a
// code or value actually consumes one of them it records that the data ran
// short (libjpeg's insufficient_data, jdhuff.c jpeg_fill_bit_buffer).
struct JpegBits
{
    const uint8_t* p;
    size_t n, pos;
    uint32_t buf = 0;
    int cnt = 0;      // bits in buf
    int real = 0;     // of them, bits read from the data (they come first)
    bool atMarker = false;
    bool ranShort = false;

    void fill()
    {
        while (cnt <= 24)
        {
            uint32_t byte = 0;
            bool fromData = false;
            if (!atMarker && pos < n)
            {
                byte = p[pos];
                if (byte == 0xff)
                {
                    size_t q = pos + 1;
                    while (q < n && p[q] == 0xff)
                        ++q;
                    if (q < n && p[q] == 0x00)
                    {
                        pos = q + 1;
                        fromData = true;
                    }
                    else
                    {
                        atMarker = true; // pos stays on the marker
                        byte = 0;
                    }
                }
                else
                {
                    ++pos;
                    fromData = true;
                }
            }
            else
                atMarker = true;
            buf |= byte << (24 - cnt);
            cnt += 8;
            if (fromData)
                real += 8;
        }
    }
    void consume(int k)
    {
        if (k > real)
            ranShort = true;
        real = std::max(real - k, 0);
        buf <<= k;
        cnt -= k;
    }
    uint32_t bits(int k)
    {
        if (k == 0)
            return 0;
        if (cnt < k)
            fill();
        const uint32_t v = buf >> (32 - k);
        consume(k);
        return v;
    }
    int decode(const JpegHuffman& h)
    {
        if (cnt < 16)
            fill();
        const uint16_t e = h.fast[buf >> (32 - 9)];
        if (e != 0)
        {
            consume(e & 15);
            return e >> 4;
        }
        int code = (int)bits(1);
        int l = 1;
        while (code > h.maxcode[l])
        {
            code = (code << 1) | (int)bits(1);
            if (++l > 16)
                return 0; // corrupt: libjpeg warns and decodes a zero
        }
        return h.vals[h.valptr[l] + code - h.mincode[l]];
    }
    // Drops the buffered bits and moves past the next marker when it is a
    // restart marker (RST0..7); the data then counts as whole again.
    void restart()
    {
        buf = 0;
        cnt = 0;
        real = 0;
        atMarker = false;
        while (pos + 1 < n && !(p[pos] == 0xff && p[pos + 1] != 0x00 && p[pos + 1] != 0xff))
            ++pos;
        if (pos + 1 < n && p[pos + 1] >= 0xd0 && p[pos + 1] <= 0xd7)
        {
            pos += 2;
            ranShort = false;
        }
    }
};

int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }
} // namespace

bool decode_jpeg(const uint8_t* b, size_t n, Jpeg* j, bool coefficients, std::string* err)
{
    if (sniff(b, n) != kJpeg)
        return fail(err, "not a JPEG");
    *j = Jpeg{};
    JpegHuffman dc[4], ac[4];
    bool qtDefined[4] = {false, false, false, false};
    bool progressive = false;
    bool sawFrame = false, sawJfif = false, sawAdobe = false;
    int adobeTransform = -1;
    uint32_t restartInterval = 0;
    size_t pos = 2;
    bool sawScan = false;
    for (;;)
    {
        // next marker
        while (pos < n && b[pos] != 0xff)
            ++pos;
        while (pos < n && b[pos] == 0xff)
            ++pos;
        if (pos >= n)
        {
            if (sawScan)
                break; // a file ending without EOI keeps what was decoded
            return fail(err, "JPEG ends before its image data");
        }
        const uint8_t m = b[pos++];
        if (m == 0xd9)
            break; // EOI
        if (m == 0x01 || (m >= 0xd0 && m <= 0xd7))
            continue;
        if (pos + 2 > n)
            return fail(err, "JPEG marker segment runs past the end");
        const size_t len = be16(b + pos);
        if (len < 2 || pos + len > n)
            return fail(err, "JPEG marker segment runs past the end");
        const uint8_t* s = b + pos + 2;
        const size_t sl = len - 2;
        pos += len;
        if (m == 0xc0 || m == 0xc1 || m == 0xc2)
        {
            progressive = m == 0xc2;
            if (sawFrame || sl < 6)
                return fail(err, "bad or repeated JPEG frame header");
            if (s[0] != 8)
                return fail(err, "only 8-bit JPEG samples are supported");
            j->height = be16(s + 1);
            j->width = be16(s + 3);
            const int nc = s[5];
            if (j->height == 0)
                return fail(err, "JPEG height from a DNL marker is not supported");
            if (j->width == 0 || nc < 1 || nc > 4 || sl < (size_t)6 + 3 * nc)
                return fail(err, "bad JPEG frame header");
            for (int c = 0; c < nc; ++c)
            {
                JpegComponent jc{};
                jc.id = s[6 + 3 * c];
                jc.h = s[7 + 3 * c] >> 4;
                jc.v = s[7 + 3 * c] & 15;
                jc.tq = s[8 + 3 * c];
                if (jc.h < 1 || jc.h > 4 || jc.v < 1 || jc.v > 4 || jc.tq > 3)
                    return fail(err, "bad JPEG component sampling or table");
                j->comps.push_back(jc);
            }
            sawFrame = true;
        }
        else if ((m >= 0xc3 && m <= 0xcb && m != 0xc4 && m != 0xc8) || (m >= 0xcd && m <= 0xcf))
            return fail(err, "lossless, hierarchical and arithmetic-coded JPEGs are not supported");
        else if (m == 0xc4)
        {
            size_t q = 0;
            while (q < sl)
            {
                if (q + 17 > sl)
                    return fail(err, "bad JPEG Huffman table");
                const int tc = s[q] >> 4, th = s[q] & 15;
                if (tc > 1 || th > 3)
                    return fail(err, "bad JPEG Huffman table class or ID");
                JpegHuffman& h = tc == 0 ? dc[th] : ac[th];
                int total = 0;
                h.bits[0] = 0;
                for (int l = 1; l <= 16; ++l)
                    total += (h.bits[l] = s[q + l]);
                if (total > 256 || q + 17 + total > sl)
                    return fail(err, "bad JPEG Huffman table");
                std::memcpy(h.vals, s + q + 17, total);
                if (!h.build())
                    return fail(err, "bad JPEG Huffman table");
                h.defined = true;
                q += 17 + total;
            }
        }
        else if (m == 0xdb)
        {
            size_t q = 0;
            while (q < sl)
            {
                const int pq = s[q] >> 4, tq = s[q] & 15;
                if (pq > 1 || tq > 3 || q + 1 + 64 * (pq + 1) > sl)
                    return fail(err, "bad JPEG quantization table");
                for (int k = 0; k < 64; ++k)
                    j->qt[tq][kNatural[k]] =
                        (uint16_t)(pq ? be16(s + q + 1 + 2 * k) : s[q + 1 + k]);
                qtDefined[tq] = true;
                q += 1 + 64 * (pq + 1);
            }
        }
        else if (m == 0xdd)
        {
            if (sl < 2)
                return fail(err, "bad JPEG restart interval");
            restartInterval = be16(s);
        }
        else if (m == 0xe0)
            sawJfif = sawJfif || (sl >= 5 && std::memcmp(s, "JFIF\0", 5) == 0);
        else if (m == 0xee)
        {
            if (sl >= 12 && std::memcmp(s, "Adobe", 5) == 0)
            {
                sawAdobe = true;
                adobeTransform = s[11];
            }
        }
        else if (m == 0xda)
        {
            if (!sawFrame)
                return fail(err, "JPEG scan before its frame header");
            if (!sawScan)
            {
                // the frame's geometry and colour space (jdinput.c,
                // jdapimin.c default_decompress_parms, jdmaster.c)
                for (const JpegComponent& c : j->comps)
                    j->hmax = std::max(j->hmax, c.h), j->vmax = std::max(j->vmax, c.v);
                const uint32_t mcusX = (j->width + 8u * j->hmax - 1) / (8u * j->hmax);
                const uint32_t mcusY = (j->height + 8u * j->vmax - 1) / (8u * j->vmax);
                size_t total = 0;
                for (JpegComponent& c : j->comps)
                {
                    if (!qtDefined[c.tq])
                        return fail(err, "JPEG component uses an undefined quantization table");
                    c.blocksW = mcusX * c.h;
                    c.blocksH = mcusY * c.v;
                    c.coefOffset = total;
                    total += (size_t)c.blocksW * c.blocksH * 64;
                    // chroma upsampled by the inverse DCT where the ratio
                    // allows (jdmaster.c, jpeg 7+), the rest by replication
                    c.scaleH = (j->hmax % (c.h * 2) == 0) ? 2 : 1;
                    c.scaleV = (j->vmax % (c.v * 2) == 0) ? 2 : 1;
                    if (j->hmax % (c.h * c.scaleH) != 0 || j->vmax % (c.v * c.scaleV) != 0)
                        return fail(err, "fractional JPEG sampling ratios are not supported");
                    c.repH = (uint8_t)(j->hmax / (c.h * c.scaleH));
                    c.repV = (uint8_t)(j->vmax / (c.v * c.scaleV));
                }
                const size_t nc = j->comps.size();
                if (nc == 2 || nc == 4)
                    return fail(err, "JPEG colour space cannot be converted to RGB");
                if (nc == 3)
                {
                    const int c0 = j->comps[0].id, c1 = j->comps[1].id, c2 = j->comps[2].id;
                    if (c0 == 1 && c1 == 2 && c2 == 3)
                        j->ycc = true;
                    else if (c0 == 1 && c1 == 0x22 && c2 == 0x23)
                        return fail(err, "big-gamut YCC JPEGs are not supported");
                    else if (c0 == 'R' && c1 == 'G' && c2 == 'B')
                        j->ycc = false;
                    else if (c0 == 'r' && c1 == 'g' && c2 == 'b')
                        return fail(err, "big-gamut RGB JPEGs are not supported");
                    else if (sawJfif)
                        j->ycc = true;
                    else if (sawAdobe)
                        j->ycc = adobeTransform != 0;
                    else
                        j->ycc = true;
                }
                sawScan = true;
                if (!coefficients)
                {
                    j->coefCount = total;
                    break; // the header and layout are all a probe needs
                }
                j->coefCount = total;
                j->coef.assign(total, 0);
            }
            // ---- one scan
            if (sl < 1)
                return fail(err, "bad JPEG scan header");
            const int ns = s[0];
            if (ns < 1 || ns > 4 || sl < (size_t)1 + 2 * ns + 3)
                return fail(err, "bad JPEG scan header");
            int scomp[4], td[4], ta[4];
            for (int i = 0; i < ns; ++i)
            {
                const int id = s[1 + 2 * i];
                scomp[i] = -1;
                for (size_t c = 0; c < j->comps.size(); ++c)
                    if (j->comps[c].id == id)
                        scomp[i] = (int)c;
                if (scomp[i] < 0)
                    return fail(err, "JPEG scan names an unknown component");
                td[i] = s[2 + 2 * i] >> 4;
                ta[i] = s[2 + 2 * i] & 15;
            }
            const int ss = s[1 + 2 * ns], se = s[2 + 2 * ns], ah = s[3 + 2 * ns] >> 4, al = s[3 + 2 * ns] & 15;
            if (!progressive && (ss != 0 || se != 63 || ah != 0 || al != 0))
                return fail(err, "bad sequential JPEG scan parameters");
            // progressive scans (T.81 G.1): a DC scan (Ss = Se = 0) or one
            // component's band of AC coefficients, each a first pass or a
            // refinement of one more bit
            if (progressive && (se < ss || se > 63 || (ss == 0 && se != 0) || (ss != 0 && ns != 1) || al > 13 ||
                                ah > 13 || (ah != 0 && ah != al + 1)))
                return fail(err, "bad progressive JPEG scan parameters");
            const bool needDc = ss == 0, needAc = se > 0;
            for (int i = 0; i < ns; ++i)
                if (td[i] > 3 || ta[i] > 3 || (needDc && ah == 0 && !dc[td[i]].defined) ||
                    (needAc && !ac[ta[i]].defined))
                    return fail(err, "JPEG scan uses an undefined Huffman table");
            // MCU geometry: interleaved scans cover the MCU grid, a
            // single-component scan the component's own block grid
            uint32_t mcusX, mcusY;
            if (ns == 1)
            {
                const JpegComponent& c = j->comps[scomp[0]];
                const uint32_t cw = (j->width * c.h + j->hmax - 1) / j->hmax;
                const uint32_t chh = (j->height * c.v + j->vmax - 1) / j->vmax;
                mcusX = (cw + 7) / 8;
                mcusY = (chh + 7) / 8;
            }
            else
            {
                mcusX = (j->width + 8u * j->hmax - 1) / (8u * j->hmax);
                mcusY = (j->height + 8u * j->vmax - 1) / (8u * j->vmax);
            }
            JpegBits in{b, n, pos};
            int pred[4] = {0, 0, 0, 0};
            uint32_t eobrun = 0; // progressive AC: blocks left in the current end-of-band run
            const int p1 = 1 << al, m1 = -(1 << al);
            // one block of a scan, by scan type (libjpeg jdhuff.c / jdphuff.c)
            auto refine = [&](int16_t* coefp) {
                if (*coefp != 0 && in.bits(1) != 0 && (*coefp & p1) == 0)
                    *coefp = (int16_t)(*coefp + (*coefp >= 0 ? p1 : m1));
            };
            auto block = [&](int16_t* blk, int i) {
                if (!progressive)
                {
                    const int t = in.decode(dc[td[i]]);
                    pred[i] += extend((int)in.bits(t & 15), t & 15);
                    blk[0] = (int16_t)pred[i];
                    for (int k = 1; k < 64;)
                    {
                        const int rs = in.decode(ac[ta[i]]);
                        const int r = rs >> 4, sz = rs & 15;
                        if (sz != 0)
                        {
                            k += r;
                            blk[kNatural[k]] = (int16_t)extend((int)in.bits(sz), sz);
                            ++k;
                        }
                        else if (r == 15)
                            k += 16;
                        else
                            break;
                    }
                }
                else if (ss == 0 && ah == 0) // DC first pass
                {
                    const int t = in.decode(dc[td[i]]);
                    pred[i] += extend((int)in.bits(t & 15), t & 15);
                    blk[0] = (int16_t)(pred[i] * (1 << al));
                }
                else if (ss == 0) // DC refinement: one more bit
                {
                    if (in.bits(1) != 0)
                        blk[0] = (int16_t)(blk[0] | p1);
                }
                else if (ah == 0) // AC first pass
                {
                    if (eobrun > 0)
                    {
                        --eobrun;
                        return;
                    }
                    for (int k = ss; k <= se; ++k)
                    {
                        const int rs = in.decode(ac[ta[i]]);
                        const int r = rs >> 4, sz = rs & 15;
                        if (sz != 0)
                        {
                            k += r;
                            blk[kNatural[k]] = (int16_t)(extend((int)in.bits(sz), sz) * (1 << al));
                        }
                        else if (r == 15)
                            k += 15;
                        else
                        {
                            eobrun = (1u << r) - 1;
                            if (r != 0)
                                eobrun += in.bits(r);
                            break;
                        }
                    }
                }
                else // AC refinement (T.81 G.1.2.3)
                {
                    int k = ss;
                    if (eobrun == 0)
                    {
                        for (; k <= se; ++k)
                        {
                            const int rs = in.decode(ac[ta[i]]);
                            int r = rs >> 4;
                            int value = 0;
                            if ((rs & 15) != 0)
                                value = in.bits(1) != 0 ? p1 : m1; // a newly nonzero coefficient
                            else if (r != 15)
                            {
                                eobrun = 1u << r;
                                if (r != 0)
                                    eobrun += in.bits(r);
                                break;
                            }
                            // pass r zero-history coefficients, refining the
                            // nonzero ones on the way
                            do
                            {
                                int16_t* c = blk + kNatural[k];
                                if (*c != 0)
                                    refine(c);
                                else if (--r < 0)
                                    break;
                                ++k;
                            } while (k <= se);
                            if (value != 0)
                                blk[kNatural[k]] = (int16_t)value;
                        }
                    }
                    if (eobrun > 0)
                    {
                        for (; k <= se; ++k)
                            refine(blk + kNatural[k]);
                        --eobrun;
                    }
                }
            };
            uint32_t todo = restartInterval;
            bool skip = false; // data ran short: the rest of the interval stays zero
            for (uint32_t my = 0; my < mcusY; ++my)
                for (uint32_t mx = 0; mx < mcusX; ++mx)
                {
                    if (restartInterval != 0)
                    {
                        if (todo == 0)
                        {
                            in.restart();
                            pred[0] = pred[1] = pred[2] = pred[3] = 0;
                            eobrun = 0;
                            todo = restartInterval;
                            skip = in.ranShort;
                        }
                        --todo;
                    }
                    if (skip)
                        continue;
                    for (int i = 0; i < ns; ++i)
                    {
                        const JpegComponent& c = j->comps[scomp[i]];
                        const int bh = ns == 1 ? 1 : c.h, bv = ns == 1 ? 1 : c.v;
                        for (int v = 0; v < bv; ++v)
                            for (int h = 0; h < bh; ++h)
                            {
                                const uint32_t bx = mx * bh + h, by = my * bv + v;
                                block(j->coef.data() + c.coefOffset + ((size_t)by * c.blocksW + bx) * 64, i);
                            }
                    }
                    if (in.ranShort)
                        skip = true;
                }
            // continue after the entropy-coded data
            pos = in.pos;
            while (pos + 1 < n && !(b[pos] == 0xff && b[pos + 1] != 0x00 && !(b[pos + 1] >= 0xd0 && b[pos + 1] <= 0xd7) &&
                                    b[pos + 1] != 0xff))
                ++pos;
        }
        // other markers (APPn, COM, ...) are skipped
    }
    if (!sawFrame)
        return fail(err, "JPEG without a frame header");
    if (!sawScan)
        return fail(err, "JPEG without image data");
    return true;
}

bool probe(const uint8_t* bytes, size_t n, Format* format, uint32_t* width, uint32_t* height, std::string* err)
{
    *format = sniff(bytes, n);
    *width = *height = 0;
    switch (*format)
    {
        case kPng:
        {
            PngHeader h;
            if (!png_header(bytes, n, &h, err))
                return false;
            *width = h.w;
            *height = h.h;
            return true;
        }
        case kJpeg:
        {
            Jpeg j;
            if (!decode_jpeg(bytes, n, &j, false, err))
                return false;
            *width = j.width;
            *height = j.height;
            return true;
        }
        case kWebp:
            return fail(err, "WebP is recognised but its decoder is not built (the reference's no_rive_webp "
                             "configuration)");
        default:
            return fail(err, "unrecognised image format");
    }
}
} // namespace rivehip::image
