// Shared device-side definitions of the gfx950 backend.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rivehip
{
// ---- constants of the reference's GPU data contract
//      (renderer/src/shaders/constants.glsl, renderer/include/rive/renderer/gpu.hpp)
constexpr uint32_t kTessTextureWidth = 2048;
constexpr uint32_t kTessTextureWidthLog2 = 11;
constexpr uint32_t kPatchSpan = 8; // kMidpointFanPatchSegmentSpan
constexpr uint32_t kGradWidth = 512;

constexpr uint32_t CONTOUR_ID_MASK = 0xffffu;
constexpr uint32_t RETROFIT_TRI_STRIP_CONTOUR_FLAG = 1u << 31;
constexpr uint32_t CULL_EXCESS_TESSELLATION_SEGMENTS_CONTOUR_FLAG = 1u << 29;
constexpr uint32_t JOIN_TYPE_MASK = 7u << 26;
constexpr uint32_t MITER_CLIP_JOIN_CONTOUR_FLAG = 5u << 26;
constexpr uint32_t MITER_REVERT_JOIN_CONTOUR_FLAG = 4u << 26;
constexpr uint32_t BEVEL_JOIN_CONTOUR_FLAG = 3u << 26;
constexpr uint32_t ROUND_JOIN_CONTOUR_FLAG = 2u << 26;
constexpr uint32_t FEATHER_JOIN_CONTOUR_FLAG = 1u << 26;
constexpr uint32_t EMULATED_STROKE_CAP_CONTOUR_FLAG = 1u << 25;
constexpr uint32_t NEGATE_PATH_FILL_COVERAGE_FLAG = 1u << 24;
constexpr uint32_t MIRRORED_CONTOUR_CONTOUR_FLAG = 1u << 23;
constexpr uint32_t JOIN_TANGENT_0_CONTOUR_FLAG = 1u << 22;
constexpr uint32_t JOIN_TANGENT_INNER_CONTOUR_FLAG = 1u << 21;
constexpr uint32_t LEFT_JOIN_CONTOUR_FLAG = 1u << 20;
constexpr uint32_t RIGHT_JOIN_CONTOUR_FLAG = 1u << 19;

constexpr uint32_t CLIP_UPDATE_PAINT_TYPE = 0u;
constexpr uint32_t SOLID_COLOR_PAINT_TYPE = 1u;
constexpr uint32_t LINEAR_GRADIENT_PAINT_TYPE = 2u;
constexpr uint32_t RADIAL_GRADIENT_PAINT_TYPE = 3u;
constexpr uint32_t PAINT_FLAG_NON_ZERO_FILL = 0x100u;
constexpr uint32_t PAINT_FLAG_EVEN_ODD_FILL = 0x200u;
constexpr uint32_t PAINT_FLAG_HAS_CLIP_RECT = 0x400u;

constexpr uint32_t GRAD_SPAN_FLAG_LEFT_BORDER = 0x80000000u;
constexpr uint32_t GRAD_SPAN_FLAG_RIGHT_BORDER = 0x40000000u;
constexpr uint32_t GRAD_SPAN_FLAG_COMPLEX_BORDER = 0x20000000u;
constexpr uint32_t GRAD_SPAN_FLAGS_MASK = 0xe0000000u;

// Raster tiles: 16x16 pixels, one pixel per lane, four wave64s per tile.
constexpr int kTileLog2 = 4;
constexpr int kTile = 1 << kTileLog2;
// Screen positions are snapped to 1/256 pixel for the exact raster.
constexpr int kSubpixelLog2 = 8;

struct TessTexel
{
    float x, y;
    uint32_t z; // theta bits (radians, [0, 2pi))
    uint32_t w; // contourIDWithFlags
};

// One tessellated segment expanded to screen space, fixed point (1/256 px).
//   fill:   p = [I0, O0, I1, O1] (inner/outer AA points), cov unused
//   stroke: p = [P0-, P0+, P1-, P1+], cov = (x, y) per point
struct Prim
{
    int32_t p[8];
    float cov[8];
    uint32_t path;  // 1-based path ID
    uint32_t flags; // kPrimFill / kPrimStroke / kPrimNegate / kPrimEmpty
    uint32_t pad[2];
};
static_assert(sizeof(Prim) == 80, "Prim is 80 bytes");
constexpr uint32_t kPrimFill = 1u;
constexpr uint32_t kPrimStroke = 2u;
constexpr uint32_t kPrimNegate = 4u;
constexpr uint32_t kPrimEmpty = 8u;
// Triangle 1 / 2 of the primitive is dropped (a vertex of it had no finite
// position); the other and, for fills, the inner edge stay.
constexpr uint32_t kPrimNoT1 = 16u;
constexpr uint32_t kPrimNoT2 = 32u;
// The primitive leaves the guard box; binning skips it and its pieces inside
// the box are appended after the flush's primitives.
constexpr uint32_t kPrimFar = 64u;
// A fill primitive without its inner edge (its contour has no finite fan
// midpoint): only its border triangles draw.
constexpr uint32_t kPrimNoEdge = 128u;
// Set aside by the expansion for the pass that handles non-finite vertices
// and vertices outside the guard box.
constexpr uint32_t kPrimSpecial = 256u;

// A feathered fill's segment (draw_path_common.glsl:107-258): p = [E0, T0,
// E1, T1], the edge points (outset 0) and the outer points (outset +1) of both
// ends; the inner points (outset -1) are B = 2E - T.  Its four border
// triangles (E0,T0,E1), (E1,T0,T1), (E0,B0,E1), (E1,B0,B1) add the feathered
// coverage of their interpolated coverage vectors; the edge E0 -> E1 carries
// the fan's winding like a fill's inner edge.  An end that is a feather-join
// spoke stores its coverage vectors (kPrimJoin0/1): cov[0..3] the outer
// vector of end 0 and pad[0] the y0 of its edge vector, cov[4..7] and pad[1]
// those of end 1 (an edge vector of a spoke is (.75, -2.5, cotTheta, y0), its
// cotTheta that of the outer vector).  Other ends take the flat-edge vectors
// (.5, -2, 1e6, .5) on the edge and (0, -2, 1e6, 0) outside.
constexpr uint32_t kPrimFeather = 512u;
constexpr uint32_t kPrimJoin0 = 1024u;
constexpr uint32_t kPrimJoin1 = 2048u;
// A feathered stroke's primitive: its coverage x is biased by -2 (x < -1.5
// marks it) and evaluates to 1 - FEATHER(1 - x) - FEATHER(1 - y).
constexpr uint32_t kPrimFeatherStroke = 4096u;

// The Gaussian-integral function tables of feathering (gpu.cpp:1991-2181):
// 512 fp16 texels each, held as floats, FEATHER first, INVERSE_FEATHER
// second.
constexpr uint32_t kFeatherTableSize = 512;

// A table sampled like a 1D texture with linear filtering and clamp-to-edge
// at normalized coordinate x (the FEATHER macro, common.glsl:53-68).
__device__ __forceinline__ float feather_table(const float* __restrict__ t, float x)
{
    x = fminf(fmaxf(x, 0.f), 1.f); // NaN -> 0
    const float u = x * (float)kFeatherTableSize - .5f;
    const int r = (int)fminf(u + 1.f, (float)(kFeatherTableSize - 1));
    const int l = max(r - 1, 0);
    const float w = fminf(fmaxf(u - (float)l, 0.f), 1.f);
    return t[l] * (1.f - w) + t[r] * w;
}

// Per-path raster parameters, unpacked from PathData/PaintData/PaintAux.
// The fields the src-over resolve reads come first.
struct PathRaster
{
    uint32_t flags;      // bit0 stroke, bit1 evenOdd, bit2 clockwise, bit3 clipUpdate, bit4 clip rect,
                         // bits 8-11 blend mode
    uint32_t paintType;
    uint32_t clipID;     // draws: clip to read (0 none); clip updates: ID to write
    uint32_t outerClipID;
    float color255[4];   // premultiplied solid color, 0..255 scale
    float alpha;         // solid color alpha, 0..1
    float gradSpan;      // > 1: complex row; else simple-ramp texel offset
    uint32_t gradRow;
    uint32_t image;      // 0 none, else 1-based index into the flush's image table
    float paintMatrix[6];
    uint32_t imageSampler; // ImageSampler key: wrapX + 3*wrapY + 9*filter
    float imageMatrix[6];  // pixel -> unit uv (PaintAuxData image matrix)
    float imageLod;
    float clipRect[6];     // pixel -> the clip rect normalized to [-1, 1]^2 (ClipRectInverseMatrix)
    uint32_t pad;
};
static_assert(sizeof(PathRaster) == 132, "PathRaster layout");

// One image of the flush's image table: RGBA8 premultiplied, mip levels
// packed one after another.
constexpr int kMaxImageLevels = 15;
struct ImageDesc
{
    const uint32_t* texels;
    uint32_t width, height, levels, pad;
    uint32_t levelOffset[kMaxImageLevels + 1];
};

// Per-path tile-cell grid (bounding box in tiles).
struct PathCells
{
    int32_t tx0, ty0, tx1, ty1; // [tx0, tx1) x [ty0, ty1); empty if tx0 >= tx1
    uint32_t base;              // first cell index
    uint32_t width;             // tx1 - tx0
    uint32_t pad[2];
};

// One canvas of a flush: its target, tile grid and path range.  Tiles of all
// canvases form one launch; canvas c owns tiles [tileBase, tileBase + tilesX*tilesY).
struct CanvasDev
{
    uint32_t* pixels;
    int32_t width, height, tilesX, tilesY;
    uint32_t tileBase;
    uint32_t firstPath, pathEnd; // paths [firstPath, pathEnd)
    uint32_t clearPremul, loadPreserve;
    uint32_t pad;
};
static_assert(sizeof(CanvasDev) == 48, "CanvasDev layout");

// One draw's tessellation range (the reference's DrawBatch, gpu.hpp).
struct DrawTess
{
    uint32_t pathID;
    uint32_t tessLocation;    // first texel of the draw's allocation
    uint32_t tessVertexCount; // both directions
    uint32_t dirs;            // ContourDirections
    uint32_t isStroke;
    uint32_t forwardBase;     // first forward texel
    uint32_t forwardCount;    // forward texels
    uint32_t primBase;        // first primitive slot
};
} // namespace rivehip
