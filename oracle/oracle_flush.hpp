// CPU ORACLE (test infrastructure only): one logical flush as the reference's
// RenderContext lays it out (render_context.cpp:1116-1568), plus the raster.
#pragma once

#include "oracle.hpp"
#include "../include/rive_hip.h"

#include <memory>
#include <vector>

namespace oracle
{
struct DrawInfo
{
    uint32_t pathID;
    uint32_t tessLocation;
    uint32_t tessVertexCount; // both directions
    Dirs dirs;
    bool isStroke;
    bool featheredFill = false;
    bool outerCubics = false; // laid out as outer-cubic patches (interior triangulation)
    // A feather drawn through the atlas (draw.cpp:387-418): its coverage is
    // rendered into its own atlas region (aw x ah texels, atlas coordinate =
    // pixel * atlasScale + atlasTranslate) and blitted over the pixel bounds
    // [blitL, blitR) x [blitT, blitB) with bilinear filtering.
    bool atlas = false;
    float atlasScale = 0, atlasTx = 0, atlasTy = 0;
    uint32_t aw = 0, ah = 0;
    int blitL = 0, blitT = 0, blitR = 0, blitB = 0;
};

// An image with its mip chain (2x2 box filter with rounding), premultiplied
// RGBA8 texels packed level after level.
struct OracleImage
{
    uint32_t width = 0, height = 0, levels = 0;
    std::vector<uint32_t> levelOffset;
    std::vector<uint32_t> texels;
};
void build_image_mips(uint32_t width, uint32_t height, const uint8_t* rgba, OracleImage* out);
// Vulkan-style sampling with the reference's sampler state (sampler key =
// wrapX + 3*wrapY + 9*filter; mip mode nearest).
void sample_image(const OracleImage& img, float u, float v, float lod, uint32_t sampler, float out[4]);

// One logical flush (LogicalFlush, render_context.hpp): its own path,
// contour and clip IDs, gradient texture and tessellation texture.  A frame
// is one or more logical flushes; the first clears the target, later ones
// load it (render_context.cpp:1218-1222).
struct Flush
{
    uint32_t width = 0, height = 0, clearColor = 0, dither = 0;
    bool loadPreserve = false;
    std::shared_ptr<const std::vector<OracleImage>> images; // shared by a frame's flushes
    std::vector<uint32_t> pathImage; // per path: image index (1-based, 0 none) | sampler << 16
    std::vector<rive_hip_path_data> pathData; // [0] reserved
    std::vector<rive_hip_paint_data> paintData;
    std::vector<rive_hip_paint_aux> paintAux;
    std::vector<ContourData> contours;
    std::vector<TessVertexSpan> spans;
    std::vector<CubicRecord> records;
    std::vector<rive_hip_grad_span> gradSpans;
    uint32_t gradSimpleRows = 0, gradComplexRows = 0, gradHeight = 1;
    std::vector<DrawInfo> draws;
    std::vector<uint32_t> pathTessVertexCounts;
    uint32_t pathCount = 0;
    uint32_t tessEnd = 0;
    uint32_t tessVertexCount = 0;
};

// Replays a stream through the front end, splitting it into logical flushes
// where the reference's resource limits require (rive_renderer.cpp:496-570).
void build_flushes(const rive_stream& s, std::vector<Flush>* out);
void render_ramps(const Flush& f, std::vector<uint32_t>* texture);
void tessellate_flush(const Flush& f, std::vector<TessVertex>* tex);
// Renders one logical flush into rgba (cleared first unless f.loadPreserve).
// edgeMask (optional, one byte per pixel) gets 1 where some draw's final
// coverage lies strictly between 0 and 1.  Rows are split over the host's
// threads.
void render_flush(const Flush& f, const std::vector<TessVertex>& tex, uint32_t* rgba, uint8_t* edgeMask);
} // namespace oracle
