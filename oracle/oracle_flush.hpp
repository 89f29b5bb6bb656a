// CPU ORACLE (test infrastructure only): one logical flush as the reference's
// RenderContext lays it out (render_context.cpp:1116-1568), plus the raster.
#pragma once

#include "oracle.hpp"
#include "../include/rive_hip.h"

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
};

struct Flush
{
    uint32_t width = 0, height = 0, clearColor = 0, dither = 0;
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

void build_flush(const rive_stream& s, Flush* f);
void render_ramps(const Flush& f, std::vector<uint32_t>* texture);
void tessellate_flush(const Flush& f, std::vector<TessVertex>* tex);
void render_flush(const Flush& f, const std::vector<TessVertex>& tex, uint32_t* rgba);
} // namespace oracle
