// CPU ORACLE for the RenderContext::flush hot path.
//
// Test infrastructure only: this is the checker that the HIP backend is
// compared against.  Nothing in the product (rive-runtime_amd/) links, loads or
// calls it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it.
//
// It is a plain, scalar restatement of the reference's algorithm, function by
// function, compiled with -ffp-contract=off (IEEE, no fused multiply-add) like
// the reference's unit-test build (tests/unit_tests/test.sh:105).
// Each function cites the reference file:line it follows.
//
// Pinning (which checks hold each part fixed) is listed in oracle/README.md.
#pragma once

#include <cstdint>
#include <cstddef>
#include <vector>

namespace oracle
{
struct V2
{
    float x = 0, y = 0;
};
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }
inline V2 operator*(float s, V2 a) { return {s * a.x, s * a.y}; }
inline bool operator==(V2 a, V2 b) { return a.x == b.x && a.y == b.y; }
inline bool operator!=(V2 a, V2 b) { return !(a == b); }

// rive::Mat2D layout: [xx, xy, yx, yy, tx, ty]; x' = xx*x + yx*y + tx,
// y' = xy*x + yy*y + ty  (include/rive/math/mat2d.hpp).
struct Mat
{
    float m[6] = {1, 0, 0, 1, 0, 0};
};

// RiveRenderer::IsAABB restated (render.cpp).
bool is_aabb(const uint8_t* verbs, uint32_t verbCount, const struct V2* pts, uint32_t ptCount, float rect[4]);

// Verb values of rive::PathVerb (include/rive/math/path_types.hpp:29-38).
enum Verb : uint8_t
{
    kMove = 0,
    kLine = 1,
    kQuad = 2,
    kCubic = 4,
    kClose = 5,
};

// ---- constants (renderer/include/rive/renderer/gpu.hpp:56-116,
//      renderer/src/shaders/constants.glsl) ----
constexpr float kParametricPrecision = 4;
constexpr float kPolarPrecision = 8;
constexpr uint32_t kMaxParametricSegments = 1023;
constexpr uint32_t kMaxPolarSegments = 1023;
constexpr uint32_t kTessTextureWidth = 2048;
constexpr uint32_t kMidpointFanPatchSegmentSpan = 8;
constexpr uint32_t kGradTextureWidth = 512;

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

// Shader blend modes (constants.glsl:237-252), PaintData params bits 4..7.
enum : uint32_t
{
    BLEND_SRC_OVER = 0,
    BLEND_MODE_SCREEN = 1,
    BLEND_MODE_OVERLAY = 2,
    BLEND_MODE_DARKEN = 3,
    BLEND_MODE_LIGHTEN = 4,
    BLEND_MODE_COLORDODGE = 5,
    BLEND_MODE_COLORBURN = 6,
    BLEND_MODE_HARDLIGHT = 7,
    BLEND_MODE_SOFTLIGHT = 8,
    BLEND_MODE_DIFFERENCE = 9,
    BLEND_MODE_EXCLUSION = 10,
    BLEND_MODE_MULTIPLY = 11,
    BLEND_MODE_HUE = 12,
    BLEND_MODE_SATURATION = 13,
    BLEND_MODE_COLOR = 14,
    BLEND_MODE_LUMINOSITY = 15,
};
// rive::BlendMode (include/rive/shapes/paint/blend_mode.hpp:5-23) to the
// shader mode (ConvertBlendModeToPLSBlendMode, gpu.cpp:717-753).
uint32_t shader_blend_mode(uint32_t riveBlendMode);
// advanced_blend_coeffs (advanced_blend.glsl:108-280), evaluated in fp32.
void advanced_blend_coeffs(const float src[3], const float dstPremul[4], uint32_t mode, float out[3]);

enum class Join : uint32_t { miter = 0, round = 1, bevel = 2 };
enum class Cap : uint32_t { butt = 0, round = 1, square = 2 };
enum class FillRule : uint32_t { nonZero = 0, evenOdd = 1, clockwise = 2 };

// gpu::ContourDirections (renderer/include/rive/renderer/gpu.hpp).
enum class Dirs : uint32_t
{
    forward = 0,
    reverse = 1,
    forwardThenReverse = 2,
    reverseThenForward = 3,
};
inline bool double_sided(Dirs d)
{
    return d == Dirs::forwardThenReverse || d == Dirs::reverseThenForward;
}

// ---- maths restated from the reference ----
float fast_acos(float x);                                   // simd.hpp:496-506
float wang_cubic_pow4(const V2 p[4], float precision, const Mat& m); // wangs_formula.hpp:157-169
float wang_cubic(const V2 p[4], float precision, const Mat& m);
float calc_polar_segments_per_radian(float approxDevStrokeRadius); // bezier_utils.hpp:108-113
float find_max_scale(const Mat& m);                         // mat2d_find_max_scale.cpp:24-67
V2 eval_cubic_at(const V2 p[4], float t);                   // bezier_utils.cpp:21-32
void chop_cubic_at(const V2 src[4], V2 dst[7], float t);    // bezier_utils.cpp:34-60
void chop_cubic_at2(const V2 src[4], V2 dst[10], float t0, float t1); // :62-101
void chop_cubic_at_n(const V2 src[4], V2 dst[], const float t[], int n); // :103-158
int find_cubic_convex_180_chops(const V2 p[4], float T[2], bool* areCusps); // :180-330
V2 find_cubic_tan0(const V2 p[4]);                          // bezier_utils.hpp:147-151
V2 find_cubic_tan1(const V2 p[4]);                          // bezier_utils.hpp:152-156
void chop_cubic_around_cusps(const V2 p[4], V2 dst[], const float cuspT[], int n,
                             float matrixMaxScale);          // draw.cpp:139-174

// ---- the integer contract ----

// One TessellationWriter::pushCubic() call (render_context.cpp:3160-3229).
struct CubicRecord
{
    V2 pts[4];
    V2 joinTangent;
    uint32_t parametric, polar, join;
    uint32_t contourIDWithFlags;
    uint32_t totalVertexCount;  // padding + parametric + polar + join - 1
    uint32_t forwardLocation;   // first vertex (forward), or ~0u
    uint32_t mirroredLocation;  // one past the mirrored run, or ~0u
};

// gpu::TessVertexSpan, 64 bytes (gpu.hpp:291-375).
struct TessVertexSpan
{
    V2 pts[4];
    V2 joinTangent;
    float y;
    float reflectionY;
    int32_t x0x1;
    int32_t reflectionX0X1;
    uint32_t segmentCounts;
    uint32_t contourIDWithFlags;
};
static_assert(sizeof(TessVertexSpan) == 64, "TessVertexSpan is 64 bytes");

// gpu::ContourData, 16 bytes (gpu.hpp:1698-1714).
struct ContourData
{
    V2 midpoint;
    uint32_t pathID;
    uint32_t vertexIndex0;
};
static_assert(sizeof(ContourData) == 16, "ContourData is 16 bytes");

// One path draw as seen by PathDraw (draw.cpp:561-757).
struct PathDrawDesc
{
    const uint8_t* verbs = nullptr;
    uint32_t verbCount = 0;
    const V2* pts = nullptr;
    uint32_t ptCount = 0;
    Mat matrix;
    float strokeRadius = 0; // 0 => fill
    float featherRadius = 0; // find_feather_radius(paint feather), 0 => none
    Join join = Join::miter;
    Cap cap = Cap::butt;
    FillRule fillRule = FillRule::nonZero;
    uint32_t pathID = 0;
    // Derived (PathDraw ctor):
    Dirs dirs = Dirs::forward;
    uint32_t contourFlags = 0;
};

// Per-path result of initForMidpointFan + pushMidpointFanTessellationData.
struct PathTess
{
    std::vector<CubicRecord> records;
    std::vector<ContourData> contours; // in contour-ID order
    std::vector<uint32_t> contourFirstRecord;
    std::vector<uint32_t> contourPadding;
    uint32_t tessVertexCount = 0; // one direction (draw.cpp:1375-1391)
    uint32_t midpointFanTessVertexCount = 0; // x2 when double sided
    uint32_t maxTessellatedSegmentCount = 0;
};

// Interior triangulation (draw.cpp:510-538): off by default (the reference's
// frameBudgetMs = 0); when on, eligible fills are laid out as outer-cubic
// patches (iterateOuterCubics) by tessellate_path_records.
void set_interior_triangulation(bool enabled);
bool uses_interior_triangulation(const PathDrawDesc& d);

// Derive contourDirections / contourFlags the way PathDraw's constructor does
// for the rasterOrdering interlock mode (draw.cpp:651-719).
void derive_directions(PathDrawDesc* d);

// Count + emit the tessellation records for one path (draw.cpp:768-1392 and
// 1992-2375).  firstContourID is the ID the first contour will receive.
// tessLocation is the path's allocation start (draw.cpp:1899-1990).
PathTess tessellate_path_records(const PathDrawDesc& d,
                                 uint32_t firstContourID,
                                 uint32_t tessLocation);

// Span writing with 2048-wide line wrapping (render_context.cpp:3271-3402).
void write_spans(const CubicRecord& r, Dirs dirs, std::vector<TessVertexSpan>* out);
// Padding spans (render_context.cpp:3404-3424).
void write_padding_span(uint32_t count, uint32_t location,
                        std::vector<TessVertexSpan>* out);

// ---- feathering (feather.cpp) ----
constexpr uint32_t GAUSSIAN_TABLE_SIZE = 512;                  // gpu.hpp:2088
constexpr float kPI_OVER_2 = 1.57079632679f;                   // common.glsl:9
constexpr float FEATHER_COVERAGE_BIAS = -2.f;                  // draw_path_common.glsl:9
constexpr float FEATHER_COVERAGE_THRESHOLD = -1.5f;            // draw_path_common.glsl:13
constexpr float FEATHER_X_COORD_BIAS = .25f;                   // draw_path_common.glsl:17
constexpr float HORIZONTAL_COTANGENT_THRESHOLD = 1e3f;         // draw_path_common.glsl:21
constexpr float HORIZONTAL_COTANGENT_VALUE = 1e6f;             // draw_path_common.glsl:24-25
constexpr uint32_t FEATHER_JOIN_HELPER_VERTEX_COUNT = 3;       // constants.glsl:31
constexpr uint32_t FEATHER_JOIN_MIN_SEGMENT_COUNT = 2 + 4;     // constants.glsl:35-40
constexpr float FEATHER_POLAR_SEGMENT_MIN_ANGLE = 3.14159265f / 16; // gpu.hpp:69
constexpr float COS_FEATHER_POLAR_SEGMENT_MIN_ANGLE_OVER_2 = 0.99518472667f; // gpu.hpp:72-73

uint16_t cast_f32_to_f16(float x);                            // gpu.cpp:1959-1980
float cast_f16_to_f32(uint16_t h);                            // gpu.cpp:1944-1957
void gaussian_integral_table(float table[GAUSSIAN_TABLE_SIZE]);         // gpu.cpp:1991-2063
void inverse_gaussian_integral_table(float table[GAUSSIAN_TABLE_SIZE]); // gpu.cpp:2126-2181
const uint16_t* feather_table_f16();         // the two tables above, as fp16 texels
const uint16_t* inverse_feather_table_f16();
float table_lookup(const uint16_t* table, float x); // linear filter, clamp to edge
float feather(float x);                      // FEATHER (common.glsl:53-60)
float inverse_feather(float x);              // INVERSE_FEATHER (common.glsl:61-68)
void pack_feathered_fill_coverages(float cornerTheta, V2 spokeNorm, float outset, float out[4]);
float eval_feathered_fill(const float coverages[4]);   // draw_path_common.glsl:153-240
float eval_feathered_stroke(const float coverages[4]); // draw_path_common.glsl:242-258
float clamped_divide(float a, float b);                // bezier_utils.glsl:158-169
float find_cubic_max_height(V2 p0, V2 p1, V2 p2, V2 p3, float* outT); // bezier_utils.glsl:171-246
float measure_cubic_local_curvature(V2 p0, V2 p1, V2 p2, V2 p3, float T,
                                    float desiredSpread); // bezier_utils.glsl:72-156
void soften_feathered_cubic(V2* p0, V2* p1, V2* p2, V2* p3, float featherRadius); // tessellate.glsl:146-193
// RiveRenderPath::makeSoftenedCopyForFeathering (rive_render_path.cpp:233-449).
void soften_path_for_feathering(const uint8_t* verbs, uint32_t verbCount, const V2* pts, uint32_t ptCount,
                                float feather, float matrixMaxScale, std::vector<uint8_t>* outVerbs,
                                std::vector<V2>* outPts);
// find_feather_radius (draw.cpp:313-318): the paint's feather is two
// standard deviations wide, the radius covers three.
inline float find_feather_radius(float paintFeather) { return paintFeather * (3.f / 2); }

// ---- tessellation shader restatement (tessellate.glsl:61-567) ----
struct TessVertex
{
    float x, y;
    uint32_t z; // theta bits (or packed feather join data)
    uint32_t contourIDWithFlags;
};
// Fills the vertices covered by one span into a linear 2048-wide texture.
void tessellate_span(const TessVertexSpan& s,
                     const ContourData* contours,      // indexed contourID-1
                     const float* pathRecords,         // PathData as floats, 16/path
                     TessVertex* texture, uint32_t textureRows);
} // namespace oracle
