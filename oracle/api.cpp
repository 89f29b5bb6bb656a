// CPU ORACLE (test infrastructure only): extern "C" entry points used by the
// tests and by bench.py's cpu_baseline leg.  Never linked into the product.
#include "oracle.hpp"
#include "oracle_flush.hpp"

#include <chrono>
#include <cstring>

using namespace oracle;

// A frame: its logical flushes and the one the record getters read.
struct oracle_flush
{
    std::vector<Flush> flushes;
    size_t sel = 0;
    const Flush& cur() const { return flushes[sel]; }
};

namespace
{
template <typename T> size_t copy_out(const std::vector<T>& v, void* dst, size_t cap)
{
    size_t bytes = v.size() * sizeof(T);
    if (dst != nullptr)
        memcpy(dst, v.data(), bytes < cap ? bytes : cap);
    return bytes;
}
} // namespace

extern "C" {

int oracle_is_aabb(const uint8_t* verbs, uint32_t verbCount, const float* pts, uint32_t ptCount, float* rect)
{
    return oracle::is_aabb(verbs, verbCount, reinterpret_cast<const oracle::V2*>(pts), ptCount, rect) ? 1 : 0;
}

oracle_flush* oracle_build(const rive_stream* s)
{
    auto* of = new oracle_flush;
    build_flushes(*s, &of->flushes);
    return of;
}

uint32_t oracle_flush_count(const oracle_flush* of) { return (uint32_t)of->flushes.size(); }

// Selects the logical flush the record getters read; -1 if out of range.
int oracle_select(oracle_flush* of, uint32_t index)
{
    if (index >= of->flushes.size())
        return -1;
    of->sel = index;
    return 0;
}

void oracle_free(oracle_flush* of) { delete of; }

// Interior triangulation on or off for the frames built from here on
// (off by default, the reference's frameBudgetMs = 0).
void oracle_set_interior_triangulation(int enabled) { set_interior_triangulation(enabled != 0); }

// what: 0 spans, 1 contours, 2 PathData, 3 PaintData, 4 PaintAux, 5 grad spans,
// 6 records (CubicRecord), 7 draws (DrawInfo as 5 u32), 8 info (u32 x 10)
size_t oracle_get(oracle_flush* of, int what, void* dst, size_t cap)
{
    const Flush& f = of->cur();
    switch (what)
    {
        case 0: return copy_out(f.spans, dst, cap);
        case 1: return copy_out(f.contours, dst, cap);
        case 2: return copy_out(f.pathData, dst, cap);
        case 3: return copy_out(f.paintData, dst, cap);
        case 4: return copy_out(f.paintAux, dst, cap);
        case 5: return copy_out(f.gradSpans, dst, cap);
        case 6: return copy_out(f.records, dst, cap);
        case 7:
        {
            std::vector<uint32_t> v;
            for (const DrawInfo& d : f.draws)
            {
                v.push_back(d.pathID);
                v.push_back(d.tessLocation);
                v.push_back(d.tessVertexCount);
                v.push_back((uint32_t)d.dirs);
                // rive_hip_draw.is_stroke: 0 fill, 1 stroke, 2 feathered fill,
                // 3 fill laid out as outer-cubic patches
                v.push_back(d.isStroke ? 1u : (d.featheredFill ? 2u : (d.outerCubics ? 3u : 0u)));
            }
            return copy_out(v, dst, cap);
        }
        case 8:
        {
            std::vector<uint32_t> v = {f.pathCount,
                                       (uint32_t)f.contours.size(),
                                       (uint32_t)f.spans.size(),
                                       f.tessVertexCount,
                                       f.gradSimpleRows,
                                       f.gradComplexRows,
                                       f.gradHeight,
                                       (uint32_t)f.draws.size(),
                                       f.width,
                                       f.height};
            return copy_out(v, dst, cap);
        }
    }
    return 0;
}

size_t oracle_tessellate(oracle_flush* of, void* dst, size_t cap)
{
    std::vector<TessVertex> tex;
    tessellate_flush(of->cur(), &tex);
    return copy_out(tex, dst, cap);
}

size_t oracle_ramps(oracle_flush* of, void* dst, size_t cap)
{
    std::vector<uint32_t> tex;
    render_ramps(of->cur(), &tex);
    return copy_out(tex, dst, cap);
}

// Renders the frame -- every logical flush in order -- into rgba
// (width*height*4 bytes); edges (optional, width*height bytes, zeroed by the
// caller) marks pixels some draw covers fractionally.  Returns seconds spent.
double oracle_render_edges(oracle_flush* of, uint8_t* rgba, uint8_t* edges)
{
    auto t0 = std::chrono::steady_clock::now();
    for (const Flush& f : of->flushes)
    {
        std::vector<TessVertex> tex;
        tessellate_flush(f, &tex);
        render_flush(f, tex, reinterpret_cast<uint32_t*>(rgba), edges);
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

double oracle_render(oracle_flush* of, uint8_t* rgba) { return oracle_render_edges(of, rgba, nullptr); }

// ---- maths entry points for golden-vector tests ----
static void load_pts(const float* p, V2 out[4])
{
    for (int i = 0; i < 4; ++i)
        out[i] = V2{p[2 * i], p[2 * i + 1]};
}
static Mat load_mat(const float* m)
{
    Mat r;
    if (m != nullptr)
        memcpy(r.m, m, sizeof(r.m));
    return r;
}
float oracle_wang_cubic(const float* pts8, float precision, const float* m6)
{
    V2 p[4];
    load_pts(pts8, p);
    return wang_cubic(p, precision, load_mat(m6));
}
float oracle_wang_cubic_pow4(const float* pts8, float precision, const float* m6)
{
    V2 p[4];
    load_pts(pts8, p);
    return wang_cubic_pow4(p, precision, load_mat(m6));
}
float oracle_fast_acos(float x) { return fast_acos(x); }
float oracle_polar_segments_per_radian(float r) { return calc_polar_segments_per_radian(r); }
float oracle_find_max_scale(const float* m6) { return find_max_scale(load_mat(m6)); }
void oracle_eval_cubic(const float* pts8, float t, float* out2)
{
    V2 p[4];
    load_pts(pts8, p);
    V2 r = eval_cubic_at(p, t);
    out2[0] = r.x;
    out2[1] = r.y;
}
void oracle_chop_cubic_at(const float* pts8, float t, float* out14)
{
    V2 p[4];
    load_pts(pts8, p);
    V2 d[7];
    chop_cubic_at(p, d, t);
    memcpy(out14, d, sizeof(d));
}
void oracle_chop_cubic_at2(const float* pts8, float t0, float t1, float* out20)
{
    V2 p[4];
    load_pts(pts8, p);
    V2 d[10];
    chop_cubic_at2(p, d, t0, t1);
    memcpy(out20, d, sizeof(d));
}
int oracle_convex_180_chops(const float* pts8, float* T2, int* areCusps)
{
    V2 p[4];
    load_pts(pts8, p);
    bool c = false;
    float T[2] = {0, 0};
    int n = find_cubic_convex_180_chops(p, T, &c);
    T2[0] = T[0];
    T2[1] = T[1];
    *areCusps = c ? 1 : 0;
    return n;
}
void oracle_blend_coeffs(const float* src3, const float* dstPremul4, uint32_t mode, float* out3)
{
    advanced_blend_coeffs(src3, dstPremul4, mode, out3);
}
uint32_t oracle_shader_blend_mode(uint32_t riveBlendMode) { return shader_blend_mode(riveBlendMode); }

// ---- feathering ----
void oracle_feather_tables(uint16_t* gauss512, uint16_t* inverse512)
{
    memcpy(gauss512, feather_table_f16(), GAUSSIAN_TABLE_SIZE * 2);
    memcpy(inverse512, inverse_feather_table_f16(), GAUSSIAN_TABLE_SIZE * 2);
}
float oracle_table_lookup(const uint16_t* table512, float x) { return table_lookup(table512, x); }
float oracle_feather(float x) { return feather(x); }
float oracle_inverse_feather(float x) { return inverse_feather(x); }
float oracle_eval_feathered_fill(const float* c4) { return eval_feathered_fill(c4); }
float oracle_eval_feathered_stroke(const float* c4) { return eval_feathered_stroke(c4); }
void oracle_pack_feathered_fill(float cornerTheta, float spokeX, float spokeY, float outset, float* out4)
{
    pack_feathered_fill_coverages(cornerTheta, V2{spokeX, spokeY}, outset, out4);
}
float oracle_cubic_max_height(const float* pts8, float* outT)
{
    V2 p[4];
    load_pts(pts8, p);
    return find_cubic_max_height(p[0], p[1], p[2], p[3], outT);
}
float oracle_cubic_local_curvature(const float* pts8, float T, float spread)
{
    V2 p[4];
    load_pts(pts8, p);
    return measure_cubic_local_curvature(p[0], p[1], p[2], p[3], T, spread);
}
// Softens one path for feathering; returns the number of output points and
// the number of verbs in *verbCountOut, or the required sizes when the
// output buffers are too small.
uint32_t oracle_soften_path(const uint8_t* verbs, uint32_t verbCount, const float* pts, uint32_t ptCount,
                            float feather, float matrixMaxScale, uint8_t* outVerbs, uint32_t verbCap,
                            float* outPts, uint32_t ptCap, uint32_t* verbCountOut)
{
    std::vector<uint8_t> v;
    std::vector<V2> p;
    soften_path_for_feathering(verbs, verbCount, reinterpret_cast<const V2*>(pts), ptCount, feather,
                               matrixMaxScale, &v, &p);
    *verbCountOut = (uint32_t)v.size();
    if (v.size() <= verbCap && p.size() <= ptCap)
    {
        memcpy(outVerbs, v.data(), v.size());
        memcpy(outPts, p.data(), p.size() * sizeof(V2));
    }
    return (uint32_t)p.size();
}
}
