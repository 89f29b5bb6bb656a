// CHECKER ONLY (test infrastructure): extern "C" entry points over the
// reference's own maths sources, compiled from the reference tree by
// oracle/ref.mk into oracle/_ref/.  Used by tests to pin the oracle's
// restatement bit for bit.  Wraps:
//   include/rive/math/wangs_formula.hpp:157-178 (cubic_pow4, cubic)
//   include/rive/math/simd.hpp:496-506 (fast_acos)
//   include/rive/math/bezier_utils.hpp:108-152 (polar density, eval, chops)
//   src/math/mat2d_find_max_scale.cpp:24-67 (Mat2D::findMaxScale)
#include "rive/math/bezier_utils.hpp"
#include "rive/math/mat2d.hpp"
#include "rive/math/simd.hpp"
#include "rive/math/wangs_formula.hpp"

#include <cstring>

using namespace rive;

namespace
{
void load(const float* p, Vec2D out[4])
{
    for (int i = 0; i < 4; ++i)
        out[i] = Vec2D(p[2 * i], p[2 * i + 1]);
}
Mat2D mat(const float* m) { return m ? Mat2D(m[0], m[1], m[2], m[3], m[4], m[5]) : Mat2D(); }
} // namespace

extern "C" {
float ref_wang_cubic_pow4(const float* pts8, float precision, const float* m6)
{
    Vec2D p[4];
    load(pts8, p);
    return wangs_formula::cubic_pow4(p, precision, wangs_formula::VectorXform(mat(m6)));
}
float ref_wang_cubic(const float* pts8, float precision, const float* m6)
{
    Vec2D p[4];
    load(pts8, p);
    return wangs_formula::cubic(p, precision, wangs_formula::VectorXform(mat(m6)));
}
float ref_fast_acos(float x) { return simd::fast_acos(float4(x))[0]; }
float ref_polar_segments_per_radian(float r) { return math::calc_polar_segments_per_radian<8>(r); }
float ref_find_max_scale(const float* m6) { return mat(m6).findMaxScale(); }
void ref_eval_cubic(const float* pts8, float t, float* out2)
{
    Vec2D p[4];
    load(pts8, p);
    Vec2D r = math::eval_cubic_at(p, t);
    out2[0] = r.x;
    out2[1] = r.y;
}
void ref_chop_cubic_at(const float* pts8, float t, float* out14)
{
    Vec2D p[4], d[7];
    load(pts8, p);
    math::chop_cubic_at(p, d, t);
    for (int i = 0; i < 7; ++i)
        out14[2 * i] = d[i].x, out14[2 * i + 1] = d[i].y;
}
void ref_chop_cubic_at2(const float* pts8, float t0, float t1, float* out20)
{
    Vec2D p[4], d[10];
    load(pts8, p);
    math::chop_cubic_at(p, d, t0, t1);
    for (int i = 0; i < 10; ++i)
        out20[2 * i] = d[i].x, out20[2 * i + 1] = d[i].y;
}
int ref_convex_180_chops(const float* pts8, float* T2, int* areCusps)
{
    Vec2D p[4];
    load(pts8, p);
    float T[2] = {0, 0};
    bool c = false;
    int n = math::find_cubic_convex_180_chops(p, T, &c);
    T2[0] = T[0];
    T2[1] = T[1];
    *areCusps = c ? 1 : 0;
    return n;
}
}
