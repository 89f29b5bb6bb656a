// Device-side counting for RIVE_HIP_FLUSH_GEOMETRY: the work the reference
// does on the CPU in PathDraw::initForMidpointFan (draw.cpp:768-1392),
// PathDraw::pushMidpointFanTessellationData (draw.cpp:1992-2375) and the
// TessellationWriter (render_context.cpp:3079-3424), done on gfx950.
//
//   count_kernel    one thread per draw: per-curve Wang's formula, convex-180
//                   chops, polar segments, joins, caps and per-contour padding
//                   -> tessellation vertex, contour and record counts.
//   (scan)          draw offsets: tessellation locations, contour IDs,
//                   record slots.
//   emit_kernel     one thread per draw: the same walk again, writing
//                   cubic records (with their locations) and ContourData.
//   span kernels    one thread per record: 2048-wide row wrapping into
//                   TessVertexSpan records, byte-compatible with the
//                   reference's.
//
// All arithmetic is plain IEEE fp32 with no contraction (the file is built
// with -ffp-contract=off); division and sqrt are correctly rounded, so the
// integer outputs match the reference's strict-IEEE CPU build bit for bit.
#pragma once

#include "curve_math.hpp"

namespace rivehip
{
// One cubic record (a TessellationWriter::pushCubic call) as written by the
// device: the span fields before row wrapping.
struct CubicRec
{
    float pts[8];
    float jt[2];
    uint32_t counts; // join<<20 | polar<<10 | parametric
    uint32_t flags;  // contourIDWithFlags
    uint32_t total;  // padding + parametric + polar + join - 1
    uint32_t fwdLoc; // forward location or ~0
    uint32_t mirLoc; // one past the mirrored run, or ~0
    uint32_t dirs;
};
static_assert(sizeof(CubicRec) == 64, "CubicRec is 64 bytes");

struct DrawCounts
{
    uint32_t tessVertexCount; // one direction
    uint32_t contours;
    uint32_t records;
    uint32_t pad;
};

struct Emit
{
    CubicRec* recs;
    ContourV* contours;
    uint32_t* overflow; // set when the emit walk outgrows the counted ranges
    uint32_t recCap, contourCap;
    uint32_t rec, contourIndex, contourID;
    uint32_t fwd, mir;
    uint32_t dirs;
    uint32_t nextPad;
};

__device__ __forceinline__ void push_cubic(Emit* e, const V2d* pts, V2d jt, uint32_t param, uint32_t polar,
                                           uint32_t join, uint32_t flags)
{
    uint32_t total = e->nextPad + param + polar + join - 1u;
    e->nextPad = 0;
    CubicRec r;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        r.pts[2 * k] = pts[k].x;
        r.pts[2 * k + 1] = pts[k].y;
    }
    r.jt[0] = jt.x;
    r.jt[1] = jt.y;
    r.counts = (join << 20) | (polar << 10) | param;
    r.flags = flags;
    r.total = total;
    r.fwdLoc = ~0u;
    r.mirLoc = ~0u;
    r.dirs = e->dirs;
    if (e->dirs != 1u)
    {
        r.fwdLoc = e->fwd;
        e->fwd += total;
    }
    if (e->dirs != 0u)
    {
        r.mirLoc = e->mir;
        e->mir -= total;
    }
    // The counting walk sized this draw's range; never write past it.
    if (e->rec < e->recCap)
        e->recs[e->rec] = r;
    else
        atomicOr(e->overflow, 1u);
    ++e->rec;
}

// An interior-triangulated fill (is_stroke 3; iterateOuterCubics,
// draw.cpp:2401-2650): per contour, every line and the closing line as one
// outer-cubic patch (16 parametric segments, one polar, one bowtie join,
// flagged to cull the segments it does not need), every cubic cut uniformly
// in T into clamp(ceil(Wang / 16), 1, 64) pieces, one patch each, followed by
// ceil((n - 1) / 3) triangle strips retrofitted into patches between the
// cubic's chord and its pieces' chords.  Contours have no midpoint and no
// padding.  The interior triangles are not part of the tessellation.
template <bool EMIT>
__device__ DrawCounts walk_outer_draw(const rive_hip_draw& d, const uint8_t* __restrict__ verbs,
                                      const V2d* __restrict__ P, Emit* e)
{
    constexpr uint32_t kSpan = 16, kMaxSubdivisions = (1023 + kSpan - 1) / kSpan;
    constexpr uint32_t kCull = CULL_EXCESS_TESSELLATION_SEGMENTS_CONTOUR_FLAG;
    DrawCounts out{0, 0, 0, 0};
    uint32_t cid = 0;
    auto patch = [&](const V2d* p) {
        ++out.records;
        if (EMIT)
            push_cubic(e, p, V2d{0.f, 0.f}, kSpan, 1, 1, cid | kCull);
    };
    auto line = [&](V2d a, V2d b) {
        // convert_line_to_cubic (draw.cpp:110-121)
        const float third = 1.f / 3.f;
        const V2d c[4] = {a, (b - a) * third + a, (a - b) * third + b, b};
        patch(c);
    };
    auto same = [](V2d a, V2d b) { return a.x == b.x && a.y == b.y; };
    V2d p0{0.f, 0.f};
    uint32_t contours = 0, pi = 0;
    for (uint32_t k = 0; k < d.verb_count; ++k)
    {
        const uint8_t v = verbs[k];
        if (v == 0)
        {
            if (contours != 0 && !same(P[pi - 1], p0))
                line(P[pi - 1], p0);
            ++out.contours;
            if (EMIT)
            {
                ContourV cdat;
                cdat.mx = 0.f;
                cdat.my = 0.f;
                cdat.pathID = d.path_id;
                cdat.vertexIndex0 = e->dirs != 1u ? e->fwd : e->mir - 1u;
                if (e->contourIndex < e->contourCap)
                    e->contours[e->contourIndex] = cdat;
                else
                    atomicOr(e->overflow, 1u);
                ++e->contourIndex;
                cid = d.contour_flags | e->contourID;
                ++e->contourID;
            }
            p0 = P[pi];
            ++contours;
            pi += 1;
        }
        else if (v == 1)
        {
            line(P[pi - 1], P[pi]);
            pi += 1;
        }
        else if (v == 4)
        {
            const V2d* c = P + pi - 1;
            const float w = sqrtf(sqrtf(wang_pow4(c, 4.f, d.matrix)));
            const uint32_t n = (uint32_t)sclamp(ceilf(w * (1.f / kSpan)), 1.f, (float)kMaxSubdivisions);
            if (n == 1)
                patch(c);
            else
            {
                out.records += n - 1 + (n + 1) / 3;
                if (EMIT)
                {
                    // chop uniformly in T, two chops at a time
                    // (bezier_utils.cpp:104-158 with null tValues)
                    V2d piece[4] = {c[0], c[1], c[2], c[3]}, start = c[0];
                    V2d ends[kMaxSubdivisions + 1];
                    ends[0] = c[0];
                    const int tCount = (int)n - 1;
                    int i = 0, made = 0;
                    for (; i < tCount - 1; i += 2)
                    {
                        V2d ch[10];
                        const float kk = (float)(tCount + 1 - i);
                        chop_at2(piece, ch, 1.f / kk, 2.f / kk);
                        push_cubic(e, ch, V2d{0.f, 0.f}, kSpan, 1, 1, cid | kCull);
                        push_cubic(e, ch + 3, V2d{0.f, 0.f}, kSpan, 1, 1, cid | kCull);
                        ends[++made] = ch[3];
                        ends[++made] = ch[6];
                        piece[0] = ch[6], piece[1] = ch[7], piece[2] = ch[8], piece[3] = ch[9];
                    }
                    if (i < tCount)
                    {
                        V2d ch[7];
                        chop_at(piece, ch, .5f);
                        push_cubic(e, ch, V2d{0.f, 0.f}, kSpan, 1, 1, cid | kCull);
                        ends[++made] = ch[3];
                        piece[0] = ch[3], piece[1] = ch[4], piece[2] = ch[5], piece[3] = ch[6];
                    }
                    push_cubic(e, piece, V2d{0.f, 0.f}, kSpan, 1, 1, cid | kCull);
                    ends[++made] = piece[3];
                    // the fan of glue strips anchored at the cubic's start
                    for (uint32_t a = 1; a < n; a += 3)
                    {
                        const uint32_t remaining = n - a;
                        const uint32_t b = remaining == 1 ? a : a + 1;
                        V2d strip[5];
                        uint32_t m = 0;
                        strip[m++] = ends[b];
                        strip[m++] = ends[b + 1];
                        if (remaining > 1)
                            strip[m++] = ends[b - 1];
                        if (remaining > 2)
                            strip[m++] = ends[b + 2];
                        strip[m++] = start;
                        // pushRetrofitCubicTriStrip (render_context.cpp:3231-3257)
                        const V2d q[4] = {strip[0], strip[1], strip[m < 4 ? m - 1 : 3], strip[2]};
                        push_cubic(e, q, strip[m < 5 ? m - 1 : 4], kSpan, 1, 1,
                                   cid | RETROFIT_TRI_STRIP_CONTOUR_FLAG);
                    }
                }
                ++out.records; // the last piece (the others are counted above)
            }
            pi += 3;
        }
        else if (v == 2)
            pi += 2;
    }
    if (contours != 0 && d.point_count != 0 && !same(P[d.point_count - 1], p0))
        line(P[d.point_count - 1], p0);
    out.tessVertexCount = out.records * (kSpan + 1);
    return out;
}

// Walks one draw.  EMIT = false counts; EMIT = true writes records.
template <bool EMIT>
__device__ DrawCounts walk_draw(const rive_hip_draw& d, const uint8_t* __restrict__ verbs,
                                const V2d* __restrict__ P, Emit* e)
{
    DrawCounts out{0, 0, 0, 0};
    // is_stroke: 0 fill, 1 stroke, 2 feathered fill (counted like a stroke
    // whose joins are feather joins, draw.cpp:768-1392)
    const bool isStroke = d.is_stroke == 1u;
    const bool feather = d.is_stroke == 2u;
    const bool strokeLike = isStroke || feather;
    const bool roundJoin = isStroke && d.join == 1;
    const float spr = d.polar_segments_per_radian;
    const float maxScale = d.matrix_max_scale;
    const float prec = d.parametric_precision;
    const uint32_t nv = d.verb_count;
    uint32_t vi = 0, pi = 0;
    while (vi < nv)
    {
        // ---- one contour: [cs, ce) verbs starting at point ps
        uint32_t cs = vi, ps = pi;
        // Contour end: the next move after cs.
        uint32_t ce = cs + 1, pe = ps + (verbs[cs] == 0 ? 1u : 0u);
        if (verbs[cs] != 0)
        {
            // A contour always starts with a move in a RawPath.
            ce = cs;
            pe = ps;
        }
        while (ce < nv && verbs[ce] != 0)
        {
            uint8_t v = verbs[ce];
            pe += v == 1 ? 1u : (v == 2 ? 2u : (v == 4 ? 3u : 0u));
            ++ce;
        }
        if (ce == cs)
        {
            ++vi; // malformed: skip a stray verb
            continue;
        }
        bool closed = !isStroke;
        for (uint32_t k = cs; k < ce; ++k)
            if (verbs[k] == 5)
                closed = true;
        const V2d movePt = P[ps];
        const V2d finalPt = P[pe - 1];
        const bool implicitClose =
            closed && (__float_as_uint(movePt.x) != __float_as_uint(finalPt.x) ||
                       __float_as_uint(movePt.y) != __float_as_uint(finalPt.y));

        // ---- walk 1: counts (draw.cpp:905-1376)
        uint32_t lines = 0, curves = 0, paramSum = 0, polarSum = 0, preChop = 0;
        V2d endpointsSum{0.f, 0.f};
        V2d lastTan{0.f, 1.f}, firstTan{0.f, 1.f};
        {
            uint32_t q = ps + 1;
            for (uint32_t k = cs + 1; k < ce; ++k)
            {
                uint8_t v = verbs[k];
                if (v == 1)
                {
                    const V2d* lp = P + q - 1;
                    if (roundJoin)
                    {
                        V2d tg = lp[1] - lp[0];
                        if (preChop == 0)
                            firstTan = tg;
                        else
                            polarSum += polar_count(lastTan, tg, spr);
                        lastTan = tg;
                    }
                    ++preChop;
                    endpointsSum = endpointsSum + lp[1];
                    ++lines;
                    q += 1;
                }
                else if (v == 4)
                {
                    const V2d* cp = P + q - 1;
                    ++preChop;
                    endpointsSum = endpointsSum + cp[3];
                    if (!strokeLike)
                    {
                        paramSum += parametric_count(wang_pow4(cp, prec, d.matrix));
                        ++curves;
                    }
                    else if (feather)
                    {
                        // feathered fills were chopped on the CPU and have no
                        // polar segments: a zero rotation (draw.cpp:1062-1128)
                        paramSum += parametric_count(wang_pow4(cp, prec, d.matrix));
                        polarSum += polar_count(V2d{0.f, 1.f}, V2d{0.f, 1.f}, spr);
                        ++curves;
                    }
                    else
                    {
                        if (roundJoin)
                        {
                            V2d t0 = tan0(cp), t1 = tan1(cp);
                            if (preChop == 1)
                                firstTan = t0;
                            else
                                polarSum += polar_count(lastTan, t0, spr);
                            lastTan = t1;
                        }
                        V2d sub[16];
                        int n = stroke_subcubics(cp, maxScale, sub);
                        for (int s = 0; s < n; ++s)
                        {
                            paramSum += parametric_count(wang_pow4(sub + 3 * s, prec, d.matrix));
                            polarSum += polar_count(tan0(sub + 3 * s), tan1(sub + 3 * s), spr);
                            ++curves;
                        }
                    }
                    q += 3;
                }
                else if (v == 2)
                {
                    q += 2; // quads are converted to cubics by the path object
                }
            }
        }
        // finishAndAppendContour (draw.cpp:916-972)
        if (implicitClose)
        {
            if (roundJoin)
            {
                V2d tg = movePt - finalPt;
                polarSum += polar_count(lastTan, tg, spr);
                lastTan = tg;
            }
            ++lines;
            ++preChop;
            endpointsSum = endpointsSum + movePt;
        }
        if (closed && roundJoin && preChop != 0)
            polarSum += polar_count(lastTan, firstTan, spr);
        uint32_t strokeJoinCount = preChop;
        if (!closed)
            strokeJoinCount = (strokeJoinCount > 1u ? strokeJoinCount : 1u) - 1u;
        const bool empty = lines == 0 && curves == 0;
        uint32_t v = lines * 2 + paramSum;
        uint32_t capSegs = 0;
        uint32_t capFlags = 0;
        if (strokeLike)
        {
            v += polarSum;
            if (feather)
                v += strokeJoinCount * (feather_join_segments(spr) - 1u);
            else if (roundJoin)
                v -= strokeJoinCount;
            else
                v += strokeJoinCount * (kMiterBevelJoinSegments - 1);
            uint32_t cap = d.cap;
            bool needsCaps = false;
            if (!empty)
                needsCaps = !closed;
            else if (isStroke)
            {
                // empty_stroke_cap (draw.cpp:253-281)
                if (closed)
                    cap = d.join == 1 ? 1u : (d.join == 0 ? 2u : 0u);
                needsCaps = cap != 0u;
            }
            if (needsCaps)
            {
                if (cap == 1u)
                {
                    float n = ceilf(spr * kMathPI);
                    n += 2.f;
                    n = fminf(n, 1023.f);
                    capSegs = (uint32_t)n;
                }
                else
                {
                    capSegs = kMiterBevelJoinSegments;
                }
                v += (capSegs - 1u) * 2u;
                capFlags = (cap == 0u ? BEVEL_JOIN_CONTOUR_FLAG
                                      : (cap == 2u ? MITER_CLIP_JOIN_CONTOUR_FLAG : ROUND_JOIN_CONTOUR_FLAG)) |
                           EMULATED_STROKE_CAP_CONTOUR_FLAG;
            }
        }
        else
        {
            v += curves;
        }
        const uint32_t padding = (kPatchSpan - (v % kPatchSpan)) % kPatchSpan;
        v += padding;
        out.tessVertexCount += v;
        out.contours += 1;
        // records in this contour
        uint32_t recs = (capSegs != 0 && !empty) ? 1u : 0u;
        recs += (capSegs != 0 && empty) ? 2u : 0u;
        recs += lines - (implicitClose ? 1u : 0u); // explicit lines
        recs += curves;
        recs += implicitClose ? 1u : 0u;
        out.records += recs;

        if (EMIT)
        {
            // ---- walk 2: records (draw.cpp:1992-2375)
            e->nextPad = padding;
            ContourV cdat;
            cdat.mx = isStroke ? (closed ? 1.f : 0.f) : endpointsSum.x * (1.f / (float)preChop);
            cdat.my = isStroke ? 0.f : endpointsSum.y * (1.f / (float)preChop);
            cdat.pathID = d.path_id;
            cdat.vertexIndex0 = e->dirs != 1u ? e->fwd : e->mir - 1u;
            if (e->contourIndex < e->contourCap)
                e->contours[e->contourIndex] = cdat;
            else
                atomicOr(e->overflow, 1u);
            ++e->contourIndex;
            const uint32_t cid = (d.contour_flags | e->contourID);
            ++e->contourID;
            uint32_t joinTypeFlags =
                isStroke ? (d.join == 0 ? MITER_REVERT_JOIN_CONTOUR_FLAG
                                        : (d.join == 1 ? ROUND_JOIN_CONTOUR_FLAG : BEVEL_JOIN_CONTOUR_FLAG))
                         : (feather ? FEATHER_JOIN_CONTOUR_FLAG : 0u);
            const uint32_t notRoundJoinSegs = feather ? feather_join_segments(spr) : kMiterBevelJoinSegments;
            bool needsFirstCap = capSegs != 0;
            V2d joinTangent{0.f, 1.f};
            uint32_t joinSegs = 1;
            const V2d* cpts = P + ps;
            const V2d* cend = P + pe;
            // Tangent of the first verb after `k` (for round joins), or the
            // closing line's / contour start's tangent.
            V2d lastT{0.f, 1.f};
            auto capBefore = [&](const V2d* cubic) {
                V2d rev[4] = {cubic[3], cubic[2], cubic[1], cubic[0]};
                push_cubic(e, rev, tan0(cubic), 0, 0, capSegs, cid | capFlags);
            };
            // find_join_tangent (draw.cpp:210-251)
            auto joinTangentAt = [&](const V2d* jp) -> V2d {
                const V2d* np = jp + 1;
                np = np != cend ? np : cpts;
                V2d tg = *np - *jp;
                if (tg.x != 0.f || tg.y != 0.f)
                    return tg;
                for (const V2d* s = jp + 1; s != cend; ++s)
                    if (vne(*s, *jp))
                        return *s - *jp;
                if (closed)
                    for (const V2d* s = cpts; s != jp; ++s)
                        if (vne(*s, *jp))
                            return *s - *jp;
                return V2d{0.f, 0.f};
            };
            auto endingTangent = [&]() -> V2d { // draw.cpp:195-208
                const V2d ep = cend[-1];
                const V2d* s = cend;
                while (--s > cpts)
                {
                    V2d pp = s[-1];
                    if (vne(pp, ep))
                        return ep - pp;
                }
                return V2d{-1.f, 0.f};
            };
            auto startingTangent = [&]() -> V2d { // draw.cpp:178-191
                const V2d p0 = cpts[0];
                for (const V2d* s = cpts + 1; s < cend; ++s)
                    if (vne(*s, p0))
                        return *s - p0;
                return V2d{1.f, 0.f};
            };
            // First tangent of the verb at index k (k < ce), for round joins.
            auto verbFirstTangent = [&](uint32_t k, uint32_t q) -> V2d {
                const V2d* pp = P + q - 1;
                if (verbs[k] == 1)
                    return pp[1] - pp[0];
                return tan0(pp);
            };
            uint32_t q = ps + 1;
            for (uint32_t k = cs + 1; k < ce; ++k)
            {
                uint8_t vb = verbs[k];
                if (vb != 1 && vb != 4)
                {
                    if (vb == 2)
                        q += 2;
                    continue;
                }
                const bool isFinal = k + 1 == ce || (k + 2 == ce && verbs[k + 1] == 5);
                const uint32_t qn = q + (vb == 1 ? 1u : 3u);
                // Flags of the verb's last record: the contour's join type, or
                // the emulated cap at the end of an open contour.
                uint32_t lastFlags = joinTypeFlags;
                // The join that follows this verb.
                if (strokeLike)
                {
                    if (closed || !isFinal)
                    {
                        if (roundJoin)
                        {
                            V2d myLast = vb == 1 ? (P[q] - P[q - 1]) : tan1(P + q - 1);
                            V2d nextFirst;
                            // next line/cubic verb in the contour, else the closing line or firstTan
                            uint32_t kk = k + 1, qq = qn;
                            while (kk < ce && verbs[kk] != 1 && verbs[kk] != 4)
                            {
                                if (verbs[kk] == 2)
                                    qq += 2;
                                ++kk;
                            }
                            if (kk < ce)
                                nextFirst = verbFirstTangent(kk, qq);
                            else if (implicitClose)
                                nextFirst = movePt - finalPt;
                            else
                                nextFirst = firstTan;
                            joinTangent = nextFirst;
                            joinSegs = polar_count(myLast, nextFirst, spr);
                        }
                        else
                        {
                            joinTangent = joinTangentAt(P + qn - 1);
                            joinSegs = notRoundJoinSegs;
                        }
                    }
                    else
                    {
                        V2d et = endingTangent();
                        joinTangent = V2d{-et.x, -et.y};
                        lastFlags = capFlags;
                        joinSegs = capSegs;
                    }
                }
                if (vb == 1)
                {
                    const V2d* lp = P + q - 1;
                    const float third = 1.f / 3.f;
                    V2d cub[4] = {lp[0], (lp[1] - lp[0]) * third + lp[0], (lp[0] - lp[1]) * third + lp[1],
                                  lp[1]};
                    if (needsFirstCap)
                    {
                        capBefore(cub);
                        needsFirstCap = false;
                    }
                    push_cubic(e, cub, joinTangent, 1, 1, joinSegs, cid | lastFlags);
                }
                else
                {
                    const V2d* cp = P + q - 1;
                    if (!strokeLike)
                    {
                        push_cubic(e, cp, V2d{0.f, 0.f}, parametric_count(wang_pow4(cp, prec, d.matrix)), 1, 1,
                                   cid);
                    }
                    else if (feather)
                    {
                        push_cubic(e, cp, joinTangent, parametric_count(wang_pow4(cp, prec, d.matrix)),
                                   polar_count(V2d{0.f, 1.f}, V2d{0.f, 1.f}, spr), joinSegs, cid | lastFlags);
                    }
                    else
                    {
                        V2d sub[16];
                        int n = stroke_subcubics(cp, maxScale, sub);
                        if (needsFirstCap)
                        {
                            capBefore(sub);
                            needsFirstCap = false;
                        }
                        // Chops before the last one keep the previous join
                        // tangent and the contour's join type.
                        for (int s = 0; s + 1 < n; ++s)
                        {
                            const V2d* sc = sub + 3 * s;
                            push_cubic(e, sc, lastT, parametric_count(wang_pow4(sc, prec, d.matrix)),
                                       polar_count(tan0(sc), tan1(sc), spr), 1, cid | joinTypeFlags);
                        }
                        const V2d* sc = sub + 3 * (n - 1);
                        push_cubic(e, sc, joinTangent, parametric_count(wang_pow4(sc, prec, d.matrix)),
                                   polar_count(tan0(sc), tan1(sc), spr), joinSegs, cid | lastFlags);
                    }
                }
                lastT = joinTangent;
                q = qn;
            }
            if (needsFirstCap)
            {
                V2d p0 = cpts[0], left{p0.x - 1.f, p0.y}, right{p0.x + 1.f, p0.y};
                V2d a[4] = {p0, right, right, right};
                V2d b[4] = {p0, left, left, left};
                capBefore(a);
                capBefore(b);
            }
            else if (closed && implicitClose)
            {
                V2d cub[4];
                const float third = 1.f / 3.f;
                cub[0] = finalPt;
                cub[1] = (movePt - finalPt) * third + finalPt;
                cub[2] = (finalPt - movePt) * third + movePt;
                cub[3] = movePt;
                if (roundJoin)
                {
                    joinTangent = firstTan;
                    joinSegs = polar_count(movePt - finalPt, firstTan, spr);
                }
                else if (strokeLike)
                {
                    joinTangent = startingTangent();
                    joinSegs = notRoundJoinSegs;
                }
                push_cubic(e, cub, joinTangent, 1, 1, joinSegs, cid | joinTypeFlags);
            }
        }
        vi = ce;
        pi = pe;
    }
    return out;
}

__global__ void count_kernel(const rive_hip_draw* __restrict__ draws, uint32_t n,
                             const uint8_t* __restrict__ verbs, uint64_t verbCapacity,
                             const V2d* __restrict__ points, uint64_t pointCapacity,
                             uint32_t* __restrict__ invalid,
                             uint32_t* __restrict__ tessCounts, uint32_t* __restrict__ contourCounts,
                             uint32_t* __restrict__ recordCounts)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const rive_hip_draw d = draws[i];
    // Validate the host-provided ranges before walking them: the verbs must
    // stay inside the VERB buffer and consume exactly point_count points.
    bool ok = (uint64_t)d.first_verb + d.verb_count <= verbCapacity &&
              (uint64_t)d.first_point + d.point_count <= pointCapacity && d.verb_count > 0 &&
              verbs[d.first_verb] == 0;
    if (ok)
    {
        uint32_t pts = 0;
        for (uint32_t k = 0; k < d.verb_count; ++k)
        {
            uint8_t v = verbs[d.first_verb + k];
            pts += v == 0 ? 1u : v == 1 ? 1u : v == 2 ? 2u : v == 4 ? 3u : 0u;
            if (v == 3 || v > 5)
                ok = false;
        }
        ok = ok && pts == d.point_count;
    }
    if (!ok)
    {
        atomicAdd(invalid, 1u);
        tessCounts[i] = 0;
        contourCounts[i] = 0;
        recordCounts[i] = 0;
        return;
    }
    DrawCounts c = d.is_stroke == 3u ? walk_outer_draw<false>(d, verbs + d.first_verb, points + d.first_point, nullptr)
                                     : walk_draw<false>(d, verbs + d.first_verb, points + d.first_point, nullptr);
    // midpointFanTessVertexCount: doubled for double-sided draws (draw.cpp:1384-1390)
    const bool doubleSided = d.contour_directions == 2u || d.contour_directions == 3u;
    tessCounts[i] = doubleSided ? c.tessVertexCount * 2u : c.tessVertexCount;
    contourCounts[i] = c.contours;
    recordCounts[i] = c.records;
}

// tessOffsets/contourOffsets/recordOffsets: exclusive scans over the draws
// (record and contour storage); tessLocal/contourLocal: each draw's
// tessellation location and first contour ID within its own logical flush
// (the same scans when the draws form one flush).
__global__ void emit_kernel(rive_hip_draw* __restrict__ draws, uint32_t n, const uint8_t* __restrict__ verbs,
                            const V2d* __restrict__ points, const uint32_t* __restrict__ tessOffsets,
                            const uint32_t* __restrict__ contourOffsets,
                            const uint32_t* __restrict__ recordOffsets, const uint32_t* __restrict__ tessLocal,
                            const uint32_t* __restrict__ contourLocal, uint32_t firstLocation,
                            CubicRec* __restrict__ recs, ContourV* __restrict__ contours,
                            uint32_t* __restrict__ overflow)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    rive_hip_draw d = draws[i];
    const uint32_t loc = firstLocation + tessLocal[i];
    const uint32_t count = tessOffsets[i + 1] - tessOffsets[i];
    Emit e;
    e.recs = recs + recordOffsets[i];
    e.contours = contours + contourOffsets[i];
    e.overflow = overflow;
    e.recCap = recordOffsets[i + 1] - recordOffsets[i];
    e.contourCap = contourOffsets[i + 1] - contourOffsets[i];
    e.rec = 0;
    e.contourIndex = 0;
    e.contourID = contourLocal[i] + 1u; // contour IDs are 1-based per logical flush
    e.dirs = d.contour_directions;
    e.nextPad = 0;
    // pushTessellationData (draw.cpp:1899-1990)
    switch (d.contour_directions)
    {
        case 0:
            e.fwd = loc;
            e.mir = 0;
            break;
        case 1:
            e.fwd = 0;
            e.mir = loc + count;
            break;
        case 3: // reverseThenForward
            e.fwd = e.mir = loc + count / 2u;
            break;
        default: // forwardThenReverse
            e.fwd = loc;
            e.mir = loc + count;
            break;
    }
    if (d.is_stroke == 3u)
        walk_outer_draw<true>(d, verbs + d.first_verb, points + d.first_point, &e);
    else
        walk_draw<true>(d, verbs + d.first_verb, points + d.first_point, &e);
    if (e.rec != e.recCap || e.contourIndex != e.contourCap)
        atomicOr(overflow, 1u);
    draws[i].tess_location = loc;
    draws[i].tess_vertex_count = count;
}

// Row wrapping (render_context.cpp:3271-3402): number of spans of a record.
__device__ __forceinline__ uint32_t record_span_count(const CubicRec& r)
{
    const int32_t W = (int32_t)kTessTextureWidth;
    uint32_t n = 0;
    if (r.dirs == 0u)
    {
        int32_t x1 = (int32_t)(r.fwdLoc % kTessTextureWidth) + (int32_t)r.total;
        for (;;)
        {
            ++n;
            if (x1 > W)
            {
                x1 -= W;
                continue;
            }
            break;
        }
    }
    else if (r.dirs == 1u)
    {
        int32_t rx1 = (int32_t)((r.mirLoc - 1u) % kTessTextureWidth + 1u) - (int32_t)r.total;
        for (;;)
        {
            ++n;
            if (rx1 < 0)
            {
                rx1 += W;
                continue;
            }
            break;
        }
    }
    else
    {
        int32_t x1 = (int32_t)(r.fwdLoc % kTessTextureWidth) + (int32_t)r.total;
        int32_t rx1 = (int32_t)((r.mirLoc - 1u) % kTessTextureWidth + 1u) - (int32_t)r.total;
        for (;;)
        {
            ++n;
            if (x1 > W || rx1 < 0)
            {
                x1 -= W;
                rx1 += W;
                continue;
            }
            break;
        }
    }
    return n;
}

__global__ void span_count_kernel(const CubicRec* __restrict__ recs, uint32_t n, uint32_t* __restrict__ counts)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    counts[i] = record_span_count(recs[i]);
}

__device__ __forceinline__ void write_span(SpanV* out, const float* pts, const float* jt, float y, int32_t x0,
                                           int32_t x1, float ry, int32_t rx0, int32_t rx1, uint32_t counts,
                                           uint32_t flags)
{
    SpanV s;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        s.p[k] = pts[k];
    s.jt[0] = jt[0];
    s.jt[1] = jt[1];
    s.y = y;
    s.ry = ry;
    s.x0x1 = (int32_t)(((uint32_t)x1 << 16) | ((uint32_t)x0 & 0xffffu));
    s.rx0x1 = (int32_t)(((uint32_t)rx1 << 16) | ((uint32_t)rx0 & 0xffffu));
    s.counts = counts;
    s.flags = flags;
    *out = s;
}

// Up to this many logical flushes are laid out by one frame launch.
constexpr uint32_t kMaxFrameFlushes = 64;
// Per-flush boundaries passed by value: v[k] is where flush k starts (in
// draws, records, ...), v[n] the end.
struct FrameBounds
{
    uint32_t n;
    uint32_t v[kMaxFrameFlushes + 1];
};
__device__ __forceinline__ uint32_t frame_flush_of(const FrameBounds& b, uint32_t i)
{
    uint32_t k = 0;
    while (k + 1 < b.n && b.v[k + 1] <= i)
        ++k;
    return k;
}

// Spans of the records of several logical flushes, each flush's spans after
// its two padding spans: record i of flush k writes at
// 2 * (k + 1) + spanOffsets[i].  One flush: recordBounds {1, {0, n}}.
__global__ void span_write_kernel(const CubicRec* __restrict__ recs, uint32_t n,
                                  const uint32_t* __restrict__ spanOffsets, const FrameBounds recordBounds,
                                  SpanV* __restrict__ spans)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const CubicRec r = recs[i];
    SpanV* out = spans + 2u * (frame_flush_of(recordBounds, i) + 1u) + spanOffsets[i];
    const int32_t W = (int32_t)kTessTextureWidth;
    const float qnan = __uint_as_float(0x7fc00000u);
    if (r.dirs == 0u)
    {
        uint32_t y = r.fwdLoc / kTessTextureWidth;
        int32_t x0 = (int32_t)(r.fwdLoc % kTessTextureWidth);
        int32_t x1 = x0 + (int32_t)r.total;
        for (;;)
        {
            write_span(out++, r.pts, r.jt, (float)y, x0, x1, qnan, -1, -1, r.counts, r.flags);
            if (x1 > W)
            {
                ++y;
                x0 -= W;
                x1 -= W;
                continue;
            }
            break;
        }
    }
    else if (r.dirs == 1u)
    {
        uint32_t ry = (r.mirLoc - 1u) / kTessTextureWidth;
        int32_t rx0 = (int32_t)((r.mirLoc - 1u) % kTessTextureWidth + 1u);
        int32_t rx1 = rx0 - (int32_t)r.total;
        for (;;)
        {
            write_span(out++, r.pts, r.jt, (float)ry, rx0, rx1, qnan, -1, -1, r.counts, r.flags);
            if (rx1 < 0)
            {
                --ry;
                rx0 += W;
                rx1 += W;
                continue;
            }
            break;
        }
    }
    else
    {
        int32_t y = (int32_t)(r.fwdLoc / kTessTextureWidth);
        int32_t x0 = (int32_t)(r.fwdLoc % kTessTextureWidth);
        int32_t x1 = x0 + (int32_t)r.total;
        uint32_t ry = (r.mirLoc - 1u) / kTessTextureWidth;
        int32_t rx0 = (int32_t)((r.mirLoc - 1u) % kTessTextureWidth + 1u);
        int32_t rx1 = rx0 - (int32_t)r.total;
        for (;;)
        {
            write_span(out++, r.pts, r.jt, (float)y, x0, x1, (float)ry, rx0, rx1, r.counts, r.flags);
            if (x1 > W || rx1 < 0)
            {
                ++y;
                x0 -= W;
                x1 -= W;
                --ry;
                rx0 += W;
                rx1 += W;
                continue;
            }
            break;
        }
    }
}

// Per draw of a frame: its location and first contour ID within its own
// logical flush (the frame-wide scans less their values at the flush start).
__global__ void frame_local_kernel(uint32_t n, const FrameBounds drawBounds, const uint32_t* __restrict__ tessScan,
                                   const uint32_t* __restrict__ contourScan, uint32_t* __restrict__ tessLocal,
                                   uint32_t* __restrict__ contourLocal)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t first = drawBounds.v[frame_flush_of(drawBounds, i)];
    tessLocal[i] = tessScan[i] - tessScan[first];
    contourLocal[i] = contourScan[i] - contourScan[first];
}

// The scans' values at the flush boundaries: out[j * (n + 1) + k] = scan_j at
// bounds.v[k] for the three scans.
__global__ void frame_bounds_kernel(const FrameBounds bounds, const uint32_t* __restrict__ a,
                                    const uint32_t* __restrict__ b, const uint32_t* __restrict__ c,
                                    uint32_t* __restrict__ out)
{
    const uint32_t k = threadIdx.x;
    if (k > bounds.n)
        return;
    const uint32_t m = bounds.n + 1;
    out[k] = a[bounds.v[k]];
    out[m + k] = b[bounds.v[k]];
    out[2 * m + k] = c != nullptr ? c[bounds.v[k]] : 0u;
}

// Padding spans (render_context.cpp:1550-1568): 8 vertices at 0 and one after
// the last path.
__global__ void padding_span_kernel(SpanV* __restrict__ spans, const uint32_t* __restrict__ tessTotal,
                                    uint32_t firstLocation)
{
    const float zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const float qnan = __uint_as_float(0x7fc00000u);
    const uint32_t counts = (1u << 20); // join 1, polar 0, parametric 0
    if (threadIdx.x == 0)
        write_span(&spans[0], zero, zero, 0.f, 0, (int32_t)kPatchSpan, qnan, -1, -1, counts, 0u);
    if (threadIdx.x == 1)
    {
        uint32_t endLoc = firstLocation + *tessTotal;
        uint32_t y = endLoc / kTessTextureWidth;
        int32_t x0 = (int32_t)(endLoc % kTessTextureWidth);
        write_span(&spans[1], zero, zero, (float)y, x0, x0 + 1, qnan, -1, -1, counts, 0u);
    }
}

// The padding spans of every logical flush of a frame: block k writes flush
// k's two at spanBase.v[k], its end padding after tessTotals.v[k] texels.
__global__ void frame_padding_kernel(SpanV* __restrict__ spans, const FrameBounds spanBase,
                                     const FrameBounds tessTotals, uint32_t firstLocation)
{
    const uint32_t k = blockIdx.x;
    const float zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const float qnan = __uint_as_float(0x7fc00000u);
    const uint32_t counts = (1u << 20);
    SpanV* out = spans + spanBase.v[k];
    if (threadIdx.x == 0)
        write_span(&out[0], zero, zero, 0.f, 0, (int32_t)kPatchSpan, qnan, -1, -1, counts, 0u);
    if (threadIdx.x == 1)
    {
        uint32_t endLoc = firstLocation + tessTotals.v[k];
        uint32_t y = endLoc / kTessTextureWidth;
        int32_t x0 = (int32_t)(endLoc % kTessTextureWidth);
        write_span(&out[1], zero, zero, (float)y, x0, x0 + 1, qnan, -1, -1, counts, 0u);
    }
}
} // namespace rivehip

namespace
{
// GEOMETRY-mode front half: count -> scan -> emit -> spans.  Leaves spans and
// contours in ctx scratch and the tessellation ranges in the DRAW buffer.
int geometry_counting(rive_hip_context* ctx, const rive_hip_flush_desc* d, const rivehip::SpanV** spans,
                      uint32_t* spanCount, const rivehip::ContourV** contours, uint32_t* tessRows)
{
    using namespace rivehip;
    const uint32_t n = d->draw_count;
    auto* draws = ctx->in.draws;
    const auto* verbs = reinterpret_cast<const uint8_t*>(ctx->bufs[RIVE_HIP_BUFFER_VERB].dev);
    const auto* points = reinterpret_cast<const V2d*>(ctx->bufs[RIVE_HIP_BUFFER_POINT].dev);
    if (n > 0 && (verbs == nullptr || points == nullptr))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "GEOMETRY flush without VERB/POINT buffers");
    uint32_t* scal = P<uint32_t>(ctx->scalars);
    // Per-draw counts: tess, contours, records (3 arrays of n) + 3 offset arrays of n+1.
    RIVE_TRY(ensure(ctx, ctx->geoCounts, (size_t)(6 * (n + 1) + 16) * 4));
    uint32_t* tessCounts = P<uint32_t>(ctx->geoCounts);
    uint32_t* contourCounts = tessCounts + (n + 1);
    uint32_t* recordCounts = contourCounts + (n + 1);
    uint32_t* tessOffsets = recordCounts + (n + 1);
    uint32_t* contourOffsets = tessOffsets + (n + 1);
    uint32_t* recordOffsets = contourOffsets + (n + 1);
    if (n > 0)
        // one thread walks one draw: half-wave blocks spread the draws over
        // more CUs (the walks are serial and latency-bound)
        count_kernel<<<(n + 31) / 32, 32, 0, ctx->stream>>>(
            draws, n, verbs, ctx->bufs[RIVE_HIP_BUFFER_VERB].size, points,
            ctx->bufs[RIVE_HIP_BUFFER_POINT].size / sizeof(V2d), scal + 12, tessCounts, contourCounts,
            recordCounts);
    HIP_TRY(hipGetLastError());
    int rc = device_scan(ctx, tessCounts, n, tessOffsets, scal + 8);
    if (rc)
        return rc;
    rc = device_scan(ctx, contourCounts, n, contourOffsets, scal + 9);
    if (rc)
        return rc;
    rc = device_scan(ctx, recordCounts, n, recordOffsets, scal + 10);
    if (rc)
        return rc;
    RIVE_TRY(fetch(ctx, ctx->hostScalars, scal + 8, 20));
    RIVE_TRY(settle(ctx));
    if (ctx->hostScalars[4] != 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT,
                    "draw record with verb/point ranges outside the VERB/POINT buffers");
    const uint32_t tessTotal = ctx->hostScalars[0];
    const uint32_t contourTotal = ctx->hostScalars[1];
    const uint32_t recordTotal = ctx->hostScalars[2];
    const uint32_t firstLocation = d->first_path_tess_location ? d->first_path_tess_location : kPatchSpan;
    if (contourTotal > 0xffffu)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "more than 65535 contours in one flush");
    if (tessTotal > kMaxFlushTexels || recordTotal == kScanSaturated)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "flush too large: more than 2^31 tessellation vertices");
    RIVE_TRY(ensure(ctx, ctx->geoRecords, (size_t)std::max<uint32_t>(recordTotal, 1) * sizeof(CubicRec)));
    RIVE_TRY(ensure(ctx, ctx->geoContours, (size_t)std::max<uint32_t>(contourTotal, 1) * sizeof(ContourV)));
    if (n > 0)
        emit_kernel<<<(n + 31) / 32, 32, 0, ctx->stream>>>(draws, n, verbs, points, tessOffsets,
                                                              contourOffsets, recordOffsets, tessOffsets,
                                                              contourOffsets, firstLocation,
                                                              P<CubicRec>(ctx->geoRecords),
                                                              P<ContourV>(ctx->geoContours), scal + 13);
    HIP_TRY(hipGetLastError());
    uint32_t overflow = 0;
    rc = read_scalar(ctx, scal + 13, &overflow);
    if (rc)
        return rc;
    if (overflow != 0)
        return fail(RIVE_HIP_ERROR_DEVICE, "curve emission disagreed with the counting pass");
    // spans
    RIVE_TRY(ensure(ctx, ctx->geoSpanCounts, (size_t)(recordTotal + 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->geoSpanOffsets, (size_t)(recordTotal + 1) * 4));
    if (recordTotal > 0)
        span_count_kernel<<<(recordTotal + 255) / 256, 256, 0, ctx->stream>>>(
            P<CubicRec>(ctx->geoRecords), recordTotal, P<uint32_t>(ctx->geoSpanCounts));
    rc = device_scan(ctx, P<uint32_t>(ctx->geoSpanCounts), recordTotal, P<uint32_t>(ctx->geoSpanOffsets),
                     scal + 11);
    if (rc)
        return rc;
    uint32_t spanTotal = 0;
    rc = read_total(ctx, scal + 11, &spanTotal, "tessellation spans");
    if (rc)
        return rc;
    const uint32_t allSpans = spanTotal + 2; // + padding spans
    RIVE_TRY(ensure(ctx, ctx->geoSpans, (size_t)allSpans * sizeof(SpanV)));
    padding_span_kernel<<<1, 64, 0, ctx->stream>>>(P<SpanV>(ctx->geoSpans), scal + 8, firstLocation);
    FrameBounds oneFlush{};
    oneFlush.n = 1;
    oneFlush.v[0] = 0;
    oneFlush.v[1] = recordTotal;
    if (recordTotal > 0)
        span_write_kernel<<<(recordTotal + 255) / 256, 256, 0, ctx->stream>>>(
            P<CubicRec>(ctx->geoRecords), recordTotal, P<uint32_t>(ctx->geoSpanOffsets), oneFlush,
            P<SpanV>(ctx->geoSpans));
    HIP_TRY(hipGetLastError());
    RIVE_TRY(mark(ctx, ctx->ev[1]));
    const uint32_t texelsUsed = firstLocation + tessTotal + 1u;
    *tessRows = std::max<uint32_t>((texelsUsed + kTessTextureWidth - 1) / kTessTextureWidth, 1);
    *spans = P<SpanV>(ctx->geoSpans);
    *spanCount = allSpans;
    *contours = P<ContourV>(ctx->geoContours);
    ctx->lastContourCount = contourTotal;
    ctx->lastRecordCount = recordTotal;
    ctx->lastTessVertexCount = texelsUsed;
    return RIVE_HIP_OK;
}
} // namespace

namespace
{
// One logical flush's layout from a frame's counting pass.
struct GeoLayout
{
    const rivehip::SpanV* spans;
    uint32_t spanCount;
    const rivehip::ContourV* contours;
    uint32_t contourCount;
    uint32_t recordCount;
    uint32_t texelsUsed; // incl. the 8 + 1 padding texels
    uint32_t tessRows;
};

// GEOMETRY-mode front half for several logical flushes whose draws are
// contiguous in the DRAW buffer (descs[k].first_draw follow each other):
// one counting launch, frame-wide scans, one emission launch and one span
// pass, each flush laid out from location firstLocation and contour ID 1 on
// its own, exactly as n separate flushes would be.
int geometry_counting_frame(rive_hip_context* ctx, const rive_hip_flush_desc* descs, uint32_t nf,
                            GeoLayout* out)
{
    using namespace rivehip;
    const uint64_t firstDraw = descs[0].first_draw;
    FrameBounds drawBounds{};
    drawBounds.n = nf;
    uint64_t n64 = 0;
    for (uint32_t k = 0; k < nf; ++k)
    {
        drawBounds.v[k] = (uint32_t)n64;
        n64 += descs[k].draw_count;
    }
    drawBounds.v[nf] = (uint32_t)n64;
    const uint32_t n = (uint32_t)n64;
    auto* draws = reinterpret_cast<rive_hip_draw*>(ctx->bufs[RIVE_HIP_BUFFER_DRAW].dev) + firstDraw;
    const auto* verbs = reinterpret_cast<const uint8_t*>(ctx->bufs[RIVE_HIP_BUFFER_VERB].dev);
    const auto* points = reinterpret_cast<const V2d*>(ctx->bufs[RIVE_HIP_BUFFER_POINT].dev);
    if (n > 0 && (verbs == nullptr || points == nullptr))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "GEOMETRY flush without VERB/POINT buffers");
    const uint32_t firstLocation = descs[0].first_path_tess_location ? descs[0].first_path_tess_location : kPatchSpan;
    RIVE_TRY(ensure(ctx, ctx->scalars, kScalarBytes));
    uint32_t* scal = P<uint32_t>(ctx->scalars);
    RIVE_TRY(clear(ctx, ctx->scalars.ptr, 0, kScalarBytes));
    // 5 arrays of n+1 (tess, contour, record counts; tess, contour locals) +
    // 3 scans of n+1 + bounds values.
    RIVE_TRY(ensure(ctx, ctx->geoCounts, ((size_t)8 * (n + 1) + 4 * (kMaxFrameFlushes + 1) + 16) * 4));
    uint32_t* tessCounts = P<uint32_t>(ctx->geoCounts);
    uint32_t* contourCounts = tessCounts + (n + 1);
    uint32_t* recordCounts = contourCounts + (n + 1);
    uint32_t* tessOffsets = recordCounts + (n + 1);
    uint32_t* contourOffsets = tessOffsets + (n + 1);
    uint32_t* recordOffsets = contourOffsets + (n + 1);
    uint32_t* tessLocal = recordOffsets + (n + 1);
    uint32_t* contourLocal = tessLocal + (n + 1);
    uint32_t* boundVals = contourLocal + (n + 1);
    if (n > 0)
        count_kernel<<<(n + 31) / 32, 32, 0, ctx->stream>>>(
            draws, n, verbs, ctx->bufs[RIVE_HIP_BUFFER_VERB].size, points,
            ctx->bufs[RIVE_HIP_BUFFER_POINT].size / sizeof(V2d), scal + 12, tessCounts, contourCounts,
            recordCounts);
    HIP_TRY(hipGetLastError());
    RIVE_TRY(device_scan(ctx, tessCounts, n, tessOffsets, scal + 8));
    RIVE_TRY(device_scan(ctx, contourCounts, n, contourOffsets, scal + 9));
    RIVE_TRY(device_scan(ctx, recordCounts, n, recordOffsets, scal + 10));
    frame_bounds_kernel<<<1, kMaxFrameFlushes + 1, 0, ctx->stream>>>(drawBounds, tessOffsets, contourOffsets,
                                                                     recordOffsets, boundVals);
    HIP_TRY(hipGetLastError());
    // One readback: the totals, the bounds and the range check flag.
    const uint32_t m = nf + 1;
    std::vector<uint32_t> hv(3 * m + 8);
    RIVE_TRY(fetch(ctx, hv.data(), boundVals, (size_t)3 * m * 4));
    RIVE_TRY(fetch(ctx, hv.data() + 3 * m, scal + 8, 20));
    RIVE_TRY(settle(ctx));
    if (hv[3 * m + 4] != 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT,
                    "draw record with verb/point ranges outside the VERB/POINT buffers");
    const uint32_t* tb = hv.data();
    const uint32_t* cb = hv.data() + m;
    const uint32_t* rb = hv.data() + 2 * m;
    if (hv[3 * m + 2] == kScanSaturated)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "flush too large: records exceed 2^32-1");
    FrameBounds recordBounds{};
    FrameBounds tessTotals{};
    recordBounds.n = tessTotals.n = nf;
    for (uint32_t k = 0; k < nf; ++k)
    {
        const uint32_t tess = tb[k + 1] - tb[k], contoursK = cb[k + 1] - cb[k];
        if (contoursK > 0xffffu)
            return fail(RIVE_HIP_ERROR_UNSUPPORTED, "more than 65535 contours in one flush");
        if (tess > kMaxFlushTexels || tb[k + 1] < tb[k])
            return fail(RIVE_HIP_ERROR_UNSUPPORTED, "flush too large: more than 2^31 tessellation vertices");
        recordBounds.v[k] = rb[k];
        tessTotals.v[k] = tess;
        out[k].contourCount = contoursK;
        out[k].recordCount = rb[k + 1] - rb[k];
        out[k].texelsUsed = firstLocation + tess + 1u;
        out[k].tessRows = std::max<uint32_t>((out[k].texelsUsed + kTessTextureWidth - 1) / kTessTextureWidth, 1);
    }
    recordBounds.v[nf] = rb[nf];
    const uint32_t recordTotal = rb[nf], contourTotal = cb[nf];
    RIVE_TRY(ensure(ctx, ctx->geoRecords, (size_t)std::max<uint32_t>(recordTotal, 1) * sizeof(CubicRec)));
    RIVE_TRY(ensure(ctx, ctx->geoContours, (size_t)std::max<uint32_t>(contourTotal, 1) * sizeof(ContourV)));
    if (n > 0)
    {
        frame_local_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, drawBounds, tessOffsets, contourOffsets,
                                                                     tessLocal, contourLocal);
        emit_kernel<<<(n + 31) / 32, 32, 0, ctx->stream>>>(draws, n, verbs, points, tessOffsets, contourOffsets,
                                                              recordOffsets, tessLocal, contourLocal, firstLocation,
                                                              P<CubicRec>(ctx->geoRecords),
                                                              P<ContourV>(ctx->geoContours), scal + 13);
    }
    HIP_TRY(hipGetLastError());
    // spans
    RIVE_TRY(ensure(ctx, ctx->geoSpanCounts, (size_t)(recordTotal + 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->geoSpanOffsets, (size_t)(recordTotal + 1) * 4));
    if (recordTotal > 0)
        span_count_kernel<<<(recordTotal + 255) / 256, 256, 0, ctx->stream>>>(
            P<CubicRec>(ctx->geoRecords), recordTotal, P<uint32_t>(ctx->geoSpanCounts));
    RIVE_TRY(device_scan(ctx, P<uint32_t>(ctx->geoSpanCounts), recordTotal, P<uint32_t>(ctx->geoSpanOffsets),
                         scal + 11));
    frame_bounds_kernel<<<1, kMaxFrameFlushes + 1, 0, ctx->stream>>>(recordBounds, P<uint32_t>(ctx->geoSpanOffsets),
                                                                     P<uint32_t>(ctx->geoSpanOffsets), nullptr,
                                                                     boundVals);
    HIP_TRY(hipGetLastError());
    RIVE_TRY(fetch(ctx, hv.data(), boundVals, (size_t)m * 4));
    RIVE_TRY(fetch(ctx, hv.data() + 3 * m, scal + 11, 12));
    RIVE_TRY(settle(ctx));
    if (hv[3 * m + 2] != 0) // scal[13]: emission overflow
        return fail(RIVE_HIP_ERROR_DEVICE, "curve emission disagreed with the counting pass");
    if (hv[3 * m] == kScanSaturated)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "flush too large: tessellation spans exceed 2^32-1");
    const uint32_t spanTotal = hv[3 * m];
    FrameBounds spanBase{};
    spanBase.n = nf;
    for (uint32_t k = 0; k < nf; ++k)
    {
        // flush k: its two padding spans, then its records' spans
        spanBase.v[k] = hv[k] + 2u * k;
        out[k].spanCount = hv[k + 1] - hv[k] + 2u;
    }
    RIVE_TRY(ensure(ctx, ctx->geoSpans, ((size_t)spanTotal + 2u * nf) * sizeof(SpanV)));
    frame_padding_kernel<<<nf, 64, 0, ctx->stream>>>(P<SpanV>(ctx->geoSpans), spanBase, tessTotals, firstLocation);
    if (recordTotal > 0)
        span_write_kernel<<<(recordTotal + 255) / 256, 256, 0, ctx->stream>>>(
            P<CubicRec>(ctx->geoRecords), recordTotal, P<uint32_t>(ctx->geoSpanOffsets), recordBounds,
            P<SpanV>(ctx->geoSpans));
    HIP_TRY(hipGetLastError());
    for (uint32_t k = 0; k < nf; ++k)
    {
        out[k].spans = P<SpanV>(ctx->geoSpans) + spanBase.v[k];
        out[k].contours = P<ContourV>(ctx->geoContours) + cb[k];
    }
    return RIVE_HIP_OK;
}
} // namespace

extern "C" int rive_hip_count_draws(rive_hip_context* ctx, uint32_t n, uint32_t* tessOut, uint32_t* contourOut)
{
    using namespace rivehip;
    if (ctx == nullptr || (n > 0 && (tessOut == nullptr || contourOut == nullptr)))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    if (n == 0)
        return RIVE_HIP_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    for (int k = 0; k < RIVE_HIP_BUFFER_KIND_COUNT; ++k)
        if (ctx->bufs[k].mapped)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "count with a mapped buffer");
    if (ctx->bufs[RIVE_HIP_BUFFER_DRAW].size < (size_t)n * sizeof(rive_hip_draw))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "DRAW buffer smaller than draw_count records");
    const auto* verbs = reinterpret_cast<const uint8_t*>(ctx->bufs[RIVE_HIP_BUFFER_VERB].dev);
    const auto* points = reinterpret_cast<const V2d*>(ctx->bufs[RIVE_HIP_BUFFER_POINT].dev);
    if (verbs == nullptr || points == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "count without VERB/POINT buffers");
    RIVE_TRY(ensure(ctx, ctx->scalars, kScalarBytes));
    RIVE_TRY(ensure(ctx, ctx->geoCounts, (size_t)(3 * (size_t)(n + 1) + 16) * 4));
    uint32_t* tessCounts = P<uint32_t>(ctx->geoCounts);
    uint32_t* contourCounts = tessCounts + (n + 1);
    uint32_t* recordCounts = contourCounts + (n + 1);
    uint32_t* scal = P<uint32_t>(ctx->scalars);
    RIVE_TRY(clear(ctx, scal + 12, 0, 4));
    count_kernel<<<(n + 31) / 32, 32, 0, ctx->stream>>>(
        reinterpret_cast<const rive_hip_draw*>(ctx->bufs[RIVE_HIP_BUFFER_DRAW].dev), n, verbs,
        ctx->bufs[RIVE_HIP_BUFFER_VERB].size, points, ctx->bufs[RIVE_HIP_BUFFER_POINT].size / sizeof(V2d),
        scal + 12, tessCounts, contourCounts, recordCounts);
    HIP_TRY(hipGetLastError());
    RIVE_TRY(fetch(ctx, ctx->hostScalars, scal + 12, 4));
    RIVE_TRY(fetch(ctx, tessOut, tessCounts, (size_t)n * 4));
    RIVE_TRY(fetch(ctx, contourOut, contourCounts, (size_t)n * 4));
    RIVE_TRY(settle(ctx));
    if (ctx->hostScalars[0] != 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT,
                    "draw record with verb/point ranges outside the VERB/POINT buffers");
    return RIVE_HIP_OK;
}
