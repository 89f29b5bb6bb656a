// Device-side counting (RIVE_HIP_FLUSH_GEOMETRY).  Placeholder until the
// counting kernels land.
#pragma once
namespace
{
int geometry_counting(rive_hip_context* ctx, const rive_hip_flush_desc* d, const rivehip::SpanV** spans,
                      uint32_t* spanCount, const rivehip::ContourV** contours, uint32_t* tessRows)
{
    (void)ctx;
    (void)d;
    (void)spans;
    (void)spanCount;
    (void)contours;
    (void)tessRows;
    return fail(RIVE_HIP_ERROR_UNSUPPORTED, "geometry mode not built yet");
}
} // namespace
