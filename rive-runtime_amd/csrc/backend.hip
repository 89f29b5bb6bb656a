// Host side of the gfx950 backend: the C ABI of include/rive_hip.h.
//
// One rive_hip_context owns one HIP stream.  Buffers follow the reference's
// BufferRing contract (renderer/include/rive/renderer/buffer_ring.hpp:22-80):
// map returns write-only pinned host memory, unmap enqueues the upload on the
// context stream.  rive_hip_flush enqueues the whole pipeline on that stream.
#include "../../include/rive_hip.h"
#include "device_common.hpp"
#include "raster_kernels.hip"
#include "image_kernels.hip"
#include "../host/image_decode.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <memory>
#include <vector>

using namespace rivehip;

namespace
{
thread_local std::string g_lastError;

int fail(int code, const std::string& msg)
{
    g_lastError = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do                                                                                       \
    {                                                                                        \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            return fail(RIVE_HIP_ERROR_DEVICE,                                               \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

struct DevBuf
{
    void* ptr = nullptr;
    size_t cap = 0;
    bool carved = false; // carved from the no-alloc scratch by the current flush
};

template <typename T> T* P(DevBuf& b) { return reinterpret_cast<T*>(b.ptr); }

// Bump allocator over the per-device no-alloc scratch array, active only
// inside rive_hip_flush_noalloc.
struct Arena
{
    bool active = false;
    unsigned char* base = nullptr;
    size_t used = 0, cap = 0;
};

// A mapped record buffer: one device copy and a ring of kRingSlots pinned
// host staging slots (the reference keeps a ring of 3 per buffer,
// gpu.hpp:77), so the host can write the next flush's records while the
// device still reads the previous ones.  Mapping waits only for the copy out
// of the slot it reuses.  `host` is the slot mapped last.
constexpr int kRingSlots = 3;
struct MappedBuffer
{
    void* dev = nullptr;
    void* host = nullptr;
    void* slots[kRingSlots] = {};
    hipEvent_t copied[kRingSlots] = {}; // recorded after each slot's upload
    int slot = 0;
    size_t size = 0; // requested size (exact, as the reference requires)
    size_t cap = 0;  // allocated bytes (device and each slot)
    bool mapped = false;
};

constexpr int kEvents = 7;
constexpr int kRasterCandidateCount = RIVE_HIP_RASTER_CANDIDATES;
// Bounds on one flush's records (checked before any size arithmetic).
constexpr uint32_t kMaxFlushPaths = 1u << 24;
constexpr uint32_t kMaxFlushTexels = 1u << 31;
// Image table entries staged at context creation (what a no-alloc flush may use).
constexpr uint32_t kNoAllocImages = 16;
// Event sets kept for the logical flushes of one rive_hip_flush_frame call.
constexpr uint32_t kMaxFrameEventSets = 64;
} // namespace

struct rive_hip_target
{
    rive_hip_context* ctx;
    uint32_t width, height;
    uint32_t* pixels;
};

struct rive_hip_context
{
    int device = 0;
    hipStream_t stream = nullptr;
    MappedBuffer bufs[RIVE_HIP_BUFFER_KIND_COUNT];
    // pipeline scratch (grow-only)
    DevBuf spanTexelCounts, spanTexelOffsets, scanPartials, scalars;
    DevBuf tess, prims, cells, cellCount, cellOffset, cellCursor, cellBackdrop, binned, cellPacked, pathDraws;
    DevBuf pathRaster, pathPaint, ramps, drawTess, verts;
    DevBuf geoCounts, geoRecords, geoContours, geoSpans, geoSpanCounts, geoSpanOffsets;
    uint32_t lastMode = 0, lastRecordCount = 0, lastTessVertexCount = 0;
    uint32_t* hostScalars = nullptr; // pinned
    DevBuf imageTable, pathImage;
    ImageDesc* hostImages = nullptr; // pinned staging of the image table
    uint32_t hostImagesCap = 0;
    DevBuf canvasTable, pathCanvas;
    DevBuf patchTable, patchWork; // SPANS flushes described by patch batches
    // primitives leaving the guard box: their list, pieces per draw, the
    // pieces' offsets and cursors, and the draw list with the pieces' draws
    DevBuf specialList, farList, drawPieces, drawPieceOffset, pieceCursor, drawTessAll;
    // the raster's tile order: cost class per tile, the order, and the
    // class histogram, class cursors and tile queue head (2 * 256 + 1 words)
    DevBuf tileClass, tileOrder, tileQueue;
    rive_hip_raster_config rasterConfig = {0u, 0u, 6u, 0u};
    bool interiorTriangulation = false; // the front end's layout of large fills
    int cuCount = 256;
    CanvasDev* hostCanvases = nullptr; // pinned staging of the canvas table
    uint32_t hostCanvasesCap = 0;
    std::vector<rive_hip_canvas> lastCanvases;
    // HIP events per flush: the set of each logical flush of the last
    // rive_hip_flush / rive_hip_flush_frame call (ev: the current flush's).
    std::vector<std::array<hipEvent_t, kEvents>> evSets;
    uint32_t evSetsUsed = 0;
    hipEvent_t* ev = nullptr;
    float passMs[6] = {};
    bool timed = false;
    // last flush sizes, for debug readback
    uint32_t lastSpanCount = 0, lastContourCount = 0, lastTexels = 0, lastDrawCount = 0;
    const void* lastSpans = nullptr;    // the last flush's TessVertexSpans (device)
    const void* lastContours = nullptr; // and ContourData
    uint32_t lastGradRows = 0;
    rive_hip_flush_stats stats = {};
    std::vector<rive_hip_canvas> flushCanvases; // the current flush's canvas table (capacity kept)
    // No-alloc flushes: this device's scratch array (device address) and the
    // context's own scratch buffers, set aside while a flush carves its own.
    unsigned char* scratchPtr = nullptr;
    Arena arena;
    bool lastFlushNoAlloc = false;
    // The current (or last) flush's records: device pointers at the
    // descriptor's first_* offsets, and the matching host staging.
    struct FlushInputs
    {
        const rivehip::PathDataV* paths = nullptr;
        const uint2* paints = nullptr;
        const float* aux = nullptr;
        const uint4* gradSpans = nullptr;
        rive_hip_draw* draws = nullptr;
        const rivehip::SpanV* spans = nullptr;
        const rivehip::ContourV* contours = nullptr;
        const rive_hip_paint_data* paintsHost = nullptr;
        const rive_hip_path_data* pathsHost = nullptr;
        const rive_hip_grad_span* gradSpansHost = nullptr;
    } in;
    // The descriptors of the last frame's logical flushes (a frame starts at
    // a flush whose records start at offset 0), with the image and canvas
    // tables they point at copied here.
    // Host reads of device totals during a flush.  Live: copy and wait.
    // Record: the same, keeping every value read on the tape.  Replay: no
    // device access, the values come off the tape in the same order (a frame
    // graph's capture re-runs the host logic of a recorded frame this way).
    enum class Reads
    {
        live,
        record,
        replay
    } reads = Reads::live;
    std::vector<uint8_t> tape;
    size_t tapePos = 0;
    std::vector<std::pair<void*, size_t>> pendingReads;
    bool capturing = false; // inside hipStreamBeginCapture: no allocation, no sync
    std::vector<void*> captureHeld; // pinned copies of tables the captured copies read
    std::vector<rive_hip_flush_desc> frameDescs;
    std::vector<std::vector<rive_hip_image*>> frameImages;
    std::vector<std::vector<rive_hip_canvas>> frameCanvases;
    size_t frameLen = 0; // flushes of the recorded frame (slots past it keep their storage)
    // The feather function tables (FEATHER, then INVERSE_FEATHER), uploaded
    // once at creation.
    float* featherTables = nullptr;
};

namespace
{
// The scratch buffers of a context (every DevBuf member), in one list.
constexpr int kDevBufCount = 40;
// The device scalars of a flush (totals, counters), 32 words: the pipeline
// uses slots 0-7, device counting 8-13, guard-box pieces kFarSlot .. +3.
constexpr size_t kScalarBytes = 128;
constexpr int kFarSlot = 16;
void devbufs(rive_hip_context* c, DevBuf** out)
{
    DevBuf* all[kDevBufCount] = {&c->spanTexelCounts, &c->spanTexelOffsets, &c->scanPartials, &c->scalars,
                                 &c->tess, &c->prims, &c->cells, &c->cellCount, &c->cellOffset,
                                 &c->cellCursor, &c->cellBackdrop, &c->binned, &c->pathRaster,
                                 &c->ramps, &c->drawTess, &c->geoCounts, &c->geoRecords,
                                 &c->geoContours, &c->geoSpans, &c->geoSpanCounts, &c->geoSpanOffsets,
                                 &c->verts, &c->imageTable, &c->pathImage, &c->canvasTable,
                                 &c->pathCanvas, &c->cellPacked, &c->pathDraws, &c->pathPaint,
                                 &c->patchTable, &c->patchWork, &c->specialList, &c->farList, &c->drawPieces,
                                 &c->drawPieceOffset, &c->pieceCursor, &c->drawTessAll, &c->tileClass,
                                 &c->tileOrder, &c->tileQueue};
    for (int i = 0; i < kDevBufCount; ++i)
        out[i] = all[i];
}

// Grows a scratch buffer to at least `bytes` (grow-only, 1.25x headroom).
// Inside a no-alloc flush it carves the buffer from the device's scratch
// array instead and never calls the allocator.
int ensure(rive_hip_context* ctx, DevBuf& b, size_t bytes)
{
    Arena& a = ctx->arena;
    if (a.active)
    {
        if (b.carved && bytes <= b.cap)
            return RIVE_HIP_OK;
        const size_t need = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
        if (need > a.cap - a.used)
            return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY,
                        "rive_hip_flush_noalloc: the flush needs more than the " +
                            std::to_string(a.cap >> 20) + " MiB no-alloc scratch (" +
                            std::to_string((a.used + need) >> 20) + " MiB so far)");
        b.ptr = a.base + a.used;
        b.cap = need;
        b.carved = true;
        a.used += need;
        return RIVE_HIP_OK;
    }
    if (bytes <= b.cap)
        return RIVE_HIP_OK;
    if (ctx->capturing)
        return fail(RIVE_HIP_ERROR_DEVICE, "a captured frame needs a buffer its recording did not size");
    if (b.ptr != nullptr)
    {
        hipError_t e = hipFree(b.ptr);
        if (e != hipSuccess)
            return fail(RIVE_HIP_ERROR_DEVICE, std::string("hipFree: ") + hipGetErrorString(e));
        b.ptr = nullptr;
        b.cap = 0;
    }
    size_t cap = std::max<size_t>(bytes + bytes / 4, 256);
    hipError_t e = hipMalloc(&b.ptr, cap);
    if (e != hipSuccess)
    {
        b.ptr = nullptr;
        return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY,
                    "scratch allocation of " + std::to_string(cap) + " bytes failed: " + hipGetErrorString(e));
    }
    b.cap = cap;
    return RIVE_HIP_OK;
}

#define RIVE_TRY(expr)                                                                       \
    do                                                                                       \
    {                                                                                        \
        int _rc = (expr);                                                                    \
        if (_rc != RIVE_HIP_OK)                                                              \
            return _rc;                                                                      \
    } while (0)
} // namespace

namespace
{
// The no-alloc flush's scratch: one array per GPU, held until the process
// ends.  Host code never passes `scratch` itself to a kernel or a copy (on
// the host the name is a stand-in, not a device address); it uses the
// address hipGetSymbolAddress returns on each device, kept in g_scratch.
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
unsigned char* g_scratch[kMaxDevices] = {};
} // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace
{
// Looks up this device's scratch address once (the device must be current).
hipError_t scratch_address(int device, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    if (g_scratch[device] == nullptr)
    {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch[device] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch[device];
    return hipSuccess;
}
} // namespace

namespace
{
// The Gaussian-integral function tables of feathering, generated the way the
// reference generates the fp16 tables it ships (gpu.cpp:1984-2181), so the
// texel values are the same: a running integral of the normal distribution
// (mu = 256, sigma = 512/6) sampled 7 times per entry, centered on 1/2,
// monotonic, clamped and eased into 0 and 1 over 8 entries; its inverse from
// 32 samples per entry; both rounded to fp16 by adding half of the dropped
// mantissa and truncating (gpu.cpp:1959-1980).
float normal_density(float x, float mu, float inverseSigma)
{
    const float y = (x - mu) * inverseSigma;
    return expf((float)(-.5 * (double)y * (double)y)) * inverseSigma * 0.398942280401433f;
}
float to_f16_and_back(float x)
{
    uint32_t b;
    memcpy(&b, &x, 4);
    b += 0x1000u;
    const uint32_t e = (b & 0x7F800000u) >> 23, m = b & 0x007FFFFFu;
    uint32_t h = (b & 0x80000000u) >> 16;
    if (e > 112u)
        h |= (((e - 112u) << 10) & 0x7C00u) | m >> 13;
    if (e < 113u && e > 101u)
        h |= (((0x007FF000u + m) >> (125u - e)) + 1u) >> 1;
    if (e > 143u)
        h |= 0x7FFFu;
    const uint32_t he = (h >> 10) & 0x1fu, hm = h & 0x3ffu;
    float v = he == 0 ? ldexpf((float)hm, -24) : ldexpf((float)(hm | 0x400u), (int)he - 25);
    return (h & 0x8000u) ? -v : v;
}
void feather_tables_host(float* out) // 2 x kFeatherTableSize
{
    const uint32_t N = kFeatherTableSize;
    const float sigma = N / (3.f * 2), inverseSigma = 1 / sigma, mu = N * .5f;
    std::vector<float> t(N);
    float integral = 0;
    for (uint32_t i = 0; i < N; ++i)
    {
        for (int k = 0; k < 7; ++k)
            integral += normal_density((float)i + (float)(k - 3) / 7, mu, inverseSigma) / 7;
        t[i] = integral;
    }
    const float shift = (float)(.5 - (t[N / 2 - 1] + t[N / 2]) / 2);
    t[0] = fminf(fmaxf(0, t[0] + shift), 1);
    for (uint32_t i = 1; i < N; ++i)
        t[i] = fminf(fmaxf(t[i - 1], t[i] + shift), 1);
    for (uint32_t i = 0; i < N; ++i)
    {
        if (i < 8)
            t[i] *= (float)i / 8.f;
        if (i > N - 8 - 1)
        {
            const float diffToOne = (float)(1.0 - t[i]);
            t[i] = t[i] + diffToOne * ((float)(i - (N - 8) + 1) / 8.f);
        }
    }
    for (uint32_t i = 0; i < N; ++i)
        out[i] = to_f16_and_back(t[i]);
    // the inverse
    const uint32_t samples = N * 32;
    integral = 0;
    for (uint32_t i = 0; i < (samples + 1) / 2; ++i)
        integral += normal_density((float)i / 32, mu, inverseSigma) / 32;
    integral = (float)(.5 - integral);
    std::fill(t.begin(), t.end(), 0.f);
    t[N - 1] = 1;
    float lastX = NAN, lastY = 0;
    for (uint32_t i = 0; i < samples; ++i)
    {
        integral += normal_density((float)i / 32, mu, inverseSigma) / 32;
        const float x = fminf(fmaxf(0, integral), 1) * N;
        const float y = (i + .5f) / samples;
        const uint32_t cell = (uint32_t)x;
        const float center = cell + .5f;
        if (center == mu)
            t[cell] = .5f;
        else if (lastX <= center && x >= center)
        {
            const float w = (center - lastX) / (x - lastX);
            t[cell] = lastY * (1 - w) + y * w;
        }
        lastX = x;
        lastY = y;
    }
    for (uint32_t i = 0; i < N; ++i)
        out[N + i] = to_f16_and_back(t[i]);
}
} // namespace

extern "C" {

const char* rive_hip_last_error(void) { return g_lastError.c_str(); }

int rive_hip_create(int device, rive_hip_context** out)
{
    if (out == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "out is null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(RIVE_HIP_ERROR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device));
    auto* ctx = new rive_hip_context;
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&ctx->hostScalars, 64, hipHostMallocDefault);
    // Staging a no-alloc flush may use without growing it: one canvas and
    // kNoAllocImages images.
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&ctx->hostCanvases, sizeof(CanvasDev), hipHostMallocDefault);
    if (e == hipSuccess)
    {
        ctx->hostCanvasesCap = 1;
        e = hipHostMalloc((void**)&ctx->hostImages, sizeof(ImageDesc) * kNoAllocImages, hipHostMallocDefault);
    }
    if (e == hipSuccess)
    {
        ctx->hostImagesCap = kNoAllocImages;
        ctx->flushCanvases.reserve(1);
        ctx->lastCanvases.reserve(1);
        e = scratch_address(device, &ctx->scratchPtr);
    }
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->cuCount, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess)
        e = hipMalloc((void**)&ctx->featherTables, sizeof(float) * 2 * kFeatherTableSize);
    if (e == hipSuccess)
    {
        std::vector<float> tables(2 * kFeatherTableSize);
        feather_tables_host(tables.data());
        e = hipMemcpy(ctx->featherTables, tables.data(), sizeof(float) * tables.size(), hipMemcpyHostToDevice);
    }
    ctx->evSets.reserve(kMaxFrameEventSets);
    ctx->evSets.emplace_back();
    ctx->evSets[0].fill(nullptr);
    for (int i = 0; i < kEvents && e == hipSuccess; ++i)
        e = hipEventCreate(&ctx->evSets[0][i]);
    ctx->ev = ctx->evSets[0].data();
    if (e != hipSuccess)
    {
        delete ctx;
        return fail(RIVE_HIP_ERROR_DEVICE, std::string("context setup: ") + hipGetErrorString(e));
    }
    *out = ctx;
    return RIVE_HIP_OK;
}

void rive_hip_destroy(rive_hip_context* ctx)
{
    if (ctx == nullptr)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->bufs)
    {
        if (b.dev)
            (void)hipFree(b.dev);
        for (int k = 0; k < kRingSlots; ++k)
        {
            if (b.slots[k])
                (void)hipHostFree(b.slots[k]);
            if (b.copied[k])
                (void)hipEventDestroy(b.copied[k]);
        }
    }
    DevBuf* all[kDevBufCount];
    devbufs(ctx, all);
    for (DevBuf* b : all)
        if (b->ptr && !b->carved)
            (void)hipFree(b->ptr);
    if (ctx->hostScalars)
        (void)hipHostFree(ctx->hostScalars);
    if (ctx->featherTables)
        (void)hipFree(ctx->featherTables);
    if (ctx->hostImages)
        (void)hipHostFree(ctx->hostImages);
    if (ctx->hostCanvases)
        (void)hipHostFree(ctx->hostCanvases);
    for (auto& set : ctx->evSets)
        for (hipEvent_t e : set)
            if (e)
                (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int rive_hip_buffer_resize(rive_hip_context* ctx, int kind, size_t bytes)
{
    if (ctx == nullptr || kind < 0 || kind >= RIVE_HIP_BUFFER_KIND_COUNT)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "bad buffer kind");
    HIP_TRY(hipSetDevice(ctx->device));
    MappedBuffer& b = ctx->bufs[kind];
    if (b.mapped)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "resize while mapped");
    if (bytes <= b.cap)
    {
        // the allocation already holds it: only the size changes
        b.size = bytes;
        return RIVE_HIP_OK;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (b.dev)
        HIP_TRY(hipFree(b.dev));
    for (int k = 0; k < kRingSlots; ++k)
        if (b.slots[k])
        {
            HIP_TRY(hipHostFree(b.slots[k]));
            b.slots[k] = nullptr;
        }
    b.dev = b.host = nullptr;
    b.size = b.cap = 0;
    HIP_TRY(hipMalloc(&b.dev, bytes));
    for (int k = 0; k < kRingSlots; ++k)
    {
        HIP_TRY(hipHostMalloc(&b.slots[k], bytes, hipHostMallocDefault));
        if (!b.copied[k])
            HIP_TRY(hipEventCreateWithFlags(&b.copied[k], hipEventDisableTiming));
    }
    b.slot = 0;
    b.host = b.slots[0];
    b.size = b.cap = bytes;
    return RIVE_HIP_OK;
}

void* rive_hip_buffer_map(rive_hip_context* ctx, int kind, size_t bytes)
{
    if (ctx == nullptr || kind < 0 || kind >= RIVE_HIP_BUFFER_KIND_COUNT)
    {
        fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "bad buffer kind");
        return nullptr;
    }
    MappedBuffer& b = ctx->bufs[kind];
    if (bytes > b.size || b.host == nullptr)
    {
        fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "map exceeds buffer size");
        return nullptr;
    }
    // The next ring slot: its previous upload must have left it before the
    // host overwrites it.
    const int next = (b.slot + 1) % kRingSlots;
    if (hipEventSynchronize(b.copied[next]) != hipSuccess)
    {
        fail(RIVE_HIP_ERROR_DEVICE, "staging slot wait failed");
        return nullptr;
    }
    b.slot = next;
    b.host = b.slots[next];
    b.mapped = true;
    return b.host;
}

int rive_hip_buffer_unmap(rive_hip_context* ctx, int kind, size_t bytes)
{
    if (ctx == nullptr || kind < 0 || kind >= RIVE_HIP_BUFFER_KIND_COUNT)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "bad buffer kind");
    MappedBuffer& b = ctx->bufs[kind];
    if (!b.mapped)
        return fail(RIVE_HIP_ERROR_NOT_MAPPED, "buffer is not mapped");
    b.mapped = false;
    if (bytes > b.size)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "unmap exceeds buffer size");
    if (bytes > 0)
        HIP_TRY(hipMemcpyAsync(b.dev, b.host, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipEventRecord(b.copied[b.slot], ctx->stream));
    return RIVE_HIP_OK;
}

int rive_hip_target_create(rive_hip_context* ctx, uint32_t width, uint32_t height,
                           rive_hip_target** out)
{
    if (ctx == nullptr || out == nullptr || width == 0 || height == 0 || width > 16384 ||
        height > 16384)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "bad target size");
    HIP_TRY(hipSetDevice(ctx->device));
    auto* t = new rive_hip_target{ctx, width, height, nullptr};
    hipError_t e = hipMalloc(&t->pixels, (size_t)width * height * 4);
    if (e != hipSuccess)
    {
        delete t;
        return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY, "target allocation failed");
    }
    HIP_TRY(hipMemsetAsync(t->pixels, 0, (size_t)width * height * 4, ctx->stream));
    *out = t;
    return RIVE_HIP_OK;
}

void rive_hip_target_destroy(rive_hip_target* t)
{
    if (t == nullptr)
        return;
    (void)hipSetDevice(t->ctx->device);
    (void)hipStreamSynchronize(t->ctx->stream);
    (void)hipFree(t->pixels);
    delete t;
}

struct rive_hip_image
{
    rive_hip_context* ctx;
    rivehip::ImageDesc desc;
};

} // extern "C"

namespace
{
// An image of width x height with its mip levels allocated (full chain when
// mipLevels is 0), texels unset.
int image_alloc(rive_hip_context* ctx, uint32_t width, uint32_t height, uint32_t mipLevels, rive_hip_image** out)
{
    using namespace rivehip;
    uint32_t full = 1;
    while ((std::max(width, height) >> full) != 0)
        ++full;
    uint32_t levels = mipLevels == 0 ? full : std::min(mipLevels, full);
    levels = std::min<uint32_t>(levels, kMaxImageLevels);
    ImageDesc desc{};
    desc.width = width;
    desc.height = height;
    desc.levels = levels;
    uint32_t total = 0;
    for (uint32_t l = 0; l < levels; ++l)
    {
        desc.levelOffset[l] = total;
        total += std::max(width >> l, 1u) * std::max(height >> l, 1u);
    }
    desc.levelOffset[levels] = total;
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t* texels = nullptr;
    if (hipMalloc(&texels, (size_t)total * 4) != hipSuccess)
        return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY, "image allocation failed");
    desc.texels = texels;
    *out = new rive_hip_image{ctx, desc};
    return RIVE_HIP_OK;
}

// Enqueues levels 1.. from level 0 (2x2 box filter).
hipError_t image_mips(rive_hip_context* ctx, rive_hip_image* img)
{
    const rivehip::ImageDesc& d = img->desc;
    uint32_t* texels = const_cast<uint32_t*>(d.texels);
    hipError_t e = hipSuccess;
    for (uint32_t l = 0; l + 1 < d.levels && e == hipSuccess; ++l)
    {
        const int sw = (int)std::max(d.width >> l, 1u), sh = (int)std::max(d.height >> l, 1u);
        const int dw = (int)std::max(d.width >> (l + 1), 1u), dh = (int)std::max(d.height >> (l + 1), 1u);
        mip_kernel<<<dim3((dw + 63) / 64, dh), 64, 0, ctx->stream>>>(texels, d.levelOffset[l], sw, sh,
                                                                      d.levelOffset[l + 1], dw, dh);
        e = hipGetLastError();
    }
    return e;
}

void image_free(rive_hip_image* img)
{
    (void)hipFree(const_cast<uint32_t*>(img->desc.texels));
    delete img;
}

// Device scratch of a decode, freed on every exit.
struct DecodeTemps
{
    std::vector<void*> ptrs;
    ~DecodeTemps()
    {
        for (void* p : ptrs)
            (void)hipFree(p);
    }
    template <typename T> T* alloc(size_t bytes)
    {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess)
            return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

// Level 0 of `img` from a decoded PNG (unpremultiplied RGBA8 on the host).
int decode_png_level0(rive_hip_context* ctx, const std::vector<uint8_t>& rgba, rive_hip_image* img, DecodeTemps& t)
{
    const size_t n = (size_t)img->desc.width * img->desc.height;
    uint32_t* staged = t.alloc<uint32_t>(n * 4);
    if (staged == nullptr)
        return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY, "image decode staging allocation failed");
    HIP_TRY(hipMemcpyAsync(staged, rgba.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
    premultiply_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, ctx->stream>>>(
        staged, const_cast<uint32_t*>(img->desc.texels), n);
    HIP_TRY(hipGetLastError());
    return RIVE_HIP_OK;
}

// Level 0 of `img` from an entropy-decoded JPEG: per component the inverse
// DCT into a sample plane, then colour conversion into the texels.
int decode_jpeg_level0(rive_hip_context* ctx, const rivehip::image::Jpeg& j, rive_hip_image* img, DecodeTemps& t)
{
    static std::once_flag basisOnce[16];
    static hipError_t basisErr[16];
    const int dev = std::min(std::max(ctx->device, 0), 15);
    std::call_once(basisOnce[dev], [&] {
        double basis[24][8];
        for (int r = 0; r < 24; ++r)
        {
            const int n = r < 8 ? 8 : 16, x = r < 8 ? r : r - 8;
            for (int u = 0; u < 8; ++u)
                basis[r][u] = (u == 0 ? 0.70710678118654752440 : 1.0) * std::cos((2 * x + 1) * u * M_PI / (2 * n));
        }
        basisErr[dev] = hipMemcpyToSymbol(HIP_SYMBOL(kIdctBasis), basis, sizeof(basis));
    });
    HIP_TRY(basisErr[dev]);
    int16_t* coef = t.alloc<int16_t>(j.coef.size() * 2);
    uint16_t* qt = t.alloc<uint16_t>(sizeof(j.qt));
    if (coef == nullptr || qt == nullptr)
        return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY, "JPEG coefficient allocation failed");
    HIP_TRY(hipMemcpyAsync(coef, j.coef.data(), j.coef.size() * 2, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(qt, j.qt, sizeof(j.qt), hipMemcpyHostToDevice, ctx->stream));
    JpegColor color{};
    color.components = (uint32_t)j.comps.size();
    color.ycc = j.ycc ? 1u : 0u;
    color.width = j.width;
    color.height = j.height;
    color.out = const_cast<uint32_t*>(img->desc.texels);
    for (size_t k = 0; k < j.comps.size(); ++k)
    {
        const rivehip::image::JpegComponent& c = j.comps[k];
        JpegPlane p{};
        p.coef = coef + c.coefOffset;
        p.qt = qt + 64 * c.tq;
        p.blocksW = c.blocksW;
        p.blocksH = c.blocksH;
        p.scaleH = c.scaleH;
        p.scaleV = c.scaleV;
        p.stride = c.blocksW * 8u * c.scaleH;
        p.samples = t.alloc<uint8_t>((size_t)p.stride * c.blocksH * 8u * c.scaleV);
        if (p.samples == nullptr)
            return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY, "JPEG sample plane allocation failed");
        const uint32_t blocks = c.blocksW * c.blocksH;
        jpeg_idct_kernel<<<(blocks + 63) / 64, 64, 0, ctx->stream>>>(p);
        HIP_TRY(hipGetLastError());
        color.plane[k] = p.samples;
        color.stride[k] = p.stride;
        color.repH[k] = c.repH;
        color.repV[k] = c.repV;
    }
    jpeg_color_kernel<<<dim3((j.width + 255) / 256, j.height), 256, 0, ctx->stream>>>(color);
    HIP_TRY(hipGetLastError());
    return RIVE_HIP_OK;
}

int decode_error(const std::string& what, rivehip::image::Format f)
{
    return fail(f == rivehip::image::kWebp ? RIVE_HIP_ERROR_UNSUPPORTED : RIVE_HIP_ERROR_INVALID_ARGUMENT,
                "image decode: " + what);
}
} // namespace

extern "C" {

int rive_hip_image_create(rive_hip_context* ctx, uint32_t width, uint32_t height, uint32_t mipLevels,
                          const uint8_t* rgba, rive_hip_image** out)
{
    if (ctx == nullptr || out == nullptr || rgba == nullptr || width == 0 || height == 0 || width > 16384 ||
        height > 16384)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "bad image arguments");
    rive_hip_image* img = nullptr;
    RIVE_TRY(image_alloc(ctx, width, height, mipLevels, &img));
    hipError_t e = hipMemcpyAsync(const_cast<uint32_t*>(img->desc.texels), rgba, (size_t)width * height * 4,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = image_mips(ctx, img);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
    {
        image_free(img);
        return fail(RIVE_HIP_ERROR_DEVICE, hipGetErrorString(e));
    }
    *out = img;
    return RIVE_HIP_OK;
}

int rive_image_probe(const uint8_t* bytes, size_t n, uint32_t* format, uint32_t* width, uint32_t* height)
{
    if (bytes == nullptr || format == nullptr || width == nullptr || height == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    rivehip::image::Format f;
    std::string err;
    const bool ok = rivehip::image::probe(bytes, n, &f, width, height, &err);
    *format = f;
    return ok ? RIVE_HIP_OK : decode_error(err, f);
}

int rive_png_decode(const uint8_t* bytes, size_t n, uint8_t* rgba, size_t capacity)
{
    if (bytes == nullptr || rgba == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> px;
    std::string err;
    if (!rivehip::image::decode_png(bytes, n, &w, &h, &px, &err))
        return decode_error(err, rivehip::image::kPng);
    if (capacity < px.size())
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "rgba capacity is below width * height * 4");
    std::memcpy(rgba, px.data(), px.size());
    return RIVE_HIP_OK;
}

int rive_jpeg_decode_info(const uint8_t* bytes, size_t n, rive_jpeg_info* info, int16_t* coef, size_t capacity)
{
    if (bytes == nullptr || info == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    rivehip::image::Jpeg j;
    std::string err;
    const bool full = coef != nullptr;
    if (!rivehip::image::decode_jpeg(bytes, n, &j, false, &err))
        return decode_error(err, rivehip::image::kJpeg);
    if (full && capacity < j.coefCount)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "coefficient capacity is below coef_count");
    if (full && !rivehip::image::decode_jpeg(bytes, n, &j, true, &err))
        return decode_error(err, rivehip::image::kJpeg);
    *info = rive_jpeg_info{};
    info->width = j.width;
    info->height = j.height;
    info->component_count = (uint32_t)j.comps.size();
    info->ycc = j.ycc ? 1u : 0u;
    info->hmax = j.hmax;
    info->vmax = j.vmax;
    info->coef_count = j.coefCount;
    for (size_t k = 0; k < j.comps.size(); ++k)
    {
        const rivehip::image::JpegComponent& c = j.comps[k];
        info->components[k] = {c.id,      c.h,     c.v,       c.tq,      c.scaleH,        c.scaleV,
                               c.repH,    c.repV,  c.blocksW, c.blocksH, (uint64_t)c.coefOffset};
    }
    std::memcpy(info->quant, j.qt, sizeof(j.qt));
    if (full)
        std::memcpy(coef, j.coef.data(), j.coef.size() * 2);
    return RIVE_HIP_OK;
}

int rive_hip_image_decode(rive_hip_context* ctx, const uint8_t* bytes, size_t n, rive_hip_image** out)
{
    if (ctx == nullptr || bytes == nullptr || out == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    rivehip::image::Format f;
    uint32_t w = 0, h = 0;
    std::string err;
    if (!rivehip::image::probe(bytes, n, &f, &w, &h, &err))
        return decode_error(err, f);
    if (w > 16384 || h > 16384)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "image decode: " + std::to_string(w) + "x" + std::to_string(h) +
                                                    " exceeds the 16384-texel image limit");
    std::vector<uint8_t> rgba;
    rivehip::image::Jpeg j;
    if (f == rivehip::image::kPng ? !rivehip::image::decode_png(bytes, n, &w, &h, &rgba, &err)
                                  : !rivehip::image::decode_jpeg(bytes, n, &j, true, &err))
        return decode_error(err, f);
    rive_hip_image* img = nullptr;
    RIVE_TRY(image_alloc(ctx, w, h, 0, &img));
    int rc;
    {
        DecodeTemps temps;
        rc = f == rivehip::image::kPng ? decode_png_level0(ctx, rgba, img, temps)
                                       : decode_jpeg_level0(ctx, j, img, temps);
        hipError_t e = rc == RIVE_HIP_OK ? image_mips(ctx, img) : hipSuccess;
        // the temporaries are freed only once the stream is done with them
        const hipError_t es = hipStreamSynchronize(ctx->stream);
        if (rc == RIVE_HIP_OK && (e != hipSuccess || es != hipSuccess))
            rc = fail(RIVE_HIP_ERROR_DEVICE, hipGetErrorString(e != hipSuccess ? e : es));
    }
    if (rc != RIVE_HIP_OK)
    {
        image_free(img);
        return rc;
    }
    *out = img;
    return RIVE_HIP_OK;
}

int rive_hip_image_read(rive_hip_image* img, uint32_t level, uint8_t* rgba, size_t capacity, uint32_t* width,
                        uint32_t* height)
{
    if (img == nullptr || width == nullptr || height == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    if (level >= img->desc.levels)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "level beyond the image's mip chain");
    *width = std::max(img->desc.width >> level, 1u);
    *height = std::max(img->desc.height >> level, 1u);
    const size_t bytes = (size_t)*width * *height * 4;
    if (rgba == nullptr)
        return RIVE_HIP_OK;
    if (capacity < bytes)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "rgba capacity is below the level's size");
    HIP_TRY(hipSetDevice(img->ctx->device));
    HIP_TRY(hipStreamSynchronize(img->ctx->stream));
    HIP_TRY(hipMemcpy(rgba, img->desc.texels + img->desc.levelOffset[level], bytes, hipMemcpyDeviceToHost));
    return RIVE_HIP_OK;
}

void rive_hip_image_destroy(rive_hip_image* img)
{
    if (img == nullptr)
        return;
    (void)hipSetDevice(img->ctx->device);
    (void)hipStreamSynchronize(img->ctx->stream);
    (void)hipFree(const_cast<uint32_t*>(img->desc.texels));
    delete img;
}

int rive_hip_sync(rive_hip_context* ctx)
{
    if (ctx == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "ctx is null");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RIVE_HIP_OK;
}

int rive_hip_read_pixels(rive_hip_context* ctx, rive_hip_target* t, uint8_t* rgba)
{
    if (ctx == nullptr || t == nullptr || rgba == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemcpyAsync(rgba, t->pixels, (size_t)t->width * t->height * 4, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RIVE_HIP_OK;
}

} // extern "C"

namespace
{
uint32_t premul_clear(uint32_t argb)
{
    float a = float(argb >> 24) / 255.f;
    auto q = [](float v) { return (uint32_t)floorf(std::min(std::max(v, 0.f), 1.f) * 255.f + .5f); };
    float r = float((argb >> 16) & 0xff) / 255.f * a;
    float g = float((argb >> 8) & 0xff) / 255.f * a;
    float b = float(argb & 0xff) / 255.f * a;
    return q(r) | (q(g) << 8) | (q(b) << 16) | (q(a) << 24);
}

// Exclusive scan of n u32 values into out[0..n], out[n] = total.  partials and
// total live in ctx scratch.
int device_scan(rive_hip_context* ctx, const uint32_t* in, uint32_t n, uint32_t* out,
                uint32_t* totalSlot)
{
    uint32_t blocks = std::max<uint32_t>((n + kScanTile - 1) / kScanTile, 1);
    RIVE_TRY(ensure(ctx, ctx->scanPartials, (size_t)blocks * 4));
    scan_reduce_kernel<<<blocks, kScanBlock, 0, ctx->stream>>>(in, n, P<uint32_t>(ctx->scanPartials));
    scan_partials_kernel<<<1, 1024, 0, ctx->stream>>>(P<uint32_t>(ctx->scanPartials), blocks, totalSlot);
    scan_apply_kernel<<<blocks, kScanBlock, 0, ctx->stream>>>(in, n, P<uint32_t>(ctx->scanPartials),
                                                               totalSlot, out);
    HIP_TRY(hipGetLastError());
    return RIVE_HIP_OK;
}

// Grows a scratch buffer to at least `bytes`, keeping its first `keep` bytes.
int grow_keep(rive_hip_context* ctx, DevBuf& b, size_t keep, size_t bytes)
{
    if (bytes <= b.cap && (!ctx->arena.active || b.carved))
        return RIVE_HIP_OK;
    DevBuf n{};
    RIVE_TRY(ensure(ctx, n, bytes));
    if (keep > 0)
        HIP_TRY(hipMemcpyAsync(n.ptr, b.ptr, keep, hipMemcpyDeviceToDevice, ctx->stream));
    if (!ctx->arena.active && b.ptr != nullptr)
    {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipFree(b.ptr));
    }
    b = n;
    return RIVE_HIP_OK;
}

// Queues a read of device bytes into host memory; settle() completes the
// queued reads (see rive_hip_context::reads).
int fetch(rive_hip_context* ctx, void* hostDst, const void* devSrc, size_t bytes)
{
    if (ctx->reads != rive_hip_context::Reads::replay)
        HIP_TRY(hipMemcpyAsync(hostDst, devSrc, bytes, hipMemcpyDeviceToHost, ctx->stream));
    ctx->pendingReads.emplace_back(hostDst, bytes);
    return RIVE_HIP_OK;
}

int settle(rive_hip_context* ctx)
{
    if (ctx->reads == rive_hip_context::Reads::replay)
    {
        for (const auto& r : ctx->pendingReads)
        {
            if (ctx->tapePos + r.second > ctx->tape.size())
            {
                ctx->pendingReads.clear();
                return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "frame replay read past its recording");
            }
            memcpy(r.first, ctx->tape.data() + ctx->tapePos, r.second);
            ctx->tapePos += r.second;
        }
        ctx->pendingReads.clear();
        return RIVE_HIP_OK;
    }
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
    {
        ctx->pendingReads.clear();
        return fail(RIVE_HIP_ERROR_DEVICE, std::string("stream sync: ") + hipGetErrorString(e));
    }
    if (ctx->reads == rive_hip_context::Reads::record)
        for (const auto& r : ctx->pendingReads)
        {
            const uint8_t* b = static_cast<const uint8_t*>(r.first);
            ctx->tape.insert(ctx->tape.end(), b, b + r.second);
        }
    ctx->pendingReads.clear();
    return RIVE_HIP_OK;
}

// Fills `bytes` (a multiple of 4) of device memory with a repeated byte: a
// kernel rather than hipMemsetAsync, so that a captured frame holds only
// kernel and copy nodes.
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t value, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}
int clear(rive_hip_context* ctx, void* ptr, uint8_t byte, size_t bytes)
{
    if (bytes == 0)
        return RIVE_HIP_OK;
    if ((bytes & 3) != 0 || (reinterpret_cast<uintptr_t>(ptr) & 3) != 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "clear of an unaligned range");
    const size_t n = bytes / 4;
    const uint32_t v = 0x01010101u * byte;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    fill_u32_kernel<<<blocks, 256, 0, ctx->stream>>>(static_cast<uint32_t*>(ptr), v, n);
    HIP_TRY(hipGetLastError());
    return RIVE_HIP_OK;
}

// Copies a host table (canvas or image descriptors) to the device.  While a
// frame is captured, the copy reads a pinned snapshot the graph keeps, since
// the staging table is rewritten by every flush.
int upload_table(rive_hip_context* ctx, void* devDst, const void* hostSrc, size_t bytes)
{
    if (bytes == 0)
        return RIVE_HIP_OK;
    const void* src = hostSrc;
    if (ctx->capturing)
    {
        void* held = nullptr;
        HIP_TRY(hipHostMalloc(&held, bytes, hipHostMallocDefault));
        memcpy(held, hostSrc, bytes);
        ctx->captureHeld.push_back(held);
        src = held;
    }
    HIP_TRY(hipMemcpyAsync(devDst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return RIVE_HIP_OK;
}

// Records a pass-timing event, except while a frame is being captured into a
// graph (its passes are timed on eager flushes).
int mark(rive_hip_context* ctx, hipEvent_t ev)
{
    if (!ctx->capturing)
        HIP_TRY(hipEventRecord(ev, ctx->stream));
    return RIVE_HIP_OK;
}

int read_scalar(rive_hip_context* ctx, const uint32_t* devSlot, uint32_t* value)
{
    RIVE_TRY(fetch(ctx, ctx->hostScalars, devSlot, 4));
    RIVE_TRY(settle(ctx));
    *value = ctx->hostScalars[0];
    return RIVE_HIP_OK;
}

// A device total (scan or layout) as read back; totals that do not fit 32
// bits saturate on the device and fail the flush here, before any kernel
// indexes with them.
int read_total(rive_hip_context* ctx, const uint32_t* devSlot, uint32_t* value, const char* what)
{
    int rc = read_scalar(ctx, devSlot, value);
    if (rc)
        return rc;
    if (*value == kScanSaturated)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, std::string("flush too large: ") + what + " exceed 2^32-1");
    return RIVE_HIP_OK;
}

// ---- Per-path ranges for a SPANS flush described by its DrawBatch list
// (reference input).  Every TessVertexSpan names its contour, whose
// ContourData names the path, and covers one row's run of texels forward
// (and, for a double-sided fill, one mirrored run): a path's forward range is
// the union of its spans' forward runs, its mirrored range the union of their
// reflections.  Paths are drawn in path-ID (submission) order; the patch
// batches give the image paint of the paths whose patches they cover.
struct PatchBatchDev
{
    uint32_t firstPatch, endPatch;
    uint32_t image; // image index | sampler << 16
    uint32_t pad;
};

__global__ void span_ranges_kernel(const SpanV* __restrict__ spans, uint32_t n, const ContourV* __restrict__ contours,
                                   uint32_t contourCount, uint32_t pathCount, uint32_t texels,
                                   uint32_t* __restrict__ fwdMin, uint32_t* __restrict__ fwdMax,
                                   uint32_t* __restrict__ mirMin, uint32_t* __restrict__ mirMax,
                                   uint32_t* __restrict__ bad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const SpanV s = spans[i];
    const uint32_t cid = s.flags & CONTOUR_ID_MASK;
    if (cid == 0u)
        return; // padding
    if (cid > contourCount)
    {
        atomicAdd(bad, 1u);
        return;
    }
    const uint32_t path = contours[cid - 1].pathID;
    if (path == 0u || path > pathCount)
    {
        atomicAdd(bad, 1u);
        return;
    }
    const int32_t W = (int32_t)kTessTextureWidth;
    for (int rect = 0; rect < 2; ++rect)
    {
        const float yf = rect == 0 ? s.y : s.ry;
        if (!(yf >= 0.f) || yf >= (float)(texels / kTessTextureWidth))
            continue;
        const int32_t packed = rect == 0 ? s.x0x1 : s.rx0x1;
        const int32_t x0 = (int32_t)((uint32_t)packed << 16) >> 16, x1 = packed >> 16;
        const int32_t lo = max(min(x0, x1), 0), hi = min(max(x0, x1), W);
        if (lo >= hi)
            continue;
        const uint32_t row = (uint32_t)yf * kTessTextureWidth;
        const bool mirrored = x1 < x0;
        atomicMin(mirrored ? &mirMin[path] : &fwdMin[path], row + (uint32_t)lo);
        atomicMax(mirrored ? &mirMax[path] : &fwdMax[path], row + (uint32_t)hi);
    }
}

__global__ void span_draw_flags_kernel(uint32_t pathCount, const uint32_t* __restrict__ fwdMax,
                                       const uint32_t* __restrict__ mirMax, uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; // path i + 1
    if (i < pathCount)
        flag[i] = (fwdMax[i + 1] | mirMax[i + 1]) != 0u ? 1u : 0u;
}

// One DrawTess per drawn path (validated like draw_tess_kernel); cnt[k] = its
// forward texels for the primitive scan.
__global__ void span_draws_kernel(uint32_t pathCount, const uint32_t* __restrict__ flag,
                                  const uint32_t* __restrict__ drawIdx, const uint32_t* __restrict__ fwdMin,
                                  const uint32_t* __restrict__ fwdMax, const uint32_t* __restrict__ mirMin,
                                  const uint32_t* __restrict__ mirMax, const PathDataV* __restrict__ paths,
                                  const PatchBatchDev* __restrict__ batches, uint32_t nb, uint32_t imageCount,
                                  DrawTess* __restrict__ out, uint32_t* __restrict__ cnt,
                                  uint32_t* __restrict__ pathImage, uint32_t* __restrict__ bad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pathCount || flag[i] == 0u)
        return;
    const uint32_t path = i + 1, k = drawIdx[i];
    DrawTess t{};
    t.pathID = path;
    t.isStroke = paths[path].strokeRadius != 0.f ? 1u : 0u;
    const bool featheredFill = !t.isStroke && paths[path].featherRadius != 0.f;
    const uint32_t nf = fwdMax[path] > 0u ? fwdMax[path] - fwdMin[path] : 0u;
    const uint32_t nm = mirMax[path] > 0u ? mirMax[path] - mirMin[path] : 0u;
    t.forwardBase = nf != 0u ? fwdMin[path] : 0u;
    t.forwardCount = nf;
    t.tessVertexCount = nf + nm;
    bool ok;
    if (featheredFill && nf == 0u && nm != 0u)
    {
        // a feathered fill through the atlas under a left-handed matrix:
        // reverse only, read back to front
        t.dirs = 1u;
        t.tessLocation = mirMin[path];
        t.forwardBase = mirMin[path];
        t.forwardCount = nm;
        ok = true;
    }
    else if (nm != 0u)
    {
        // a double-sided fill: reverseThenForward when the mirrored run comes
        // first (draw.cpp:1899-1990)
        t.dirs = (nf == 0u || mirMin[path] < fwdMin[path]) ? 3u : 2u;
        t.tessLocation = nf != 0u ? min(fwdMin[path], mirMin[path]) : mirMin[path];
        ok = nf == nm && !t.isStroke;
    }
    else
    {
        t.dirs = 0u;
        t.tessLocation = t.forwardBase;
        // raster-ordering fills are double-sided, atlas feathers one way
        ok = t.isStroke != 0u || featheredFill;
    }
    // the batch holding the path's first patch supplies its image paint
    const uint32_t patch = t.tessLocation / 8u;
    for (uint32_t b = 0; b < nb; ++b)
        if (patch >= batches[b].firstPatch && patch < batches[b].endPatch)
        {
            const uint32_t img = batches[b].image;
            if ((img & 0xffffu) != 0u)
            {
                if ((img & 0xffffu) > imageCount || (img >> 16) >= 18u)
                    ok = false;
                else
                    pathImage[path] = img;
            }
            break;
        }
    if (!ok)
    {
        atomicAdd(bad, 1u);
        t.forwardCount = 0u;
        t.pathID = 0u;
    }
    out[k] = t;
    cnt[k] = t.forwardCount;
}

__global__ void draw_prim_base_kernel(uint32_t n, const uint32_t* __restrict__ base, DrawTess* __restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        out[k].primBase = base[k];
}

// Raster half of the pipeline, shared by both flush modes.  Inputs: the
// tessellation texture, contours, paths, paints and per-draw ranges.
int derive_patch_draws(rive_hip_context* ctx, const rive_hip_flush_desc* d, const SpanV* spans, uint32_t spanCount,
                       const ContourV* contours, uint32_t contourCount, uint32_t texels, uint32_t* drawCount,
                       uint32_t* totalPrims);

// The raster variant for a flush: feathered rows or not, persistent or one
// workgroup per tile, 5 or 6 waves per SIMD.
using RasterFn = decltype(&raster_kernel<false, false, 6>);
RasterFn raster_fn(bool feathered, bool persistent, int waves)
{
    static const RasterFn table[2][2][2] = {
        {{raster_kernel<false, false, 5>, raster_kernel<false, false, 6>},
         {raster_kernel<false, true, 5>, raster_kernel<false, true, 6>}},
        {{raster_kernel<true, false, 5>, raster_kernel<true, false, 6>},
         {raster_kernel<true, true, 5>, raster_kernel<true, true, 6>}}};
    return table[feathered ? 1 : 0][persistent ? 1 : 0][waves == 5 ? 0 : 1];
}

int raster_pipeline(rive_hip_context* ctx, const rive_hip_flush_desc* d, const SpanV* spans,
                    uint32_t spanCount, const ContourV* contours, uint32_t contourCount, uint32_t tessRows,
                    DrawTess* drawTess, uint32_t drawCount, uint32_t totalPrims)
{
    // The canvas table: the flush's canvases, or its one target.
    const uint32_t canvasCount = (uint32_t)ctx->lastCanvases.size();
    if (ctx->hostCanvasesCap < canvasCount)
    {
        if (ctx->arena.active)
            return fail(RIVE_HIP_ERROR_UNSUPPORTED, "rive_hip_flush_noalloc: more canvases than the staged table holds");
        if (ctx->hostCanvases)
            HIP_TRY(hipHostFree(ctx->hostCanvases));
        ctx->hostCanvases = nullptr;
        HIP_TRY(hipHostMalloc((void**)&ctx->hostCanvases, sizeof(CanvasDev) * canvasCount, hipHostMallocDefault));
        ctx->hostCanvasesCap = canvasCount;
    }
    uint32_t tileTotal = 0;
    int maxTilesY = 1;
    for (uint32_t c = 0; c < canvasCount; ++c)
    {
        const rive_hip_canvas& cv = ctx->lastCanvases[c];
        CanvasDev& cd = ctx->hostCanvases[c];
        cd = CanvasDev{};
        cd.pixels = cv.target->pixels;
        cd.width = (int)cv.target->width;
        cd.height = (int)cv.target->height;
        cd.tilesX = (cd.width + kTile - 1) / kTile;
        cd.tilesY = (cd.height + kTile - 1) / kTile;
        cd.tileBase = tileTotal;
        cd.firstPath = cv.first_path;
        cd.pathEnd = cv.first_path + cv.path_count;
        cd.clearPremul = premul_clear(cv.clear_color);
        cd.loadPreserve = cv.load_action == 1 ? 1u : 0u;
        tileTotal += (uint32_t)(cd.tilesX * cd.tilesY);
        maxTilesY = std::max(maxTilesY, cd.tilesY);
    }
    RIVE_TRY(ensure(ctx, ctx->canvasTable, sizeof(CanvasDev) * std::max<uint32_t>(canvasCount, 1)));
    RIVE_TRY(upload_table(ctx, ctx->canvasTable.ptr, ctx->hostCanvases, sizeof(CanvasDev) * canvasCount));
    uint32_t* scal = P<uint32_t>(ctx->scalars);
    const uint32_t pathCount = d->path_count;
    const PathDataV* paths = ctx->in.paths;
    const uint2* paints = ctx->in.paints;
    const float* aux = ctx->in.aux;
    const uint32_t gradRows = std::max<uint32_t>(d->grad_simple_rows + d->grad_complex_rows, 1);

    // ---- color ramps
    RIVE_TRY(mark(ctx, ctx->ev[0]));
    RIVE_TRY(ensure(ctx, ctx->ramps, (size_t)gradRows * kGradWidth * 4));
    if (d->grad_span_count > 0)
    {
        RIVE_TRY(clear(ctx, ctx->ramps.ptr, 0, (size_t)gradRows * kGradWidth * 4));
        ramp_kernel<<<d->grad_span_count, 64, 0, ctx->stream>>>(
            ctx->in.gradSpans,
            d->grad_span_count, gradRows, P<uint32_t>(ctx->ramps));
    }
    ctx->lastGradRows = gradRows;
    RIVE_TRY(mark(ctx, ctx->ev[2]));

    // ---- tessellation: texel counts -> scan -> one thread per texel
    const uint32_t texels = tessRows * kTessTextureWidth;
    RIVE_TRY(ensure(ctx, ctx->tess, (size_t)texels * sizeof(TessTexel)));
    RIVE_TRY(clear(ctx, ctx->tess.ptr, 0, (size_t)texels * sizeof(TessTexel)));
    if (spanCount > 0)
    {
        RIVE_TRY(ensure(ctx, ctx->spanTexelCounts, (size_t)spanCount * 4));
        RIVE_TRY(ensure(ctx, ctx->spanTexelOffsets, (size_t)(spanCount + 1) * 4));
        span_texel_count_kernel<<<(spanCount + 255) / 256, 256, 0, ctx->stream>>>(
            spans, spanCount, tessRows, P<uint32_t>(ctx->spanTexelCounts));
        int rc = device_scan(ctx, P<uint32_t>(ctx->spanTexelCounts), spanCount,
                             P<uint32_t>(ctx->spanTexelOffsets), scal + 0);
        if (rc)
            return rc;
        uint32_t used = 0;
        rc = read_total(ctx, scal + 0, &used, "tessellation texels");
        if (rc)
            return rc;
        if (used > 0)
            tessellate_kernel<<<(used + 255) / 256, 256, 0, ctx->stream>>>(
                spans, spanCount, P<uint32_t>(ctx->spanTexelOffsets), used, P<TessTexel>(ctx->tess),
                reinterpret_cast<const uint4*>(contours), contourCount, reinterpret_cast<const float*>(paths),
                pathCount, ctx->featherTables);
        ctx->lastTexels = texels;
    }
    HIP_TRY(hipGetLastError());
    if (d->mode == RIVE_HIP_FLUSH_SPANS && d->patch_batch_count > 0)
    {
        RIVE_TRY(derive_patch_draws(ctx, d, spans, spanCount, contours, contourCount, texels, &drawCount,
                                    &totalPrims));
        drawTess = P<DrawTess>(ctx->drawTess);
        ctx->stats.draw_count = drawCount;
        ctx->stats.prim_count = totalPrims;
    }
    RIVE_TRY(mark(ctx, ctx->ev[3]));

    // ---- expansion + binning
    RIVE_TRY(ensure(ctx, ctx->pathCanvas, (size_t)(pathCount + 1) * 4));
    RIVE_TRY(clear(ctx, ctx->pathCanvas.ptr, 0xff, (size_t)(pathCount + 1) * 4));
    path_canvas_kernel<<<canvasCount, 256, 0, ctx->stream>>>(P<CanvasDev>(ctx->canvasTable),
                                                             P<uint32_t>(ctx->pathCanvas));
    RIVE_TRY(ensure(ctx, ctx->cells, (size_t)(pathCount + 1) * sizeof(PathCells)));
    cells_init_kernel<<<(pathCount + 1 + 255) / 256, 256, 0, ctx->stream>>>(P<PathCells>(ctx->cells),
                                                                             pathCount + 1);
    RIVE_TRY(ensure(ctx, ctx->prims, (size_t)std::max<uint32_t>(totalPrims, 1) * sizeof(Prim)));
    RIVE_TRY(ensure(ctx, ctx->verts, (size_t)std::max<uint32_t>(totalPrims, 1) * sizeof(VtxPair)));
    if (totalPrims > 0)
        vertex_kernel<<<(totalPrims + 255) / 256, 256, 0, ctx->stream>>>(
            P<TessTexel>(ctx->tess), (int)texels, paths, drawTess, drawCount, totalPrims, P<VtxPair>(ctx->verts));
    RIVE_TRY(ensure(ctx, ctx->specialList, (size_t)std::max<uint32_t>(totalPrims, 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->farList, (size_t)std::max<uint32_t>(totalPrims, 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->drawPieces, (size_t)(drawCount + 1) * 4));
    RIVE_TRY(clear(ctx, ctx->drawPieces.ptr, 0, (size_t)(drawCount + 1) * 4));
    if (totalPrims > 0)
    {
        expand_kernel<<<(totalPrims + 255) / 256, 256, 0, ctx->stream>>>(
            P<TessTexel>(ctx->tess), (int)texels, P<VtxPair>(ctx->verts), contours, contourCount, paths,
            drawTess, drawCount, totalPrims, P<Prim>(ctx->prims), P<PathCells>(ctx->cells), scal + 5,
            P<uint32_t>(ctx->specialList), scal + kFarSlot + 3);
        // a grid-stride pass over the set-aside primitives (usually none)
        special_prims_kernel<<<64, 256, 0, ctx->stream>>>(
            P<TessTexel>(ctx->tess), (int)texels, contours, contourCount, paths, drawTess, drawCount,
            P<Prim>(ctx->prims), P<PathCells>(ctx->cells), P<uint32_t>(ctx->specialList), scal + kFarSlot + 3,
            P<uint32_t>(ctx->farList), scal + kFarSlot, P<uint32_t>(ctx->drawPieces));
    }
    cells_layout_kernel<<<1, 1024, 0, ctx->stream>>>(P<PathCells>(ctx->cells), pathCount + 1,
                                                     P<uint32_t>(ctx->pathCanvas), P<CanvasDev>(ctx->canvasTable),
                                                     scal + 1);
    HIP_TRY(hipGetLastError());
    // Cell total, the count of texels naming no ContourData record, and the
    // primitives leaving the guard box with their piece total.
    RIVE_TRY(fetch(ctx, ctx->hostScalars, scal + 1, 20));
    RIVE_TRY(fetch(ctx, ctx->hostScalars + 5, scal + kFarSlot, 8));
    RIVE_TRY(settle(ctx));
    const uint32_t totalCells = ctx->hostScalars[0];
    const uint32_t farPrims = ctx->hostScalars[5], pieces = ctx->hostScalars[6];
    if (ctx->hostScalars[4] != 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT,
                    "tessellated vertices name contour IDs outside 1..contour_count");
    if (totalCells == 0xffffffffu)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "flush too large: path cell grids exceed 2^32-1 cells");
    if ((uint64_t)totalPrims + pieces > 0xfffffff0ull)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "flush too large: primitives and guard-box pieces exceed 2^32");
    int rc = RIVE_HIP_OK;
    // Primitives leaving the guard box become their pieces, appended after
    // the flush's primitives, each draw's in a range binned as one more draw.
    const DrawTess* binDraws = drawTess;
    uint32_t binDrawCount = drawCount;
    ctx->stats.guard_pieces = pieces;
    if (pieces > 0)
    {
        RIVE_TRY(grow_keep(ctx, ctx->prims, (size_t)totalPrims * sizeof(Prim),
                           (size_t)(totalPrims + pieces) * sizeof(Prim)));
        RIVE_TRY(ensure(ctx, ctx->drawPieceOffset, (size_t)(drawCount + 1) * 4));
        RIVE_TRY(ensure(ctx, ctx->pieceCursor, (size_t)(drawCount + 1) * 4));
        RIVE_TRY(ensure(ctx, ctx->drawTessAll, (size_t)drawCount * 2 * sizeof(DrawTess)));
        RIVE_TRY(clear(ctx, ctx->pieceCursor.ptr, 0, (size_t)(drawCount + 1) * 4));
        rc = device_scan(ctx, P<uint32_t>(ctx->drawPieces), drawCount, P<uint32_t>(ctx->drawPieceOffset),
                         scal + kFarSlot + 2);
        if (rc)
            return rc;
        far_pieces_kernel<<<(farPrims + 255) / 256, 256, 0, ctx->stream>>>(
            P<Prim>(ctx->prims), P<uint32_t>(ctx->farList), scal + kFarSlot, drawTess, drawCount,
            P<uint32_t>(ctx->drawPieceOffset), P<uint32_t>(ctx->pieceCursor), totalPrims);
        piece_draws_kernel<<<(drawCount + 255) / 256, 256, 0, ctx->stream>>>(
            drawTess, drawCount, P<uint32_t>(ctx->drawPieces), P<uint32_t>(ctx->drawPieceOffset), totalPrims,
            P<DrawTess>(ctx->drawTessAll));
        HIP_TRY(hipGetLastError());
        binDraws = P<DrawTess>(ctx->drawTessAll);
        binDrawCount = drawCount * 2;
    }
    RIVE_TRY(ensure(ctx, ctx->cellCount, (size_t)(totalCells + 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->cellOffset, (size_t)(totalCells + 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->cellCursor, (size_t)(totalCells + 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->cellBackdrop, (size_t)(totalCells + 1) * 4));
    RIVE_TRY(clear(ctx, ctx->cellCount.ptr, 0, (size_t)(totalCells + 1) * 4));
    RIVE_TRY(clear(ctx, ctx->cellCursor.ptr, 0, (size_t)(totalCells + 1) * 4));
    RIVE_TRY(clear(ctx, ctx->cellBackdrop.ptr, 0, (size_t)(totalCells + 1) * 4));
    RIVE_TRY(ensure(ctx, ctx->pathDraws, (size_t)(pathCount + 1) * 4));
    RIVE_TRY(clear(ctx, ctx->pathDraws.ptr, 0, (size_t)(pathCount + 1) * 4));
    if (totalPrims > 0 && drawCount > 0)
        bin_count_draw_kernel<<<binDrawCount, 256, 0, ctx->stream>>>(
            P<Prim>(ctx->prims), binDraws, P<PathCells>(ctx->cells), P<uint32_t>(ctx->cellCount),
            P<int32_t>(ctx->cellBackdrop), P<uint32_t>(ctx->pathDraws));
    rc = device_scan(ctx, P<uint32_t>(ctx->cellCount), totalCells, P<uint32_t>(ctx->cellOffset), scal + 2);
    if (rc)
        return rc;
    RIVE_TRY(ensure(ctx, ctx->cellPacked, (size_t)(totalCells + 1) * sizeof(uint2)));
    {
        int maxRows = maxTilesY;
        const uint64_t threads = std::max<uint64_t>((uint64_t)(pathCount + 1) * (uint64_t)maxRows * 64u, 1);
        backdrop_prefix_kernel<<<(uint32_t)((threads + 255) / 256), 256, 0, ctx->stream>>>(
            P<PathCells>(ctx->cells), pathCount + 1, maxRows, P<int32_t>(ctx->cellBackdrop),
            P<uint32_t>(ctx->cellOffset), totalCells, P<uint2>(ctx->cellPacked));
    }
    uint32_t totalPairs = 0;
    rc = read_total(ctx, scal + 2, &totalPairs, "binned (primitive, tile) pairs");
    if (rc)
        return rc;
    ctx->stats.bin_entries = totalPairs;
    ctx->stats.prim_count = totalPrims;
    ctx->stats.tile_count = tileTotal;
    RIVE_TRY(ensure(ctx, ctx->binned, (size_t)std::max<uint32_t>(totalPairs, 1) * 4));
    if (totalPrims > 0 && drawCount > 0)
        bin_scatter_draw_kernel<<<binDrawCount, 256, 0, ctx->stream>>>(
            P<Prim>(ctx->prims), binDraws, P<PathCells>(ctx->cells), P<uint32_t>(ctx->cellOffset),
            P<uint32_t>(ctx->cellCursor), P<uint32_t>(ctx->pathDraws), P<uint32_t>(ctx->binned));
    RIVE_TRY(ensure(ctx, ctx->pathRaster, (size_t)(pathCount + 1) * sizeof(PathRaster)));
    RIVE_TRY(ensure(ctx, ctx->pathPaint, (size_t)(pathCount + 1) * sizeof(PathPaint)));
    path_raster_kernel<<<(pathCount + 1 + 255) / 256, 256, 0, ctx->stream>>>(
        paths, paints, aux, P<uint32_t>(ctx->pathImage), pathCount, gradRows, P<PathRaster>(ctx->pathRaster),
        P<PathPaint>(ctx->pathPaint));
    HIP_TRY(hipGetLastError());
    RIVE_TRY(mark(ctx, ctx->ev[4]));

    // ---- raster + resolve
    if (tileTotal > 0)
    {
        // The feather variant only when a path of the flush is feathered (or
        // the host copy of its records is not at hand).
        bool feathered = ctx->in.pathsHost == nullptr;
        for (uint32_t i = 1; !feathered && i <= d->path_count; ++i)
            feathered = ctx->in.pathsHost[i].feather_radius != 0.f;
        const rive_hip_raster_config& rc = ctx->rasterConfig;
        const bool persistent = rc.persistent != 0;
        const size_t queueBytes = (2 * kCostClasses + 1) * 4;
        RIVE_TRY(ensure(ctx, ctx->tileQueue, queueBytes));
        RIVE_TRY(clear(ctx, ctx->tileQueue.ptr, 0, queueBytes));
        uint32_t* queue = P<uint32_t>(ctx->tileQueue);
        const uint32_t* order = nullptr;
        if (rc.tile_order != 0)
        {
            RIVE_TRY(ensure(ctx, ctx->tileClass, (size_t)tileTotal * 4));
            RIVE_TRY(ensure(ctx, ctx->tileOrder, (size_t)tileTotal * 4));
            tile_cost_kernel<<<(tileTotal + 3) / 4, 256, 0, ctx->stream>>>(
                P<PathCells>(ctx->cells), P<uint2>(ctx->cellPacked), P<CanvasDev>(ctx->canvasTable), canvasCount,
                tileTotal, P<uint32_t>(ctx->tileClass), queue);
            tile_order_kernel<<<(tileTotal + 255) / 256, 256, 0, ctx->stream>>>(
                P<uint32_t>(ctx->tileClass), queue, queue + kCostClasses, tileTotal, P<uint32_t>(ctx->tileOrder));
            order = P<uint32_t>(ctx->tileOrder);
        }
        const int waves = rc.waves_per_simd == 5 ? 5 : 6;
        // a resident grid: one workgroup of four waves per SIMD's share
        const uint32_t grid =
            persistent ? std::min<uint32_t>(tileTotal, (uint32_t)std::max(ctx->cuCount, 1) * (uint32_t)waves)
                       : tileTotal;
        raster_fn(feathered, persistent, waves)<<<grid, kWgThreads, 0, ctx->stream>>>(
            P<PathCells>(ctx->cells), P<PathRaster>(ctx->pathRaster), P<PathPaint>(ctx->pathPaint),
            P<uint2>(ctx->cellPacked),
            P<uint32_t>(ctx->binned), P<Prim>(ctx->prims),
            P<uint32_t>(ctx->ramps), gradRows, P<ImageDesc>(ctx->imageTable), P<CanvasDev>(ctx->canvasTable),
            canvasCount, d->dither, ctx->featherTables, order, tileTotal, queue + 2 * kCostClasses);
    }
    HIP_TRY(hipGetLastError());
    RIVE_TRY(mark(ctx, ctx->ev[5]));
    return RIVE_HIP_OK;
}

int derive_patch_draws(rive_hip_context* ctx, const rive_hip_flush_desc* d, const SpanV* spans, uint32_t spanCount,
                       const ContourV* contours, uint32_t contourCount, uint32_t texels, uint32_t* drawCount,
                       uint32_t* totalPrims)
{
    if (d->patch_batches == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "patch_batch_count without patch_batches");
    const uint32_t nb = d->patch_batch_count;
    std::vector<PatchBatchDev> table(nb);
    for (uint32_t b = 0; b < nb; ++b)
    {
        const rive_hip_patch_batch& pb = d->patch_batches[b];
        if ((uint64_t)pb.first_patch + pb.patch_count > (uint64_t)texels / 8u)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "patch batch outside the tessellation texture");
        table[b] = {pb.first_patch, pb.first_patch + pb.patch_count, pb.image_index | (pb.image_sampler << 16), 0u};
    }
    const uint32_t np = d->path_count;
    *drawCount = 0;
    *totalPrims = 0;
    if (np == 0)
        return RIVE_HIP_OK;
    uint32_t* scal = P<uint32_t>(ctx->scalars);
    RIVE_TRY(ensure(ctx, ctx->patchTable, sizeof(PatchBatchDev) * std::max<uint32_t>(nb, 1)));
    if (nb > 0)
        HIP_TRY(hipMemcpyAsync(ctx->patchTable.ptr, table.data(), sizeof(PatchBatchDev) * nb,
                               hipMemcpyHostToDevice, ctx->stream));
    // 8 arrays of np + 2 words
    const size_t m = (size_t)np + 2;
    RIVE_TRY(ensure(ctx, ctx->patchWork, 8 * m * 4));
    uint32_t* w = P<uint32_t>(ctx->patchWork);
    uint32_t *fwdMin = w, *fwdMax = w + m, *mirMin = w + 2 * m, *mirMax = w + 3 * m;
    uint32_t *flag = w + 4 * m, *drawIdx = w + 5 * m, *cnt = w + 6 * m, *base = w + 7 * m;
    RIVE_TRY(clear(ctx, fwdMin, 0xff, m * 4));
    RIVE_TRY(clear(ctx, mirMin, 0xff, m * 4));
    RIVE_TRY(clear(ctx, fwdMax, 0, m * 4));
    RIVE_TRY(clear(ctx, mirMax, 0, m * 4));
    RIVE_TRY(clear(ctx, cnt, 0, m * 4));
    if (spanCount > 0)
        span_ranges_kernel<<<(spanCount + 255) / 256, 256, 0, ctx->stream>>>(
            spans, spanCount, contours, contourCount, np, texels, fwdMin, fwdMax, mirMin, mirMax, scal + 6);
    span_draw_flags_kernel<<<(np + 255) / 256, 256, 0, ctx->stream>>>(np, fwdMax, mirMax, flag);
    RIVE_TRY(device_scan(ctx, flag, np, drawIdx, scal + 7));
    RIVE_TRY(ensure(ctx, ctx->drawTess, (size_t)(np + 1) * sizeof(DrawTess)));
    span_draws_kernel<<<(np + 255) / 256, 256, 0, ctx->stream>>>(
        np, flag, drawIdx, fwdMin, fwdMax, mirMin, mirMax, ctx->in.paths, P<PatchBatchDev>(ctx->patchTable), nb,
        d->image_count, P<DrawTess>(ctx->drawTess), cnt, P<uint32_t>(ctx->pathImage), scal + 6);
    RIVE_TRY(device_scan(ctx, cnt, np, base, scal + 3));
    HIP_TRY(hipGetLastError());
    RIVE_TRY(fetch(ctx, ctx->hostScalars, scal + 3, 20));
    RIVE_TRY(settle(ctx));
    if (ctx->hostScalars[3] != 0)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED,
                    "spans naming no contour or path, or contour directions the raster-ordering pipeline "
                    "does not use");
    *totalPrims = ctx->hostScalars[0];
    *drawCount = ctx->hostScalars[4];
    if (*drawCount > 0)
        draw_prim_base_kernel<<<(*drawCount + 255) / 256, 256, 0, ctx->stream>>>(*drawCount, base,
                                                                                 P<DrawTess>(ctx->drawTess));
    HIP_TRY(hipGetLastError());
    return RIVE_HIP_OK;
}

// Single-block prefix over draws: forward ranges and primitive bases.
__global__ void draw_tess_kernel(const rive_hip_draw* __restrict__ draws, uint32_t n,
                                 uint32_t pathCount, uint32_t texels, uint32_t imageCount,
                                 DrawTess* __restrict__ out, uint32_t* __restrict__ pathImage,
                                 uint32_t* __restrict__ totalPrims, uint32_t* __restrict__ unsupported)
{
    __shared__ uint32_t buf[1024];
    __shared__ uint64_t carry; // 64-bit: a total beyond 2^32-1 saturates below
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += blockDim.x)
    {
        uint32_t i = base + threadIdx.x;
        DrawTess t{};
        uint32_t cnt = 0;
        if (i < n)
        {
            const rive_hip_draw d = draws[i];
            t.pathID = d.path_id;
            t.tessLocation = d.tess_location;
            t.tessVertexCount = d.tess_vertex_count;
            t.dirs = d.contour_directions;
            t.isStroke = d.is_stroke;
            switch (d.contour_directions)
            {
                case 0: // forward
                    t.forwardBase = d.tess_location;
                    t.forwardCount = d.tess_vertex_count;
                    break;
                case 2: // forwardThenReverse
                    t.forwardBase = d.tess_location;
                    t.forwardCount = d.tess_vertex_count / 2;
                    break;
                case 3: // reverseThenForward
                    t.forwardBase = d.tess_location + d.tess_vertex_count / 2;
                    t.forwardCount = d.tess_vertex_count / 2;
                    break;
                default: // reverse only: a feathered fill through the atlas
                    t.forwardBase = d.tess_location;
                    t.forwardCount = d.is_stroke == 2u ? d.tess_vertex_count : 0u;
                    if (d.is_stroke != 2u && d.tess_vertex_count != 0)
                        atomicAdd(unsupported, 1u);
                    break;
            }
            // A plain fill drawn forward only would need positive-winding-only
            // coverage; rasterOrdering fills are always double sided (a
            // feathered fill through the atlas is drawn one way).
            if (d.is_stroke == 0u && d.contour_directions == 0 && d.tess_vertex_count != 0)
                atomicAdd(unsupported, 1u);
            // Reject records that would index outside the path or texel arrays.
            if (d.path_id == 0 || d.path_id > pathCount ||
                (uint64_t)t.forwardBase + t.forwardCount + 1 > texels)
            {
                atomicAdd(unsupported, 1u);
                t.forwardCount = 0;
                t.pathID = 0;
            }
            else if (d.image_index != 0u)
            {
                if (d.image_index > imageCount || d.image_sampler >= 18u)
                    atomicAdd(unsupported, 1u);
                else
                    pathImage[d.path_id] = d.image_index | (d.image_sampler << 16);
            }
            cnt = t.forwardCount;
        }
        buf[threadIdx.x] = cnt;
        __syncthreads();
        for (uint32_t off = 1; off < blockDim.x; off <<= 1)
        {
            uint32_t v = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < n)
        {
            t.primBase = (uint32_t)carry + buf[threadIdx.x] - cnt;
            out[i] = t;
        }
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1)
            carry += buf[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *totalPrims = (uint32_t)min(carry, (uint64_t)kScanSaturated);
}
} // namespace

#include "count_kernels.hip"

extern "C" int rive_hip_last_flush_canvases(rive_hip_context* ctx, rive_hip_canvas* out, uint32_t capacity,
                                            uint32_t* count)
{
    if (ctx == nullptr || count == nullptr || (out == nullptr && capacity != 0))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    *count = (uint32_t)ctx->lastCanvases.size();
    for (uint32_t c = 0; c < std::min(capacity, *count); ++c)
        out[c] = ctx->lastCanvases[c];
    return RIVE_HIP_OK;
}

namespace
{
int flush_impl(rive_hip_context* ctx, const rive_hip_flush_desc* d, const GeoLayout* pre = nullptr)
{
    if (ctx == nullptr || d == nullptr || (d->target == nullptr && d->canvas_count == 0))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    // The canvas table: validated before anything is enqueued.  The vector
    // keeps its capacity across flushes (a no-alloc flush must not grow it).
    std::vector<rive_hip_canvas>& canvases = ctx->flushCanvases;
    canvases.clear();
    if (ctx->arena.active && std::max<uint32_t>(d->canvas_count, 1) > canvases.capacity())
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "rive_hip_flush_noalloc renders one canvas per flush");
    if (d->canvas_count == 0)
    {
        if (d->target->ctx != ctx)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "target belongs to another context");
        canvases.push_back({d->target, 1u, d->path_count, d->clear_color, d->load_action});
    }
    else
    {
        if (d->canvases == nullptr)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "canvas_count without canvases");
        uint64_t nextFree = 1;
        for (uint32_t c = 0; c < d->canvas_count; ++c)
        {
            const rive_hip_canvas& cv = d->canvases[c];
            if (cv.target == nullptr || cv.target->ctx != ctx)
                return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null or foreign canvas target");
            if (cv.first_path < nextFree || (uint64_t)cv.first_path + cv.path_count > (uint64_t)d->path_count + 1)
                return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT,
                            "canvas path ranges must ascend, not overlap and lie in 1..path_count");
            nextFree = (uint64_t)cv.first_path + cv.path_count;
            canvases.push_back(cv);
        }
    }
    for (int k = 0; k < RIVE_HIP_BUFFER_KIND_COUNT; ++k)
        if (ctx->bufs[k].mapped)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "flush with a mapped buffer");
    auto need = [&](int kind, size_t bytes) { return ctx->bufs[kind].size >= bytes; };
    // A SPANS flush described by DrawBatch patch runs has its draws derived
    // on the device; its draw_count is not read.
    const bool patchMode = d->mode == RIVE_HIP_FLUSH_SPANS && d->patch_batch_count > 0;
    const uint32_t drawCountIn = patchMode ? 0u : d->draw_count;
    // Record counts are bounded before any size arithmetic: path IDs index
    // per-path device arrays sized (path_count + 1) records.
    if (d->path_count > kMaxFlushPaths)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "path_count above 2^24");
    if (d->mode == RIVE_HIP_FLUSH_SPANS && d->tess_vertex_count > kMaxFlushTexels)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "tess_vertex_count above 2^31");
    const size_t pathRecords = (size_t)d->path_count + 1;
    // Each record run [first, first + count) must lie inside its buffer; the
    // offsets are bounded first so the products cannot wrap.
    const uint64_t kMaxOffset = (uint64_t)1 << 40;
    if (d->first_path > kMaxOffset || d->first_paint > kMaxOffset || d->first_paint_aux > kMaxOffset ||
        d->first_contour > kMaxOffset || d->first_grad_span > kMaxOffset || d->first_tess_span > kMaxOffset ||
        d->first_draw > kMaxOffset)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "record offset above 2^40");
    if (!need(RIVE_HIP_BUFFER_PATH, (d->first_path + pathRecords) * sizeof(rive_hip_path_data)) ||
        !need(RIVE_HIP_BUFFER_PAINT, (d->first_paint + pathRecords) * sizeof(rive_hip_paint_data)) ||
        !need(RIVE_HIP_BUFFER_PAINT_AUX, (d->first_paint_aux + pathRecords) * sizeof(rive_hip_paint_aux)) ||
        !need(RIVE_HIP_BUFFER_GRAD_SPAN, (d->first_grad_span + d->grad_span_count) * sizeof(rive_hip_grad_span)) ||
        !need(RIVE_HIP_BUFFER_DRAW, (d->first_draw + drawCountIn) * sizeof(rive_hip_draw)))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "buffers smaller than the flush descriptor");
    {
        auto dev = [&](int kind, uint64_t first, size_t recordBytes) -> unsigned char* {
            auto* base = static_cast<unsigned char*>(ctx->bufs[kind].dev);
            return base == nullptr ? nullptr : base + first * recordBytes;
        };
        auto host = [&](int kind, uint64_t first, size_t recordBytes) -> const unsigned char* {
            auto* base = static_cast<const unsigned char*>(ctx->bufs[kind].host);
            return base == nullptr ? nullptr : base + first * recordBytes;
        };
        auto& in = ctx->in;
        in.paths = reinterpret_cast<const PathDataV*>(dev(RIVE_HIP_BUFFER_PATH, d->first_path, 64));
        in.paints = reinterpret_cast<const uint2*>(dev(RIVE_HIP_BUFFER_PAINT, d->first_paint, 8));
        in.aux = reinterpret_cast<const float*>(dev(RIVE_HIP_BUFFER_PAINT_AUX, d->first_paint_aux, 128));
        in.gradSpans = reinterpret_cast<const uint4*>(dev(RIVE_HIP_BUFFER_GRAD_SPAN, d->first_grad_span, 16));
        in.draws = reinterpret_cast<rive_hip_draw*>(dev(RIVE_HIP_BUFFER_DRAW, d->first_draw, sizeof(rive_hip_draw)));
        in.spans = reinterpret_cast<const SpanV*>(dev(RIVE_HIP_BUFFER_TESS_SPAN, d->first_tess_span, 64));
        in.contours = reinterpret_cast<const ContourV*>(dev(RIVE_HIP_BUFFER_CONTOUR, d->first_contour, 16));
        in.paintsHost = reinterpret_cast<const rive_hip_paint_data*>(host(RIVE_HIP_BUFFER_PAINT, d->first_paint, 8));
        in.pathsHost = reinterpret_cast<const rive_hip_path_data*>(host(RIVE_HIP_BUFFER_PATH, d->first_path, 64));
        in.gradSpansHost =
            reinterpret_cast<const rive_hip_grad_span*>(host(RIVE_HIP_BUFFER_GRAD_SPAN, d->first_grad_span, 16));
    }
    RIVE_TRY(ensure(ctx, ctx->scalars, kScalarBytes));
    RIVE_TRY(ensure(ctx, ctx->drawTess, (size_t)std::max<uint32_t>(drawCountIn, 1) * sizeof(DrawTess)));
    RIVE_TRY(clear(ctx, ctx->scalars.ptr, 0, kScalarBytes));
    if (pre == nullptr) // a frame's flushes have their start and counting events set already
        RIVE_TRY(mark(ctx, ctx->ev[6])); // flush start

    const SpanV* spans;
    uint32_t spanCount;
    const ContourV* contours;
    uint32_t tessRows;
    if (d->mode == RIVE_HIP_FLUSH_SPANS)
    {
        if (!need(RIVE_HIP_BUFFER_TESS_SPAN, (d->first_tess_span + d->tess_span_count) * sizeof(rive_hip_tess_span)) ||
            !need(RIVE_HIP_BUFFER_CONTOUR, (d->first_contour + d->contour_count) * sizeof(rive_hip_contour)))
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "span/contour buffers too small");
        // Draws with texels index ContourData records (every tessellated
        // vertex names its contour), so a SPANS flush needs them.
        if ((drawCountIn > 0 || patchMode) && d->tess_span_count > 0 && d->contour_count == 0)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "SPANS flush with spans and draws but no contours");
        spans = ctx->in.spans;
        spanCount = d->tess_span_count;
        contours = ctx->in.contours;
        tessRows = std::max<uint32_t>((d->tess_vertex_count + kTessTextureWidth - 1) / kTessTextureWidth, 1);
        ctx->lastContourCount = d->contour_count;
        RIVE_TRY(mark(ctx, ctx->ev[1])); // no counting pass
    }
    else if (d->mode == RIVE_HIP_FLUSH_GEOMETRY && pre != nullptr)
    {
        // laid out by the frame's counting pass
        spans = pre->spans;
        spanCount = pre->spanCount;
        contours = pre->contours;
        tessRows = pre->tessRows;
        ctx->lastContourCount = pre->contourCount;
        ctx->lastRecordCount = pre->recordCount;
        ctx->lastTessVertexCount = pre->texelsUsed;
    }
    else if (d->mode == RIVE_HIP_FLUSH_GEOMETRY)
    {
        int rc = geometry_counting(ctx, d, &spans, &spanCount, &contours, &tessRows);
        if (rc)
            return rc;
    }
    else
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "unknown flush mode");

    uint32_t* scal = P<uint32_t>(ctx->scalars);
    // The flush's image table (the DrawBatch image textures).
    if (d->image_count > 0)
    {
        if (d->images == nullptr)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "image_count without images");
        if (ctx->hostImagesCap < d->image_count)
        {
            if (ctx->arena.active)
                return fail(RIVE_HIP_ERROR_UNSUPPORTED,
                            "rive_hip_flush_noalloc: more images than the staged table holds");
            if (ctx->hostImages)
                HIP_TRY(hipHostFree(ctx->hostImages));
            ctx->hostImages = nullptr;
            HIP_TRY(hipHostMalloc((void**)&ctx->hostImages, sizeof(ImageDesc) * d->image_count,
                                  hipHostMallocDefault));
            ctx->hostImagesCap = d->image_count;
        }
        for (uint32_t k = 0; k < d->image_count; ++k)
        {
            if (d->images[k] == nullptr || d->images[k]->ctx != ctx)
                return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null or foreign image in the image table");
            ctx->hostImages[k] = d->images[k]->desc;
        }
        RIVE_TRY(ensure(ctx, ctx->imageTable, sizeof(ImageDesc) * d->image_count));
        RIVE_TRY(upload_table(ctx, ctx->imageTable.ptr, ctx->hostImages, sizeof(ImageDesc) * d->image_count));
    }
    RIVE_TRY(ensure(ctx, ctx->pathImage, (size_t)(d->path_count + 1) * 4));
    RIVE_TRY(clear(ctx, ctx->pathImage.ptr, 0, (size_t)(d->path_count + 1) * 4));
    uint32_t totalPrims = 0;
    if (!patchMode)
    {
        draw_tess_kernel<<<1, 1024, 0, ctx->stream>>>(
            ctx->in.draws, drawCountIn,
            d->path_count, tessRows * kTessTextureWidth, d->image_count, P<DrawTess>(ctx->drawTess),
            P<uint32_t>(ctx->pathImage), scal + 3, scal + 4);
        HIP_TRY(hipGetLastError());
        RIVE_TRY(fetch(ctx, ctx->hostScalars, scal + 3, 8));
        RIVE_TRY(settle(ctx));
        totalPrims = ctx->hostScalars[0];
        if (ctx->hostScalars[1] != 0)
            return fail(RIVE_HIP_ERROR_UNSUPPORTED,
                        "draw record out of range, or contour directions the raster-ordering "
                        "pipeline does not use");
    }
    if (totalPrims > kMaxFlushTexels)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED, "flush too large: draws cover more than 2^31 texels");
    ctx->lastCanvases = canvases;
    ctx->lastSpanCount = spanCount;
    ctx->lastSpans = spans;
    ctx->lastContours = contours;
    ctx->lastDrawCount = drawCountIn;
    ctx->lastMode = d->mode;
    {
        rive_hip_flush_stats& st = ctx->stats;
        st = rive_hip_flush_stats{};
        st.width = canvases[0].target->width;
        st.height = canvases[0].target->height;
        st.load_action = canvases[0].load_action;
        st.mode = d->mode;
        st.path_count = d->path_count;
        st.draw_count = drawCountIn;
        st.contour_count = ctx->lastContourCount;
        st.span_count = spanCount;
        st.tess_vertex_count = d->mode == RIVE_HIP_FLUSH_GEOMETRY ? ctx->lastTessVertexCount : d->tess_vertex_count;
        st.curve_record_count = d->mode == RIVE_HIP_FLUSH_GEOMETRY ? ctx->lastRecordCount : 0;
        st.grad_simple_rows = d->grad_simple_rows;
        st.grad_complex_rows = d->grad_complex_rows;
        st.grad_span_count = d->grad_span_count;
        st.first_path_tess_location = d->first_path_tess_location;
        st.clear_color = canvases[0].clear_color;
        st.dither = d->dither;
        const rive_hip_grad_span* gradHost = ctx->in.gradSpansHost;
        if (gradHost != nullptr)
            for (uint32_t i = 0; i < d->grad_span_count; ++i)
                st.grad_simple_ramps += (gradHost[i].y_with_flags & 0x20000000u) == 0u ? 1u : 0u;
        // Gradient paints read PaintAuxData; the host staging copy of the
        // PAINT buffer holds this flush's records.
        const rive_hip_paint_data* paintsHost = ctx->in.paintsHost;
        if (paintsHost != nullptr)
            for (uint32_t i = 1; i <= d->path_count; ++i)
            {
                const uint32_t type = paintsHost[i].params & 0xfu;
                st.aux_path_count += (type == 2u || type == 3u) ? 1u : 0u;
            }
    }
    int rc = raster_pipeline(ctx, d, spans, spanCount, contours, ctx->lastContourCount, tessRows,
                             P<DrawTess>(ctx->drawTess), drawCountIn, totalPrims);
    if (rc)
        return rc;
    ctx->timed = true;
    return RIVE_HIP_OK;
}
} // namespace

struct rive_hip_frame_graph
{
    rive_hip_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> held; // pinned table snapshots the graph's copies read
    // Everything the graph's kernels read or write, except the targets and
    // images: its own pipeline buffers and a device snapshot of each record
    // buffer, so eager flushes between launches neither overwrite its
    // inputs nor reallocate its buffers.
    std::vector<DevBuf> scratch;
    void* records[RIVE_HIP_BUFFER_KIND_COUNT] = {};
};

extern "C" void rive_hip_frame_graph_destroy(rive_hip_frame_graph* g)
{
    if (g == nullptr)
        return;
    if (g->ctx)
        (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    for (void* p : g->held)
        (void)hipHostFree(p);
    for (DevBuf& b : g->scratch)
        if (b.ptr && !b.carved)
            (void)hipFree(b.ptr);
    for (void* p : g->records)
        if (p)
            (void)hipFree(p);
    delete g;
}

extern "C" int rive_hip_frame_graph_create(rive_hip_context* ctx, const rive_hip_flush_desc* descs, uint32_t n,
                                           rive_hip_frame_graph** out)
{
    if (ctx == nullptr || descs == nullptr || n == 0 || out == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument or empty frame");
    for (uint32_t k = 0; k < n; ++k)
        if (descs[k].patch_batch_count != 0)
            return fail(RIVE_HIP_ERROR_UNSUPPORTED, "frame graphs take no patch-batch flushes");
    for (int k = 0; k < RIVE_HIP_BUFFER_KIND_COUNT; ++k)
        if (ctx->bufs[k].mapped)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "a record buffer is mapped");
    HIP_TRY(hipSetDevice(ctx->device));
    auto g = std::make_unique<rive_hip_frame_graph>();
    g->ctx = ctx;
    // The graph's own buffers: the context's pipeline buffers are set aside
    // while the frame is recorded and captured (the flushes allocate fresh
    // ones, which the graph keeps), and the record buffers read through
    // device snapshots.
    DevBuf* all[kDevBufCount];
    devbufs(ctx, all);
    DevBuf saved[kDevBufCount];
    for (int i = 0; i < kDevBufCount; ++i)
    {
        saved[i] = *all[i];
        *all[i] = DevBuf{};
    }
    void* savedDev[RIVE_HIP_BUFFER_KIND_COUNT];
    for (int k = 0; k < RIVE_HIP_BUFFER_KIND_COUNT; ++k)
        savedDev[k] = ctx->bufs[k].dev;
    auto restore = [&]() {
        for (int i = 0; i < kDevBufCount; ++i)
        {
            g->scratch.push_back(*all[i]);
            *all[i] = saved[i];
        }
        for (int k = 0; k < RIVE_HIP_BUFFER_KIND_COUNT; ++k)
            ctx->bufs[k].dev = savedDev[k];
        ctx->lastSpans = ctx->lastContours = nullptr;
        ctx->in = rive_hip_context::FlushInputs{};
    };
    hipError_t e = hipSuccess;
    for (int k = 0; k < RIVE_HIP_BUFFER_KIND_COUNT && e == hipSuccess; ++k)
        if (ctx->bufs[k].dev != nullptr && ctx->bufs[k].size > 0)
        {
            e = hipMalloc(&g->records[k], ctx->bufs[k].size);
            if (e == hipSuccess)
                e = hipMemcpyAsync(g->records[k], ctx->bufs[k].dev, ctx->bufs[k].size, hipMemcpyDeviceToDevice,
                                   ctx->stream);
            if (e == hipSuccess)
                ctx->bufs[k].dev = g->records[k];
        }
    if (e != hipSuccess)
    {
        restore();
        rive_hip_frame_graph_destroy(g.release());
        return fail(RIVE_HIP_ERROR_OUT_OF_MEMORY, std::string("frame graph records: ") + hipGetErrorString(e));
    }
    // 1. an eager frame that sizes every buffer and records the totals it reads
    ctx->pendingReads.clear();
    ctx->tape.clear();
    ctx->reads = rive_hip_context::Reads::record;
    int rc = rive_hip_flush_frame(ctx, descs, n);
    ctx->reads = rive_hip_context::Reads::live;
    if (rc == RIVE_HIP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = fail(RIVE_HIP_ERROR_DEVICE, "frame graph: recording flush failed");
    if (rc != RIVE_HIP_OK)
    {
        restore();
        rive_hip_frame_graph_destroy(g.release());
        return rc;
    }
    // 2. the same host logic again, reading the recorded totals, with every
    //    kernel and copy captured
    ctx->reads = rive_hip_context::Reads::replay;
    ctx->tapePos = 0;
    ctx->capturing = true;
    ctx->captureHeld.clear();
    e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess)
    {
        rc = rive_hip_flush_frame(ctx, descs, n);
        e = hipStreamEndCapture(ctx->stream, &g->graph);
    }
    ctx->reads = rive_hip_context::Reads::live;
    ctx->capturing = false;
    g->held.swap(ctx->captureHeld);
    const bool consumed = ctx->tapePos == ctx->tape.size();
    ctx->pendingReads.clear();
    restore();
    if (e != hipSuccess)
    {
        rive_hip_frame_graph_destroy(g.release());
        return fail(RIVE_HIP_ERROR_DEVICE, std::string("frame capture: ") + hipGetErrorString(e));
    }
    if (rc != RIVE_HIP_OK || !consumed)
    {
        rive_hip_frame_graph_destroy(g.release());
        return rc != RIVE_HIP_OK ? rc : fail(RIVE_HIP_ERROR_DEVICE, "frame capture diverged from its recording");
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess)
    {
        rive_hip_frame_graph_destroy(g.release());
        return fail(RIVE_HIP_ERROR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    *out = g.release();
    return RIVE_HIP_OK;
}

extern "C" int rive_hip_frame_graph_launch(rive_hip_frame_graph* g)
{
    if (g == nullptr || g->exec == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null frame graph");
    HIP_TRY(hipGraphLaunch(g->exec, g->ctx->stream));
    return RIVE_HIP_OK;
}

namespace
{
void record_frame_desc(rive_hip_context* ctx, const rive_hip_flush_desc* d);

// Points ctx->ev at event set k (created on first use).
int use_event_set(rive_hip_context* ctx, uint32_t k)
{
    while (ctx->evSets.size() <= k)
    {
        if (ctx->evSets.size() >= kMaxFrameEventSets)
            return fail(RIVE_HIP_ERROR_UNSUPPORTED, "more than 64 logical flushes in one frame call");
        ctx->evSets.emplace_back();
        ctx->evSets.back().fill(nullptr);
        for (hipEvent_t& e : ctx->evSets.back())
            HIP_TRY(hipEventCreate(&e));
    }
    ctx->ev = ctx->evSets[k].data();
    return RIVE_HIP_OK;
}
} // namespace

extern "C" int rive_hip_flush(rive_hip_context* ctx, const rive_hip_flush_desc* d)
{
    if (ctx != nullptr)
    {
        ctx->lastFlushNoAlloc = false;
        ctx->evSetsUsed = 1;
        ctx->ev = ctx->evSets[0].data();
    }
    const int rc = flush_impl(ctx, d);
    if (rc == RIVE_HIP_OK)
        record_frame_desc(ctx, d);
    return rc;
}

extern "C" int rive_hip_flush_frame(rive_hip_context* ctx, const rive_hip_flush_desc* descs, uint32_t n)
{
    if (ctx == nullptr || descs == nullptr || n == 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument or empty frame");
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->lastFlushNoAlloc = false;
    // One counting launch for the whole frame needs GEOMETRY flushes of one
    // target whose draw records follow each other; anything else runs flush
    // by flush.
    bool fused = n <= kMaxFrameEventSets;
    for (uint32_t k = 0; k < n && fused; ++k)
        fused = descs[k].mode == RIVE_HIP_FLUSH_GEOMETRY && descs[k].canvas_count == 0 &&
                descs[k].first_path_tess_location == descs[0].first_path_tess_location &&
                (k == 0 || descs[k].first_draw == descs[k - 1].first_draw + descs[k - 1].draw_count) &&
                descs[k].draw_count <= 0x7fffffffu;
    if (fused)
    {
        uint64_t drawEnd = descs[n - 1].first_draw + descs[n - 1].draw_count;
        fused = (uint64_t)ctx->bufs[RIVE_HIP_BUFFER_DRAW].size >= drawEnd * sizeof(rive_hip_draw) &&
                drawEnd - descs[0].first_draw <= 0x7fffffffu;
        for (int k = 0; k < RIVE_HIP_BUFFER_KIND_COUNT; ++k)
            fused = fused && !ctx->bufs[k].mapped;
    }
    if (!fused)
    {
        for (uint32_t k = 0; k < n; ++k)
            RIVE_TRY(rive_hip_flush(ctx, &descs[k]));
        return RIVE_HIP_OK;
    }
    RIVE_TRY(use_event_set(ctx, 0));
    RIVE_TRY(mark(ctx, ctx->ev[6])); // frame start
    std::vector<GeoLayout> layouts(n);
    RIVE_TRY(geometry_counting_frame(ctx, descs, n, layouts.data()));
    RIVE_TRY(mark(ctx, ctx->ev[1])); // the frame's counting done
    for (uint32_t k = 0; k < n; ++k)
    {
        RIVE_TRY(use_event_set(ctx, k));
        if (k > 0)
        {
            RIVE_TRY(mark(ctx, ctx->ev[6]));
            RIVE_TRY(mark(ctx, ctx->ev[1]));
        }
        ctx->evSetsUsed = k + 1;
        RIVE_TRY(flush_impl(ctx, &descs[k], &layouts[k]));
        record_frame_desc(ctx, &descs[k]);
    }
    return RIVE_HIP_OK;
}

namespace
{
bool same_canvases(const std::vector<rive_hip_canvas>& a, const std::vector<rive_hip_canvas>& b)
{
    if (a.size() != b.size())
        return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].target != b[i].target || a[i].first_path != b[i].first_path ||
            a[i].path_count != b[i].path_count || a[i].clear_color != b[i].clear_color ||
            a[i].load_action != b[i].load_action)
            return false;
    return true;
}

void record_frame_desc(rive_hip_context* ctx, const rive_hip_flush_desc* d)
{
    // Copy the tables first: d may point into the recorded frame itself (a
    // frame re-issued from rive_hip_last_frame).
    const rive_hip_flush_desc desc = *d;
    std::vector<rive_hip_image*> images(d->images, d->images + (d->images ? d->image_count : 0));
    std::vector<rive_hip_canvas> canvases(d->canvases, d->canvases + (d->canvases ? d->canvas_count : 0));
    if (d->first_path == 0 && d->first_draw == 0 && d->first_tess_span == 0 && d->first_contour == 0)
        ctx->frameLen = 0;
    const size_t k = ctx->frameLen++;
    if (ctx->frameImages.size() <= k)
    {
        ctx->frameImages.emplace_back();
        ctx->frameCanvases.emplace_back();
    }
    // A slot keeps its storage while its tables stay the same, so the table
    // pointers rive_hip_last_frame handed out stay valid while that frame is
    // re-issued flush by flush (an inner vector keeps its buffer when the
    // outer one grows).
    if (ctx->frameImages[k] != images)
        ctx->frameImages[k] = std::move(images);
    if (!same_canvases(ctx->frameCanvases[k], canvases))
        ctx->frameCanvases[k] = std::move(canvases);
    ctx->frameDescs.resize(k + 1);
    ctx->frameDescs[k] = desc;
    ctx->frameDescs[k].images = ctx->frameImages[k].empty() ? nullptr : ctx->frameImages[k].data();
    ctx->frameDescs[k].canvases = ctx->frameCanvases[k].empty() ? nullptr : ctx->frameCanvases[k].data();
}
} // namespace

extern "C" int rive_hip_last_frame(rive_hip_context* ctx, rive_hip_flush_desc* out, uint32_t capacity,
                                   uint32_t* count)
{
    if (ctx == nullptr || count == nullptr || (out == nullptr && capacity != 0))
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    *count = (uint32_t)ctx->frameDescs.size();
    for (uint32_t k = 0; k < std::min(capacity, *count); ++k)
        out[k] = ctx->frameDescs[k];
    return RIVE_HIP_OK;
}

extern "C" int rive_hip_flush_noalloc(rive_hip_context* ctx, const rive_hip_flush_desc* d)
{
    if (ctx == nullptr || d == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    if (ctx->scratchPtr == nullptr)
        return fail(RIVE_HIP_ERROR_DEVICE, "no-alloc scratch address unavailable");
    // Set the context's own scratch buffers aside; every buffer this flush
    // needs is carved from the device's scratch array, and the set-aside
    // buffers come back untouched afterwards (the enqueued kernels keep the
    // carved addresses they were given).
    DevBuf* all[kDevBufCount];
    devbufs(ctx, all);
    DevBuf saved[kDevBufCount];
    for (int i = 0; i < kDevBufCount; ++i)
    {
        saved[i] = *all[i];
        *all[i] = DevBuf{};
    }
    ctx->evSetsUsed = 1;
    ctx->ev = ctx->evSets[0].data();
    ctx->arena.active = true;
    ctx->arena.base = ctx->scratchPtr;
    ctx->arena.used = 0;
    ctx->arena.cap = kNoAllocScratchBytes;
    const int rc = flush_impl(ctx, d);
    ctx->arena.active = false;
    for (int i = 0; i < kDevBufCount; ++i)
        *all[i] = saved[i];
    ctx->lastFlushNoAlloc = true;
    return rc;
}

// The raster configurations rive_hip_autotune_frame tries, in the order of
// its times_ms output.
constexpr rive_hip_raster_config kRasterCandidates[kRasterCandidateCount] = {
    {0u, 0u, 6u, 0u}, {0u, 0u, 5u, 0u}, {0u, 1u, 6u, 0u}, {0u, 1u, 5u, 0u},
    {1u, 0u, 6u, 0u}, {1u, 0u, 5u, 0u}, {1u, 1u, 6u, 0u}, {1u, 1u, 5u, 0u}};

extern "C" int rive_hip_autotune_frame(rive_hip_context* ctx, const rive_hip_flush_desc* descs, uint32_t n,
                                       uint32_t trials, rive_hip_raster_config* out, float* timesMs)
{
    if (ctx == nullptr || descs == nullptr || n == 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "autotune: null context or empty frame");
    if (trials == 0 || trials > 64)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "autotune: trials must be 1..64");
    if (ctx->capturing)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "autotune inside a stream capture");
    HIP_TRY(hipSetDevice(ctx->device));
    // The frame and its tables copied first: descs may point into the frame
    // the context records, which the trials re-record.
    std::vector<rive_hip_flush_desc> frame(descs, descs + n);
    std::vector<std::vector<rive_hip_image*>> images(n);
    std::vector<std::vector<rive_hip_canvas>> canvases(n);
    for (uint32_t k = 0; k < n; ++k)
    {
        if (frame[k].images != nullptr)
            images[k].assign(frame[k].images, frame[k].images + frame[k].image_count);
        if (frame[k].canvases != nullptr)
            canvases[k].assign(frame[k].canvases, frame[k].canvases + frame[k].canvas_count);
        frame[k].images = images[k].empty() ? nullptr : images[k].data();
        frame[k].canvases = canvases[k].empty() ? nullptr : canvases[k].data();
    }
    descs = frame.data();
    // The frame's targets, saved so that the trials leave their pixels as the
    // frame's first rendering will find them.
    std::vector<rive_hip_target*> targets;
    for (uint32_t k = 0; k < n; ++k)
    {
        const rive_hip_flush_desc& d = descs[k];
        if (d.canvas_count > 0 && d.canvases == nullptr)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "autotune: canvas_count without canvases");
        if (d.canvas_count > 0)
            for (uint32_t c = 0; c < d.canvas_count; ++c)
                targets.push_back(d.canvases[c].target);
        else
            targets.push_back(d.target);
    }
    std::sort(targets.begin(), targets.end());
    targets.erase(std::unique(targets.begin(), targets.end()), targets.end());
    size_t bytes = 0;
    for (rive_hip_target* t : targets)
    {
        if (t == nullptr || t->ctx != ctx)
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "autotune: a flush names a target of another context");
        bytes += (size_t)t->width * t->height * 4;
    }
    unsigned char* saved = nullptr;
    HIP_TRY(hipMalloc((void**)&saved, std::max<size_t>(bytes, 4)));
    size_t at = 0;
    hipError_t e = hipSuccess;
    for (rive_hip_target* t : targets)
    {
        const size_t b = (size_t)t->width * t->height * 4;
        if (e == hipSuccess && b > 0)
            e = hipMemcpyAsync(saved + at, t->pixels, b, hipMemcpyDeviceToDevice, ctx->stream);
        at += b;
    }
    const rive_hip_raster_config before = ctx->rasterConfig;
    int rc = e == hipSuccess ? RIVE_HIP_OK : fail(RIVE_HIP_ERROR_DEVICE, hipGetErrorString(e));
    float best = INFINITY;
    int bestK = 0;
    std::vector<float> t(trials);
    for (int k = 0; k < kRasterCandidateCount && rc == RIVE_HIP_OK; ++k)
    {
        ctx->rasterConfig = kRasterCandidates[k];
        rc = rive_hip_flush_frame(ctx, descs, n); // untimed: first use, buffer growth
        // a trial: the frame flush by flush, each flush's raster pass timed
        for (uint32_t r = 0; r < trials && rc == RIVE_HIP_OK; ++r)
        {
            t[r] = 0.f;
            for (uint32_t f = 0; f < n && rc == RIVE_HIP_OK; ++f)
            {
                float ms[6];
                rc = rive_hip_flush(ctx, &descs[f]);
                if (rc == RIVE_HIP_OK)
                    rc = rive_hip_pass_times(ctx, ms, 6);
                if (rc == RIVE_HIP_OK)
                    t[r] += ms[4];
            }
        }
        if (rc != RIVE_HIP_OK)
            break;
        std::sort(t.begin(), t.end());
        const float median = t[trials / 2];
        if (timesMs != nullptr)
            timesMs[k] = median;
        if (median < best)
        {
            best = median;
            bestK = k;
        }
    }
    at = 0;
    for (rive_hip_target* tg : targets)
    {
        const size_t b = (size_t)tg->width * tg->height * 4;
        if (e == hipSuccess && b > 0)
            e = hipMemcpyAsync(tg->pixels, saved + at, b, hipMemcpyDeviceToDevice, ctx->stream);
        at += b;
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(saved);
    if (rc != RIVE_HIP_OK)
    {
        ctx->rasterConfig = before;
        return rc;
    }
    if (e != hipSuccess)
        return fail(RIVE_HIP_ERROR_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
    ctx->rasterConfig = kRasterCandidates[bestK];
    if (out != nullptr)
        *out = ctx->rasterConfig;
    return RIVE_HIP_OK;
}

extern "C" int rive_hip_pass_times(rive_hip_context* ctx, float* ms, int n)
{
    if (ctx == nullptr || ms == nullptr || n < 0 || n > 6)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "bad arguments");
    if (!ctx->timed)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "no flush recorded");
    HIP_TRY(hipSetDevice(ctx->device));
    // Summed over the logical flushes of the last call (one for rive_hip_flush).
    float sum[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t k = 0; k < std::max<uint32_t>(ctx->evSetsUsed, 1); ++k)
    {
        const hipEvent_t* ev = ctx->evSets[k].data();
        HIP_TRY(hipEventSynchronize(ev[5]));
        float t[6];
        // ev: 6 start, 1 counted, 0 ramps start, 2 ramps end, 3 tess end, 4 bin end, 5 raster end
        HIP_TRY(hipEventElapsedTime(&t[0], ev[0], ev[2]));
        HIP_TRY(hipEventElapsedTime(&t[1], ev[6], ev[1]));
        HIP_TRY(hipEventElapsedTime(&t[2], ev[2], ev[3]));
        HIP_TRY(hipEventElapsedTime(&t[3], ev[3], ev[4]));
        HIP_TRY(hipEventElapsedTime(&t[4], ev[4], ev[5]));
        HIP_TRY(hipEventElapsedTime(&t[5], ev[6], ev[5]));
        for (int i = 0; i < 6; ++i)
            sum[i] += t[i];
    }
    for (int i = 0; i < n; ++i)
        ms[i] = sum[i];
    return RIVE_HIP_OK;
}

extern "C" int rive_hip_debug_readback(rive_hip_context* ctx, int what, void* dst, size_t capacity,
                                       size_t* size)
{
    if (ctx == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "ctx is null");
    HIP_TRY(hipSetDevice(ctx->device));
    if (ctx->lastFlushNoAlloc)
        return fail(RIVE_HIP_ERROR_UNSUPPORTED,
                    "intermediates of a no-alloc flush are not kept (its scratch is shared per device)");
    const bool geo = ctx->lastMode == RIVE_HIP_FLUSH_GEOMETRY;
    const void* src = nullptr;
    size_t bytes = 0;
    switch (what)
    {
        case 0: // TessVertexSpans consumed by the last flush
            src = ctx->lastSpans;
            bytes = (size_t)ctx->lastSpanCount * sizeof(rive_hip_tess_span);
            break;
        case 1: // ContourData
            src = ctx->lastContours;
            bytes = (size_t)ctx->lastContourCount * sizeof(rive_hip_contour);
            break;
        case 2: // tessellation texture
            src = ctx->tess.ptr;
            bytes = (size_t)ctx->lastTexels * sizeof(TessTexel);
            break;
        case 3: // rive_hip_draw records (GEOMETRY mode: with device-written ranges)
            src = ctx->in.draws;
            bytes = (size_t)ctx->lastDrawCount * sizeof(rive_hip_draw);
            break;
        case 4: // gradient ramps
            src = ctx->ramps.ptr;
            bytes = (size_t)ctx->lastGradRows * kGradWidth * 4;
            break;
        case 5: // device cubic records (GEOMETRY mode)
            src = geo ? ctx->geoRecords.ptr : nullptr;
            bytes = geo ? (size_t)ctx->lastRecordCount * 64 : 0;
            break;
        default:
            return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "unknown readback");
    }
    if (size)
        *size = bytes;
    if (dst != nullptr && src != nullptr && bytes > 0)
    {
        HIP_TRY(hipMemcpyAsync(dst, src, std::min(bytes, capacity), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return RIVE_HIP_OK;
}

extern "C" int rive_hip_set_raster_config(rive_hip_context* ctx, const rive_hip_raster_config* config)
{
    if (ctx == nullptr || config == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null context or config");
    if (config->tile_order > 1 || config->persistent > 1 ||
        (config->waves_per_simd != 5 && config->waves_per_simd != 6) || config->reserved != 0)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT,
                    "raster config: tile_order and persistent are 0 or 1, waves_per_simd 5 or 6, reserved 0");
    ctx->rasterConfig = *config;
    return RIVE_HIP_OK;
}

extern "C" int rive_hip_set_interior_triangulation(rive_hip_context* ctx, uint32_t enabled)
{
    if (ctx == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null context");
    ctx->interiorTriangulation = enabled != 0;
    return RIVE_HIP_OK;
}

extern "C" uint32_t rive_hip_get_interior_triangulation(const rive_hip_context* ctx)
{
    return ctx != nullptr && ctx->interiorTriangulation ? 1u : 0u;
}

extern "C" int rive_hip_get_raster_config(rive_hip_context* ctx, rive_hip_raster_config* out)
{
    if (ctx == nullptr || out == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null context or output");
    *out = ctx->rasterConfig;
    return RIVE_HIP_OK;
}

extern "C" int rive_hip_last_flush_stats(rive_hip_context* ctx, rive_hip_flush_stats* out)
{
    if (ctx == nullptr || out == nullptr)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    if (!ctx->timed)
        return fail(RIVE_HIP_ERROR_INVALID_ARGUMENT, "no flush recorded");
    *out = ctx->stats;
    return RIVE_HIP_OK;
}
