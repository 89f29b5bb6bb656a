/*
 * rive_hip.h -- C ABI of the MI355X (gfx950) backend for Rive's GPU renderer.
 *
 * The backend plugs in where the reference's per-API backends do: behind
 * rive::gpu::RenderContextImpl (renderer/include/rive/renderer/
 * render_context_impl.hpp:26-243), which rive::gpu::RenderContext owns and
 * drives once per frame (render_context.cpp:770-1044).  Every entry point
 * below names the RenderContextImpl virtual it replaces.  Nothing in these
 * signatures is a C++ or torch type: plain pointers, sizes and PODs only.
 *
 * Two flush modes share one device pipeline:
 *   RIVE_HIP_FLUSH_SPANS     the host already wrote TessVertexSpan/ContourData
 *                            records, exactly as the reference RenderContext
 *                            does on the CPU (render_context.cpp:1424-2408).
 *   RIVE_HIP_FLUSH_GEOMETRY  the host wrote raw path geometry plus one
 *                            rive_hip_draw per PathDraw; the device counts
 *                            curves (Wang's formula, polar segments, joins,
 *                            caps), scans offsets and writes the same
 *                            TessVertexSpan/ContourData records itself.
 * Both then tessellate, bin, rasterise and resolve to RGBA8 on the device.
 *
 * Status: every function returns RIVE_HIP_OK (0) or a positive error code;
 * rive_hip_last_error() gives a message.  No function throws.
 */
#ifndef RIVE_HIP_H
#define RIVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum rive_hip_status
{
    RIVE_HIP_OK = 0,
    RIVE_HIP_ERROR_INVALID_ARGUMENT = 1,
    RIVE_HIP_ERROR_DEVICE = 2,
    RIVE_HIP_ERROR_OUT_OF_MEMORY = 3,
    RIVE_HIP_ERROR_UNSUPPORTED = 4,
    RIVE_HIP_ERROR_NOT_MAPPED = 5,
};

/* ------------------------------------------------------------------------ */
/* Record formats (byte-compatible with the reference's GPU records).       */
/* ------------------------------------------------------------------------ */

/* gpu::TessVertexSpan, 64 bytes (gpu.hpp:291-375). */
typedef struct rive_hip_tess_span
{
    float pts[8];
    float join_tangent[2];
    float y;
    float reflection_y;
    int32_t x0x1;
    int32_t reflection_x0x1;
    uint32_t segment_counts; /* join<<20 | polar<<10 | parametric */
    uint32_t contour_id_with_flags;
} rive_hip_tess_span;

/* gpu::ContourData, 16 bytes (gpu.hpp:1698-1714). */
typedef struct rive_hip_contour
{
    float midpoint[2];
    uint32_t path_id;
    uint32_t vertex_index0;
} rive_hip_contour;

/* gpu::PathData, 64 bytes (gpu.hpp, PathData::set at gpu.cpp:859-877). */
typedef struct rive_hip_path_data
{
    float matrix[6];
    float stroke_radius; /* 0 => fill */
    float feather_radius;
    uint32_t z_index;
    float atlas_transform[3];
    uint32_t coverage_range[4];
} rive_hip_path_data;

/* gpu::PaintData, 8 bytes (PaintData::set at gpu.cpp:879-939). */
typedef struct rive_hip_paint_data
{
    uint32_t params; /* clipID<<16 | blend<<4 | flags | paintType */
    uint32_t value;  /* RGBA color, gradient row (float bits) or clip ID<<16 */
} rive_hip_paint_data;

/* gpu::PaintAuxData, 128 bytes (PaintAuxData::set at gpu.cpp:941-1063). */
typedef struct rive_hip_paint_aux
{
    float paint_matrix[6];
    float grad_span[2]; /* horizontal span, normalized */
    float clip_rect_inverse_matrix[6];
    float inverse_fwidth[2];
    float image_matrix[6];
    float image_lod;
    float pad[9];
} rive_hip_paint_aux;

/* gpu::GradientSpan, 16 bytes (gpu.hpp:248-270). */
typedef struct rive_hip_grad_span
{
    uint32_t horizontal_span; /* x1Fixed<<16 | x0Fixed */
    uint32_t y_with_flags;
    uint32_t color0; /* ColorInt ARGB */
    uint32_t color1;
} rive_hip_grad_span;

/* One PathDraw for device-side counting (RIVE_HIP_FLUSH_GEOMETRY).  The host
 * computes what PathDraw::PathDraw/initForMidpointFan compute once per draw
 * (draw.cpp:561-825): the matrix max scale, polar segments per radian and the
 * contour directions.  The device does the per-curve work. */
typedef struct rive_hip_draw
{
    uint32_t path_id;          /* 1-based PathData/PaintData index */
    uint32_t first_verb;       /* into the VERB buffer */
    uint32_t verb_count;
    uint32_t first_point;      /* into the POINT buffer (float pairs) */
    uint32_t point_count;
    uint32_t contour_directions; /* 0 fwd, 1 rev, 2 fwdThenRev, 3 revThenFwd */
    uint32_t contour_flags;    /* e.g. NEGATE_PATH_FILL_COVERAGE_FLAG */
    uint32_t join;             /* 0 miter, 1 round, 2 bevel */
    uint32_t cap;              /* 0 butt, 1 round, 2 square */
    uint32_t is_stroke;        /* 0 fill, 1 stroke, 2 feathered fill (counted like a
                                  stroke whose joins are feather joins), 3 fill laid
                                  out for interior triangulation (outer-cubic
                                  patches, draw.cpp:2401-2650) */
    float matrix[6];
    float stroke_radius;
    float matrix_max_scale;
    float polar_segments_per_radian;
    float parametric_precision;
    /* SPANS mode: the draw's tessellation allocation, as in the reference's
     * DrawBatch (gpu.hpp) -- first texel and texel count (both directions).
     * GEOMETRY mode: written by the device. */
    uint32_t tess_location;
    uint32_t tess_vertex_count;
    /* Image paints (the DrawBatch's imageTexture and ImageSampler): 0 = none,
     * k = rive_hip_flush_desc.images[k - 1].  The sampler is
     * ImageSampler::asKey() (image_sampler.hpp:56-61): wrapX + 3*wrapY +
     * 9*filter. */
    uint32_t image_index;
    uint32_t image_sampler;
} rive_hip_draw;

/* ------------------------------------------------------------------------ */
/* Backend context.                                                         */
/* ------------------------------------------------------------------------ */

typedef struct rive_hip_context rive_hip_context;
typedef struct rive_hip_target rive_hip_target;
typedef struct rive_hip_image rive_hip_image;

/* Message for the last error on this thread (never NULL). */
const char* rive_hip_last_error(void);

/* Creates a backend on HIP device `device` with its own stream.
 * Replaces: the backend's MakeContext factory, e.g.
 * RenderContextVulkanImpl::MakeContext (render_context_vulkan_impl.cpp:4259). */
int rive_hip_create(int device, rive_hip_context** out);
void rive_hip_destroy(rive_hip_context* ctx);

/* Buffer kinds.  The first nine mirror RenderContextImpl::resize*Buffer
 * (render_context_impl.hpp:108-120); VERB/POINT/DRAW carry the
 * RIVE_HIP_FLUSH_GEOMETRY input. */
enum rive_hip_buffer_kind
{
    RIVE_HIP_BUFFER_FLUSH_UNIFORM = 0,
    RIVE_HIP_BUFFER_PATH = 1,
    RIVE_HIP_BUFFER_PAINT = 2,
    RIVE_HIP_BUFFER_PAINT_AUX = 3,
    RIVE_HIP_BUFFER_CONTOUR = 4,
    RIVE_HIP_BUFFER_GRAD_SPAN = 5,
    RIVE_HIP_BUFFER_TESS_SPAN = 6,
    RIVE_HIP_BUFFER_TRIANGLE_VERTEX = 7,
    RIVE_HIP_BUFFER_IMAGE_DRAW = 8,
    RIVE_HIP_BUFFER_VERB = 9,
    RIVE_HIP_BUFFER_POINT = 10,
    RIVE_HIP_BUFFER_DRAW = 11,
    RIVE_HIP_BUFFER_KIND_COUNT = 12
};

/* Replaces RenderContextImpl::resize{FlushUniform,Path,Paint,PaintAux,Contour,
 * GradSpan,TessVertexSpan,TriangleVertex,ImageDrawInstance}Buffer
 * (render_context_impl.hpp:108-120).  Allocates exactly `bytes` of device
 * memory plus pinned host staging. */
int rive_hip_buffer_resize(rive_hip_context* ctx, int kind, size_t bytes);

/* Replaces RenderContextImpl::map*Buffer (render_context_impl.hpp:155-165).
 * Returns a write-only pinned host pointer, or NULL on failure (the caller
 * then drops the frame, render_context.cpp:1012-1016). */
void* rive_hip_buffer_map(rive_hip_context* ctx, int kind, size_t bytes);

/* Replaces RenderContextImpl::unmap*Buffer (render_context_impl.hpp:166-174).
 * Enqueues the host-to-device copy of `bytes` on the context stream. */
int rive_hip_buffer_unmap(rive_hip_context* ctx, int kind, size_t bytes);

/* Replaces RenderContextImpl::makeRenderTarget (render_target.hpp:15-39 /
 * the backend's RenderTarget factory): an RGBA8 image in HBM. */
int rive_hip_target_create(rive_hip_context* ctx, uint32_t width, uint32_t height,
                           rive_hip_target** out);
void rive_hip_target_destroy(rive_hip_target* target);

/* Replaces RenderContextImpl::makeImageTexture (render_context_impl.hpp:61-70):
 * a premultiplied RGBA8 image in HBM.  mip_levels = 0 builds the full chain;
 * levels after the first are built on the device with a 2x2 box filter (the
 * reference generates them on the GPU too, render_context_impl.hpp:43-70). */
int rive_hip_image_create(rive_hip_context* ctx, uint32_t width, uint32_t height,
                          uint32_t mip_levels, const uint8_t* rgba_premul,
                          rive_hip_image** out);
void rive_hip_image_destroy(rive_hip_image* image);

/* ---- Encoded images (RenderContext::decodeImage, render_context.cpp:237-262,
 * and the decoders behind Bitmap::decode, decoders/src).  PNG and
 * Huffman-coded 8-bit JPEG (baseline, extended and progressive) are decoded;
 * WebP is recognised and refused with RIVE_HIP_ERROR_UNSUPPORTED, as the
 * reference built with no_rive_webp refuses it; lossless, hierarchical and
 * arithmetic-coded JPEG are refused too. */
enum rive_image_format
{
    RIVE_IMAGE_UNKNOWN = 0,
    RIVE_IMAGE_PNG = 1,
    RIVE_IMAGE_JPEG = 2,
    RIVE_IMAGE_WEBP = 3
};

/* The format (by signature, bitmap_decoder_thirdparty.cpp:36-70) and size
 * from the header alone.  *format is set even when the call fails. */
int rive_image_probe(const uint8_t* bytes, size_t n, uint32_t* format, uint32_t* width, uint32_t* height);

/* Decodes a PNG on the host into unpremultiplied RGBA8 (capacity >= width *
 * height * 4 from rive_image_probe), after libpng's transformations in
 * decode_png.cpp:87-116: palette and low-bit-depth gray expanded, tRNS to
 * alpha, 16-bit samples stripped to their high byte, gray to RGB, alpha 255
 * when absent.  Damaged data (CRC of a critical chunk, zlib stream, filter
 * type) and data too short for the image fail, as libpng's errors make the
 * reference's DecodePng return null. */
int rive_png_decode(const uint8_t* bytes, size_t n, uint8_t* rgba, size_t capacity);

/* A JPEG after entropy decoding (host side; every scan of a progressive
 * file applied): the frame's components and quantization tables, and with
 * coef != NULL (capacity >= coef_count) the quantized coefficients of every
 * 8x8 block in natural order, component by component over each component's
 * MCU-padded block grid. */
typedef struct rive_jpeg_component_info
{
    uint32_t id, h, v, quant_table;
    uint32_t scale_h, scale_v; /* inverse-DCT output per block side / 8 */
    uint32_t rep_h, rep_v;     /* further replication to the image grid */
    uint32_t blocks_w, blocks_h;
    uint64_t coef_offset;
} rive_jpeg_component_info;

typedef struct rive_jpeg_info
{
    uint32_t width, height, component_count;
    uint32_t ycc; /* 1: YCbCr, 0: RGB (3 components) or gray (1) */
    uint32_t hmax, vmax;
    uint64_t coef_count;
    rive_jpeg_component_info components[4];
    uint16_t quant[4][64]; /* natural order */
} rive_jpeg_info;

int rive_jpeg_decode_info(const uint8_t* bytes, size_t n, rive_jpeg_info* info, int16_t* coef, size_t capacity);

/* Decodes an encoded image into a premultiplied RGBA8 image in HBM with its
 * full mip chain: what RenderContext::decodeImage does with Bitmap::decode,
 * pixelFormat(RGBAPremul) and makeImageTexture(..., generateRemainingMips).
 * The container and entropy decoding run on the host; JPEG's inverse DCT,
 * chroma upsampling and colour conversion, the premultiplication and the
 * mips run on the device.  Images beyond 16384 texels a side fail with
 * RIVE_HIP_ERROR_UNSUPPORTED before anything is decoded. */
int rive_hip_image_decode(rive_hip_context* ctx, const uint8_t* bytes, size_t n, rive_hip_image** out);

/* Reads one mip level back (premultiplied RGBA8); with rgba NULL only the
 * level's width and height are returned. */
int rive_hip_image_read(rive_hip_image* image, uint32_t level, uint8_t* rgba, size_t capacity, uint32_t* width,
                        uint32_t* height);

enum rive_hip_flush_mode
{
    RIVE_HIP_FLUSH_SPANS = 0,
    RIVE_HIP_FLUSH_GEOMETRY = 1
};

/* POD mirror of the gpu::FlushDescriptor fields the HIP pipeline reads
 * (gpu.hpp:1320-1438). */
typedef struct rive_hip_flush_desc
{
    rive_hip_target* target;
    uint32_t mode;             /* rive_hip_flush_mode */
    uint32_t load_action;      /* 0 clear, 1 preserve */
    uint32_t clear_color;      /* ColorInt ARGB, unpremultiplied */
    uint32_t dither;           /* 0 none, 1 interleaved gradient noise */
    uint32_t path_count;       /* PathData records 1..path_count are live */
    uint32_t contour_count;    /* SPANS mode: ContourData records */
    uint32_t tess_span_count;  /* SPANS mode: TessVertexSpan records */
    uint32_t tess_vertex_count;/* SPANS mode: texels used (incl. padding) */
    uint32_t draw_count;       /* GEOMETRY mode: rive_hip_draw records */
    uint32_t grad_span_count;
    uint32_t grad_simple_rows;   /* rows of simple (2-texel) ramps */
    uint32_t grad_complex_rows;  /* rows of complex ramps, after the simple rows */
    uint32_t first_path_tess_location; /* GEOMETRY mode: usually 8 */
    rive_hip_image* const* images;     /* referenced by rive_hip_draw.image_index */
    uint32_t image_count;
    /* Batched flush: canvas_count > 0 renders each canvas's paths into its own
     * target (target, load_action and clear_color above are then unused).
     * canvas_count == 0 is the reference's single-target flush. */
    const struct rive_hip_canvas* canvases;
    uint32_t canvas_count;
    /* Where this logical flush's records start in their buffers, as
     * FlushDescriptor's firstPath/firstPaint/firstPaintAux/firstContour/
     * firstGradSpan/firstTessVertexSpan (gpu.hpp:1386-1395): one frame's
     * logical flushes share the mapped buffers, each flush reading its own
     * run.  Path ID k reads PATH record first_path + k (k = 0 is the
     * flush's reserved record), contour ID k reads CONTOUR record
     * first_contour + k - 1.  first_draw offsets the DRAW records the same
     * way.  All zero for a frame of one logical flush. */
    uint64_t first_path, first_paint, first_paint_aux, first_contour;
    uint64_t first_grad_span, first_tess_span, first_draw;
    /* SPANS mode, as the reference hands a flush over: instead of per-draw
     * rive_hip_draw records, its DrawBatch list (FlushDescriptor::drawList,
     * gpu.hpp:1219-1281) -- runs of midpoint-fan patches.  When
     * patch_batch_count > 0 the device derives each path's tessellation range
     * from its TessVertexSpans (each names its contour, whose ContourData
     * names the path), draws the paths in path-ID order, takes each path's
     * image paint from the batch holding its first patch, and ignores
     * draw_count. */
    const struct rive_hip_patch_batch* patch_batches;
    uint32_t patch_batch_count;
} rive_hip_flush_desc;

/* One DrawBatch of DrawType::midpointFanPatches: patches first_patch ..
 * first_patch + patch_count - 1 of the tessellation texture (8 texels each,
 * DrawBatch::baseElement / elementCount) and the batch's image paint. */
typedef struct rive_hip_patch_batch
{
    uint32_t first_patch;
    uint32_t patch_count;
    uint32_t image_index;   /* 0 none, k = rive_hip_flush_desc.images[k - 1] */
    uint32_t image_sampler; /* ImageSampler::asKey() */
} rive_hip_patch_batch;

/* One canvas of a batched flush: paths first_path .. first_path+path_count-1
 * render into `target`.  No counterpart in the reference, whose FlushDescriptor
 * has one render target (gpu.hpp:1320-1438); it lets one flush render many
 * small canvases (one launch per pass for all of them).  Canvases must be in
 * ascending first_path order and must not overlap. */
typedef struct rive_hip_canvas
{
    rive_hip_target* target;
    uint32_t first_path;   /* >= 1 */
    uint32_t path_count;
    uint32_t clear_color;  /* ColorInt ARGB */
    uint32_t load_action;  /* 0 clear, 1 preserve */
} rive_hip_canvas;

/* Replaces RenderContextImpl::flush (render_context_impl.hpp:217): ramps,
 * (counting), tessellation, binning, rasterisation and RGBA8 resolve, all
 * enqueued on the context stream. */
int rive_hip_flush(rive_hip_context* ctx, const rive_hip_flush_desc* desc);

/* rive_hip_flush that allocates no memory: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, records and
 * results as rive_hip_flush; the caller still provides the inputs (the
 * context's mapped buffers) and the target.  Every intermediate buffer is
 * carved from one 256 MiB scratch array per GPU, held until the process
 * ends, instead of the context's grow-only buffers; a flush that needs more
 * fails with RIVE_HIP_ERROR_OUT_OF_MEMORY, one that needs more than one
 * canvas or than 16 images with RIVE_HIP_ERROR_UNSUPPORTED.  All calls on a
 * GPU share its array, so no-alloc flushes that overlap on one GPU (several
 * contexts, several threads) race.  rive_hip_debug_readback of intermediates
 * is unavailable after a no-alloc flush. */
int rive_hip_flush_noalloc(rive_hip_context* ctx, const rive_hip_flush_desc* desc);

/* Per-draw resource counts of the draws in the DRAW/VERB/POINT buffers,
 * computed on the device: tess_vertex_counts[i] = draw i's tessellation
 * vertices (both directions, with contour padding) and contour_counts[i] its
 * contours, i.e. PathDraw::resourceCounts() (draw.cpp:1378-1391).  The front
 * end uses them to split a frame into logical flushes the way
 * LogicalFlush::pushDraws does (render_context.cpp:517-541) before laying
 * each flush out with rive_hip_flush.  Synchronous. */
int rive_hip_count_draws(rive_hip_context* ctx, uint32_t draw_count, uint32_t* tess_vertex_counts,
                         uint32_t* contour_counts);

/* The descriptors of the last frame's logical flushes, in order: the flushes
 * since the last one whose records start at offset 0 (first_* all zero).
 * Their image and canvas tables are copies the context keeps: they stay
 * valid while the same frame is re-issued, and until a flush records other
 * tables in their place (the images and targets they name must still be
 * alive).  Re-issuing them with rive_hip_flush re-renders the frame from the
 * records still resident in HBM (what a serving loop or a benchmark of the
 * flush alone does).  Writes min(capacity, count) descriptors; *count gets
 * the full count. */
int rive_hip_last_frame(rive_hip_context* ctx, rive_hip_flush_desc* out, uint32_t capacity, uint32_t* count);

/* Runs a frame's logical flushes in order, with the results of n
 * rive_hip_flush calls.  When they are GEOMETRY flushes of one target whose
 * DRAW records follow each other (descs[k].first_draw = descs[k-1].first_draw
 * + descs[k-1].draw_count, at most 64 flushes), their curve counting, layout
 * scans, record emission and span writing run once for the whole frame --
 * each flush still laid out on its own from location 8 and contour ID 1 --
 * and only the per-flush passes (tessellation, binning, raster) run per
 * flush.  Replaces the per-logical-flush RenderContextImpl::flush calls of
 * one RenderContext::flush (render_context.cpp:971-1018).
 * rive_hip_pass_times then sums over the frame's flushes. */
int rive_hip_flush_frame(rive_hip_context* ctx, const rive_hip_flush_desc* descs, uint32_t n);

/* Waits for the context stream (the reference's postFlush/fence wait). */
int rive_hip_sync(rive_hip_context* ctx);

/* Copies the target's RGBA8 pixels (row-major, premultiplied) to host. */
int rive_hip_read_pixels(rive_hip_context* ctx, rive_hip_target* target, uint8_t* rgba);

/* Debug/parity readback of device intermediates from the last flush:
 *  0 TessVertexSpans, 1 ContourData, 2 tessellation texture (16 B/texel),
 *  3 rive_hip_draw records (GEOMETRY mode: with the device-written
 *  tess_location / tess_vertex_count), 4 gradient ramp texture (RGBA8,
 *  512 wide), 5 GEOMETRY-mode curve records (64 B each, before row
 *  wrapping).  Writes min(capacity, size) bytes; *size gets the full size. */
int rive_hip_debug_readback(rive_hip_context* ctx, int what, void* dst,
                            size_t capacity, size_t* size);

/* Per-pass device timings of the last flush in milliseconds (hipEvents on the
 * context stream), summed over the logical flushes of the last
 * rive_hip_flush_frame call: [0] ramps, [1] counting+spans, [2] tessellation,
 * [3] binning, [4] raster+resolve, [5] total.  n <= 6. */
int rive_hip_pass_times(rive_hip_context* ctx, float* ms, int n);

/* Sizes of the last flush, as the device laid it out (the counters the
 * reference keeps in ResourceCounters / FlushDescriptor,
 * render_context.hpp:354-397, gpu.hpp:1320-1438). */
typedef struct rive_hip_flush_stats
{
    uint32_t width, height, load_action, mode;
    uint32_t path_count, draw_count, contour_count, span_count;
    uint32_t tess_vertex_count;  /* texels laid out, incl. 8 + 1 padding */
    uint32_t curve_record_count; /* GEOMETRY mode: pushCubic records */
    uint32_t prim_count;         /* expanded fan/border/stroke primitives */
    uint32_t bin_entries;        /* (tile, primitive) pairs */
    uint32_t tile_count;
    uint32_t grad_simple_rows, grad_complex_rows;
    uint32_t aux_path_count;     /* paths whose paint reads PaintAuxData */
    uint32_t grad_span_count;
    uint32_t first_path_tess_location;
    uint32_t grad_simple_ramps;  /* two-texel ramps (one GradientSpan each) */
    uint32_t clear_color, dither;
    uint32_t guard_pieces;       /* pieces of primitives clipped to the guard box */
} rive_hip_flush_stats;

int rive_hip_last_flush_stats(rive_hip_context* ctx, rive_hip_flush_stats* out);

/* How the raster pass runs on this context.  No counterpart in the reference
 * (its raster is the GPU's fixed-function pipeline); these are the launch
 * choices of the tile raster, set by hand or by rive_hip_autotune_frame.
 * Every choice renders the same pixels.
 *  tile_order     0: tiles in canvas/row/column order; 1: heaviest first (a
 *                 per-tile cost estimate from the binning, two small kernels).
 *  persistent     0: one workgroup per tile; 1: a resident grid (waves_per_simd
 *                 workgroups per CU) taking tiles from a device queue.
 *  waves_per_simd 5 or 6: the occupancy the raster is compiled for (its
 *                 per-workgroup LDS allows at most 6). */
typedef struct rive_hip_raster_config
{
    uint32_t tile_order;
    uint32_t persistent;
    uint32_t waves_per_simd;
    uint32_t reserved; /* 0 */
} rive_hip_raster_config;

int rive_hip_set_raster_config(rive_hip_context* ctx, const rive_hip_raster_config* config);
int rive_hip_get_raster_config(rive_hip_context* ctx, rive_hip_raster_config* out);

/* Interior triangulation for frames the context's front end lays out
 * (RenderContext::FrameDescriptor::triangulationThresholds,
 * render_context.hpp:114, and PathDraw::Make, draw.cpp:510-538): off (0,
 * the reference's frameBudgetMs = 0) by default; on, a plain fill of at
 * least 512 x 512 px on screen and at most 256 verbs becomes outer-cubic
 * patches (rive_hip_draw.is_stroke = 3), decided by those guards alone.
 * The getter returns the setting (0 for a null context). */
int rive_hip_set_interior_triangulation(rive_hip_context* ctx, uint32_t enabled);
uint32_t rive_hip_get_interior_triangulation(const rive_hip_context* ctx);

/* The raster autotuner (no counterpart in the reference).  Runs a frame's
 * logical flushes (descs, as for rive_hip_flush_frame) once untimed and then
 * `trials` (1..64) times under each of RIVE_HIP_RASTER_CANDIDATES raster
 * configs -- tile_order x persistent x waves_per_simd, in the order
 * {0,0,6} {0,0,5} {0,1,6} {0,1,5} {1,0,6} {1,0,5} {1,1,6} {1,1,5} -- times
 * each flush's raster pass with the context's HIP events (a trial runs the
 * flushes one by one), and keeps the config of the
 * smallest median on the context.  The frame's targets are saved before and
 * restored after, so their pixels are as they were.  out (may be null) gets
 * the chosen config, times_ms (may be null, RIVE_HIP_RASTER_CANDIDATES
 * floats) each candidate's median raster milliseconds.  Allocates a copy of
 * the targets; synchronizes the stream. */
#define RIVE_HIP_RASTER_CANDIDATES 8
int rive_hip_autotune_frame(rive_hip_context* ctx, const rive_hip_flush_desc* descs, uint32_t n,
                            uint32_t trials, rive_hip_raster_config* out, float* times_ms);

/* The canvas table of the last flush (one entry for a single-target flush).
 * Writes min(capacity, count) entries; *count gets the full count. */
int rive_hip_last_flush_canvases(rive_hip_context* ctx, rive_hip_canvas* out, uint32_t capacity,
                                 uint32_t* count);

/* ------------------------------------------------------------------------ */
/* Draw streams: a recorded sequence of rive::Renderer calls.               */
/* ------------------------------------------------------------------------ */

/* For draw i: save(); transform(matrix); clipPath() for each clip path;
 * drawPath(path, paint); restore() -- rive::Renderer (renderer.hpp:209-236). */
typedef struct rive_stream_path
{
    uint32_t first_verb, verb_count;   /* rive::PathVerb values (move=0, line=1,
                                          quad=2, cubic=4, close=5) */
    uint32_t first_point, point_count; /* float pairs */
    uint32_t fill_rule;                /* 0 nonZero, 1 evenOdd, 2 clockwise */
} rive_stream_path;

typedef struct rive_stream_draw
{
    uint32_t path;          /* index into paths */
    float matrix[6];
    uint32_t style;         /* 0 fill, 1 stroke */
    float thickness;
    uint32_t join, cap;
    float feather;
    uint32_t blend_mode;    /* rive::BlendMode value (srcOver = 3) */
    uint32_t paint_type;    /* 0 solid, 1 linear gradient, 2 radial gradient */
    uint32_t color;         /* ColorInt ARGB for solid paints */
    uint32_t first_stop, stop_count; /* gradients: into stop arrays */
    float grad[4];          /* linear: sx, sy, ex, ey; radial: cx, cy, r, 0 */
    uint32_t clip_count;    /* 0..2 */
    uint32_t clip_paths[2];
    /* drawImage instead of drawPath when image != 0: images[image - 1] under
     * `matrix`, with `blend_mode`, ImageSampler key `image_sampler` and
     * `opacity` (renderer.hpp drawImage). */
    uint32_t image;
    uint32_t image_sampler;
    float opacity;
} rive_stream_draw;

typedef struct rive_stream_image
{
    uint32_t width, height;
    const uint8_t* rgba_premul; /* width * height * 4 bytes */
} rive_stream_image;

typedef struct rive_stream
{
    uint32_t width, height;
    uint32_t clear_color;   /* ColorInt ARGB */
    uint32_t dither;        /* 0 none, 1 interleaved gradient noise */
    const uint8_t* verbs;
    uint32_t verb_count;
    const float* points;    /* x,y pairs */
    uint32_t point_count;
    const rive_stream_path* paths;
    uint32_t path_count;
    const rive_stream_draw* draws;
    uint32_t draw_count;
    const uint32_t* stop_colors;
    const float* stop_offsets;
    uint32_t stop_count;
    const rive_stream_image* images;
    uint32_t image_count;
} rive_stream;

/* Seeded synthetic streams (std::mt19937_64 with the float/int mappings of
 * tests/common/rand.hpp:13-54).  config: 1 = 1024x1024, 200 closed cubic
 * fills; 2 = 1920x1080, 5000 paths of fills/strokes/gradients/clips;
 * 3 = 3840x2160, 20000 paths of config 2's mix with 25% feathered clockwise
 * fills (feather 1..40 px), 20% advanced blends and 64 transformed 256x256
 * images.
 * path_scale < 1 keeps only that fraction of the draws (parity-test sizes).
 * The stream owns its arrays until rive_stream_free. */
int rive_stream_generate(int config, uint64_t seed, uint32_t width, uint32_t height,
                         uint32_t draw_count, rive_stream** out);
void rive_stream_free(rive_stream* stream);

/* Renders a stream through the C++ front end (RiveRenderer -> RenderContext
 * -> RenderContextHIPImpl) into `target`, GEOMETRY mode.  Like the
 * reference's RenderContext, the front end splits the frame into logical
 * flushes where a flush would exceed MaxPathID(1) - 1 = 30719 paths or clip
 * IDs, 65535 contours, 2048 * 2048 - 25 tessellation vertices or 2048
 * gradient rows (render_context.cpp:43-53,484-492,517-541,617-654;
 * rive_renderer.cpp:496-570); each logical flush is one rive_hip_flush, the
 * later ones loading the target. */
int rive_hip_render_stream(rive_hip_context* ctx, const rive_stream* stream,
                           rive_hip_target* target);

/* rive_hip_render_stream that stops after the frame's first max_flushes
 * logical flushes (all of them when max_flushes == 0) and reports the
 * frame's logical flush count in *flush_count (when not NULL): lets a
 * caller inspect each logical flush's records with rive_hip_debug_readback. */
int rive_hip_render_stream_ex(rive_hip_context* ctx, const rive_stream* stream, rive_hip_target* target,
                              uint32_t max_flushes, uint32_t* flush_count);

/* Renders n streams into n targets with batched flushes: each stream is
 * replayed through RiveRenderer as its own canvas, and consecutive canvases
 * share one flush while their contours fit the 16-bit contour IDs
 * (rive_stream_contour_bound sums to at most 65535); a flush then renders all
 * of its canvases at once. */
int rive_hip_render_streams(rive_hip_context* ctx, const rive_stream* const* streams,
                            rive_hip_target* const* targets, uint32_t n);

/* An upper bound of the contours a stream's draws and clips produce (its
 * move verbs, plus one per image draw). */
uint32_t rive_stream_contour_bound(const rive_stream* stream);

/* Seeded batch of canvases (std::mt19937_64(seed)): per canvas, side
 * s = exp U[ln 64, ln 1024], aspect a = exp U[ln .5, ln 2],
 * w = round(s sqrt a), h = round(s / sqrt a), and round(exp U[ln 10, ln 2000])
 * closed cubic fills drawn like config 1 over w x h.  out[0..n) each own their
 * arrays until rive_stream_free. */
int rive_stream_generate_batch(uint64_t seed, uint32_t n, rive_stream** out);

/* A frame captured in a HIP graph: the frame's flushes (the descriptors of
 * rive_hip_flush_frame, whose records must stay as they are) run once
 * eagerly, recording every device total the host reads to size buffers and
 * launches; the same host logic then runs again under stream capture,
 * reading the recorded totals, so the graph holds every kernel and copy of
 * the frame and no host readback.  Launches are asynchronous on the
 * context's stream (rive_hip_sync waits).  No counterpart in the
 * reference: the batched path's frame capture.  Patch-batch flushes are
 * unsupported. */
typedef struct rive_hip_frame_graph rive_hip_frame_graph;
int rive_hip_frame_graph_create(rive_hip_context* ctx, const rive_hip_flush_desc* descs, uint32_t n,
                                rive_hip_frame_graph** out);
int rive_hip_frame_graph_launch(rive_hip_frame_graph* graph);
void rive_hip_frame_graph_destroy(rive_hip_frame_graph* graph);

/* ------------------------------------------------------------------------ */
/* Animated scenes and the serving loop.                                    */
/* ------------------------------------------------------------------------ */

/* A transform tree of nodes with keyframed affine animation and closed cubic
 * paths hanging off the nodes, generated from a seed (std::mt19937_64 with
 * the mappings above): 50-500 nodes of depth at most 6, 2-10 keys per node
 * over 4 s, each key easing linearly or along a cubic bezier towards the
 * next; 20-1500 paths of 8-32 cubics over [-256, 256]^2 with solid colours.
 * Node transforms and their composition follow the reference runtime
 * (transform_component.cpp:63-84; cubic_ease_interpolator.cpp:5-18;
 * cubic_interpolator_solver.cpp:12-95).  Not a reference interface: the
 * serving workload's scene source. */
typedef struct rive_scene rive_scene;

typedef struct rive_scene_info_t
{
    uint32_t node_count, path_count;
    uint64_t cubic_count;
    float duration; /* seconds */
    float extent;   /* side of the square the paths live in */
} rive_scene_info_t;

typedef struct rive_scene_key
{
    float time;
    float value[5]; /* x, y, rotation (radians), scaleX, scaleY */
    uint32_t cubic; /* 0 linear, 1 cubic-bezier ease towards the next key */
    float ease[4];  /* x1, y1, x2, y2 */
} rive_scene_key;

int rive_scene_generate(uint64_t seed, rive_scene** out);
void rive_scene_free(rive_scene* scene);
int rive_scene_info(const rive_scene* scene, rive_scene_info_t* info);
/* Node `node`'s parent (-1 for the root) and keys (min(capacity, count)
 * written; *key_count gets the count). */
int rive_scene_node(const rive_scene* scene, uint32_t node, int32_t* parent, rive_scene_key* keys,
                    uint32_t capacity, uint32_t* key_count);
/* World transforms of every node at time t, 6 floats each ([a b c d tx ty]). */
int rive_scene_world_transforms(const rive_scene* scene, float t, float* out);
/* The scene at time t as a draw stream for a width x height target (the
 * scene extent fits the shorter side, centred); free with rive_stream_free. */
int rive_scene_stream(const rive_scene* scene, float t, uint32_t width, uint32_t height, rive_stream** out);
/* The cubic-bezier easing of factor f (the scene keys' interpolation). */
float rive_ease_cubic(float x1, float y1, float x2, float y2, float f);
/* The serving loop's estimate of the HBM a request needs. */
uint64_t rive_scene_request_bytes(const rive_scene* scene, uint32_t width, uint32_t height);

typedef struct rive_hip_request
{
    uint32_t scene;   /* index into the scene array */
    float time;       /* animation time, seconds */
    uint32_t width, height;
    double arrival_s; /* arrival, seconds after the loop starts; non-decreasing */
} rive_hip_request;

typedef struct rive_hip_serve_config
{
    uint64_t hbm_budget_bytes; /* admission limit per batch (estimated) */
    uint32_t max_batch;        /* requests per batch */
} rive_hip_serve_config;

typedef struct rive_hip_request_result
{
    double arrival_s, admit_s, done_s; /* done: the batch's pixels are in HBM */
    uint32_t batch, batch_size;
    uint64_t batch_bytes;
} rive_hip_request_result;

/* Serves n timed requests in real time with continuous batching: whenever
 * the device is free, every request that has arrived is admitted in arrival
 * order while the batch's estimated HBM stays within the budget and its size
 * within max_batch, and the batch renders as one batched flush (one canvas
 * per request).  Output targets stay resident per size and are reused.
 * pixels: NULL, or per request NULL or width * height * 4 bytes that receive
 * the request's RGBA8 result.  Not a reference interface: the serving loop
 * around RenderContext::flush. */
int rive_hip_serve(rive_hip_context* ctx, const rive_scene* const* scenes, uint32_t scene_count,
                   const rive_hip_request* requests, uint32_t n, const rive_hip_serve_config* config,
                   rive_hip_request_result* results, uint8_t* const* pixels);

/* ------------------------------------------------------------------------ */
/* Deferred 2D command stream (record / replay).                             */
/* ------------------------------------------------------------------------ */

/* A recorded sequence of rive::Renderer calls as the reference's wire
 * format (render_commands.hpp:17-350, command_stream.hpp:34-165): opcode
 * byte + plain-data payload per command, variable-size data in an 8-byte
 * aligned blob arena.  Recorded here from a draw stream: makePath per path,
 * per draw makePaint and its paint settings (with makeLinearGradient /
 * makeRadialGradient + paintShader for gradients), save, transform,
 * clipPath per clip, drawPath, restore.  Streams with images are
 * unsupported.  The frame (size, clear colour, dither) travels beside it. */
typedef struct rive_cmd_buffer rive_cmd_buffer;

int rive_cmd_record_stream(const rive_stream* stream, rive_cmd_buffer** out);
void rive_cmd_free(rive_cmd_buffer* buffer);
/* The buffer's bytes (valid until it is freed). */
int rive_cmd_bytes(const rive_cmd_buffer* buffer, const uint8_t** commands, uint64_t* command_bytes,
                   const uint8_t** blobs, uint64_t* blob_bytes);
/* A buffer holding copies of the given bytes (e.g. read from a file). */
int rive_cmd_from_bytes(const uint8_t* commands, uint64_t command_bytes, const uint8_t* blobs,
                        uint64_t blob_bytes, uint32_t width, uint32_t height, uint32_t clear_color,
                        uint32_t dither, rive_cmd_buffer** out);
/* Replays the commands through RiveRenderer into `target` and flushes;
 * RIVE_HIP_ERROR_INVALID_ARGUMENT on an unknown opcode, a truncated payload,
 * a blob out of range or a handle naming no resource. */
int rive_hip_render_commands(rive_hip_context* ctx, const rive_cmd_buffer* buffer, rive_hip_target* target);

#ifdef __cplusplus
}
#endif

#endif /* RIVE_HIP_H */
