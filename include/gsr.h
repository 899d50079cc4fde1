/*
 * gsr.h — C ABI of the MI355X (gfx950) Gaussian-splatting renderer.
 *
 * One context owns one HIP stream on one GPU, the resident scenes and the
 * per-frame buffers.  Every function returns an int status (GSR_OK = 0);
 * on failure gsr_last_error(ctx) holds a message.  No C++ exception crosses
 * this boundary.  A context is not re-entrant: use one per host thread.
 *
 * Device-pointer arguments ("_dev") are HIP device addresses (for instance
 * torch.Tensor.data_ptr() of a CUDA tensor).  Host-pointer arguments are
 * plain host memory; the library copies them and never keeps them.
 *
 * Reference interfaces each entry point replaces are cited as file:line
 * inside KoS-Y1/Web-GS-Renderer.
 */
#ifndef GSR_H
#define GSR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define GSR_OK                 0
#define GSR_ERR_ARGUMENT       1   /* invalid argument (null, size, range)        */
#define GSR_ERR_DEVICE         2   /* no HIP device, or a HIP call failed          */
#define GSR_ERR_ALLOC          3   /* device allocation failed / exceeds the limit */
#define GSR_ERR_NOT_FOUND      4   /* unknown scene name                           */
#define GSR_ERR_CAPACITY       5   /* tile instances exceeded the instance capacity */
#define GSR_ERR_KERNEL         6   /* a kernel reported an internal failure        */
#define GSR_ERR_STATE          7   /* call out of order (e.g. backward before forward) */

/* ---- per-frame camera block ------------------------------------------ */
/*
 * Byte-for-byte the 176-byte GlobalUniforms block of the reference
 * (src/shaders/preprocess.wgsl:2-12, packed by src/renderer/renderer.js:773-797):
 * matrices are column-major float32 as produced by wgpu-matrix
 * (view = lookAt, view_proj = perspective(fovY, aspect, 0.2, 1000) * view,
 * src/renderer/camera.js:87-98, 135-142).  `count` is ignored on input (the
 * library fills in the scene's splat count); width/height are read as
 * (uint32_t)texture_size[0], (uint32_t)texture_size[1] like the shaders do.
 */
typedef struct gsr_camera {
    float    view[16];
    float    view_proj[16];
    float    eye[3];
    uint32_t count;
    float    focal[2];
    float    tan_fov[2];
    float    texture_size[2];
    uint32_t pad[2];
} gsr_camera;

/* ---- render options -------------------------------------------------- */
#define GSR_KEY_AUTO 0   /* 32-bit keys when every (view,tile) id fits 16 bits, else 64 */
#define GSR_KEY_32   4
#define GSR_KEY_64   8

typedef struct gsr_options {
    uint32_t tile_size;        /* 8, 16 or 32 pixels; 0 = 16 (the reference's 16x16)            */
    uint32_t key_bytes;        /* GSR_KEY_AUTO / GSR_KEY_32 / GSR_KEY_64                          */
    uint64_t instance_capacity;/* tile-instance slots; 0 = 64 * splat count (renderer.js:16,536) */
    uint32_t timing;           /* 1 = hipEvents around every stage (gsr_get_timings);            */
                               /* 2 = around the raster stages only (two events per raster)     */
    uint32_t save_state;       /* 1 = keep per-pixel T_final / last contributor for backward     */
    uint32_t raster_waves_per_cu; /* batches: persistent raster waves per CU; 0 = 32              */
    uint32_t sort_blocks_per_cu;  /* onesweep blocks per CU: 1, 2 or 4 (1024-, 512-, 256-thread); 0 = 1 */
    uint32_t sort_digit_bits;     /* radix digit width: 8 to 11 bits (256 to 2048 buckets); 0 = 8 for
                                     32-bit keys, for 64-bit keys the narrowest of 8..10 with the
                                     fewest passes                                                  */
} gsr_options;

/* ---- timings --------------------------------------------------------- */
typedef struct gsr_timing_row {
    char   label[32];   /* same labels as the reference profiler rows (profiler.js:96-123) */
    double ms;
} gsr_timing_row;

typedef struct gsr_context gsr_context;

/* ---- lifecycle (replaces src/gpu/device.js:3-40 requestDevice) -------- */
int         gsr_create(gsr_context** out, int device_ordinal);
void        gsr_destroy(gsr_context* ctx);
/* Message of the last failure on ctx; with ctx == NULL, of the last failed
 * gsr_create on this thread.  Never NULL. */
const char* gsr_last_error(const gsr_context* ctx);
/* Waits for the context's stream.  When a gsr_forward ran since its counters
 * were last read, also reads them and reports what gsr_render would:
 * GSR_ERR_CAPACITY (instances above the capacity) or GSR_ERR_KERNEL (a bounded
 * wait gave up); the failed forward's saved state is then dropped. */
int         gsr_synchronize(gsr_context* ctx);
/* The context's HIP stream (hipStream_t) so a caller can order its own work. */
void*       gsr_stream(gsr_context* ctx);

/* ---- scenes (replaces renderer.js:461-479 setGs / uploadGsData) ------- */
/*
 * `packed` is the loader's block layout (src/loader/loader.js:75-111):
 * pos[3N] xyz-interleaved | log-scale[3N] | rot[4N] (w,x,y,z) | f_dc[3N] |
 * opacity-logit[N] | f_rest[45N] (per splat: R0..14, G0..14, B0..14).
 * It is copied to the device and stored planar ([59][N] float32).
 */
int gsr_scene_upload(gsr_context* ctx, const char* name, const float* packed, uint32_t count);
int gsr_scene_select(gsr_context* ctx, const char* name);
int gsr_scene_free(gsr_context* ctx, const char* name);
/* Device address of a resident scene's planar [59][N] parameters. */
int gsr_scene_params(gsr_context* ctx, const char* name, const float** params_dev, uint32_t* count);

/* ---- whole-frame forward (replaces renderer.js:481-508, 822-932) ------ */
/*
 * Renders the selected scene for one camera.  Writes planar RGB float32
 * [3][H][W] (row y = pixel y of preprocess.wgsl:207-209, no flip) into
 * out_rgb_host (host, may be NULL) and/or out_rgb_dev (device, may be NULL).
 * Blocks until the frame is done.  opts may be NULL (defaults).
 */
int gsr_render(gsr_context* ctx, const gsr_camera* cam, const gsr_options* opts,
               float* out_rgb_host, float* out_rgb_dev);

/*
 * Forward from caller-owned planar device parameters ([59][N], see
 * gsr_scene_upload) — the entry the differentiable harness uses.  Enqueues
 * on the context stream and returns without waiting; an instance overflow or
 * kernel failure of the frame is reported by the next gsr_synchronize (or
 * gsr_last_instance_count).  When opts->save_state is 1 the per-pixel state
 * for gsr_backward is kept.
 */
int gsr_forward(gsr_context* ctx, const float* params_dev, uint32_t count,
                const gsr_camera* cam, const gsr_options* opts, float* out_rgb_dev);

/* Number of tile instances (raw scan total I) of the last forward; syncs. */
int gsr_last_instance_count(gsr_context* ctx, uint64_t* instances);

/* Per-pass GPU times of the last frame(s) with opts.timing = 1, or of the raster passes with
   opts.timing = 2 (profiler.js:52-123). */
int gsr_get_timings(gsr_context* ctx, gsr_timing_row* rows, int capacity, int* count);

/* rgba8 conversion with WebGPU unorm8 semantics (raseter.wgsl:19,109): clamp to
 * [0,1], x255, round to nearest; alpha 255.  flip_y = 1 applies blit.wgsl:22.
 * rgb_dev == NULL converts the context's last rendered frame (which must be
 * width x height). */
int gsr_to_rgba8(gsr_context* ctx, const float* rgb_dev, uint32_t width, uint32_t height,
                 int flip_y, uint8_t* out_rgba_host);

/* ---- stage entry points (parity; device arrays; synchronous) ---------- */
/* preprocess.wgsl:58-144.  Outputs per splat: tiles_touched u32[N],
 * meta u32[4N] ((y0<<16)|x0, (y1<<16)|x1, depth16, 0), conic_opacity f32[4N],
 * xy f32[2N], rgb f32[4N] (r,g,b,0).  Splats culled by the near plane have
 * every output zero; splats culled by the frustum keep xy/conic_opacity and
 * have zero rgb/meta. */
int gsr_preprocess(gsr_context* ctx, const float* params_dev, uint32_t count,
                   const gsr_camera* cam, uint32_t tile_size,
                   uint32_t* tiles_touched_dev, uint32_t* meta_dev, float* conic_opacity_dev,
                   float* xy_dev, float* rgb_dev);

/* offset_scan.wgsl:22-68: exclusive scan of counts[0..n), out[n] = total.
 * in/out may alias.  total_dev (u32[1]) may be NULL. */
int gsr_scan_offsets(gsr_context* ctx, const uint32_t* counts_dev, uint32_t n,
                     uint32_t* offsets_dev, uint32_t* total_dev);

/* emit.wgsl:28-67: one (key, value) per covered tile, row-major, at slot
 * offsets[g]+k, slots >= capacity dropped.  key_bytes 4: (tile<<16)|depth16
 * as u32; 8: the same value as u64. */
int gsr_emit(gsr_context* ctx, const uint32_t* offsets_dev, const uint32_t* meta_dev,
             uint32_t count, uint32_t width, uint32_t tile_size, uint64_t capacity,
             uint32_t key_bytes, void* keys_dev, uint32_t* values_dev);

/* radix_parallel.wgsl:26-160 + renderer.js:869-917: stable LSD sort of
 * (key, value) pairs on bits [begin_bit, end_bit).  keys0/values0 hold the
 * input; keys1/values1 are the ping-pong buffers.  *result_index tells which
 * pair (0 or 1) holds the sorted output.  key_bytes is 4 or 8. */
int gsr_sort_pairs(gsr_context* ctx, uint32_t key_bytes, void* keys0_dev, uint32_t* values0_dev,
                   void* keys1_dev, uint32_t* values1_dev, uint64_t n,
                   uint32_t begin_bit, uint32_t end_bit, int* result_index);

/* gsr_sort_pairs with an explicit onesweep geometry: blocks_per_cu 1, 2 or 4
 * (1024-, 512- or 256-thread blocks) and digit_bits 8 to 11 (the sort
 * dimensions of gsr_options / gsr_autotune_batch).  Same result for every
 * geometry.  No reference counterpart beyond gsr_sort_pairs'. */
int gsr_sort_pairs_geometry(gsr_context* ctx, uint32_t key_bytes, void* keys0_dev, uint32_t* values0_dev,
                            void* keys1_dev, uint32_t* values1_dev, uint64_t n,
                            uint32_t begin_bit, uint32_t end_bit, uint32_t blocks_per_cu,
                            uint32_t digit_bits, int* result_index);

/* gsr_sort_pairs (1 block per CU, 8-bit digits, the same result) with no
 * device allocation: no reference counterpart; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its scratch (counters, digit histograms, look-back
 * words) is one fixed array of GSR_SORT_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends, which bounds n: at most 133,955,584 keys of 4 bytes
 * or 92,094,464 of 8 bytes (larger n: GSR_ERR_ARGUMENT, nothing launched).
 * Every call on a GPU shares that array, so calls that overlap on one GPU
 * (from several contexts or threads) race; calls in sequence are safe. */
#define GSR_SORT_NOALLOC_SCRATCH_BYTES (16u << 20)
int gsr_sort_pairs_noalloc(gsr_context* ctx, uint32_t key_bytes, void* keys0_dev, uint32_t* values0_dev,
                           void* keys1_dev, uint32_t* values1_dev, uint64_t n,
                           uint32_t begin_bit, uint32_t end_bit, int* result_index);

/* tile_ranges.wgsl:7-35: ranges[2t] = first, ranges[2t+1] = last+1 of tile t
 * (= key >> 16) over sorted keys[0..n); untouched tiles are (0,0). */
int gsr_tile_ranges(gsr_context* ctx, const void* keys_dev, uint32_t key_bytes, uint64_t n,
                    uint32_t num_tiles, uint32_t* ranges_dev);

/* raseter.wgsl:35-117: front-to-back blend per tile.  Writes planar RGB
 * f32[3][H][W]; T_final_dev (f32[H][W]) and last_dev (u32[H][W], the
 * 1-based position in the tile's range of the last blended splat, 0 = none)
 * may be NULL.  `count` is the length of the per-splat arrays and
 * `instances` the length of values_dev; they bound every read. */
int gsr_rasterize(gsr_context* ctx, uint32_t width, uint32_t height, uint32_t tile_size,
                  uint32_t count, uint64_t instances,
                  const uint32_t* ranges_dev, const uint32_t* values_dev,
                  const float* conic_opacity_dev, const float* xy_dev, const float* rgb_dev,
                  float* out_rgb_dev, float* T_final_dev, uint32_t* last_dev);

/* ---- differentiation (no reference counterpart: the reference is forward-only) ---- */
/*
 * Gradients of the last gsr_forward (which must have had save_state = 1, and
 * the same params pointer, count and camera, else GSR_ERR_STATE) with respect
 * to the planar parameters.  A backward enqueued behind a forward that failed
 * (see gsr_forward) writes quiet NaNs into grad_params_dev and grad2d_dev.
 * dL_drgb_dev: f32[3][H][W].  grad_params_dev: f32[59][N], overwritten.
 * grad2d_dev (may be NULL) receives the per-splat 2-D gradient, f32[N][12]:
 * pixel x, y; conic A, B, C; opacity; r, g, b; 3 unused.  Enqueues on the
 * context stream like gsr_forward.  Gradients accumulate with float atomics,
 * so the last bits may differ between runs.
 */
int gsr_backward(gsr_context* ctx, const float* params_dev, uint32_t count,
                 const gsr_camera* cam, const float* dL_drgb_dev, float* grad_params_dev,
                 float* grad2d_dev);

/* ---- multi-view batch (no reference counterpart: the reference renders one view) ---- */
/*
 * Renders `views` cameras of the same parameters in one launch sequence with
 * keys view|tile|depth (one sort serves all views) and a persistent (view,
 * tile) work queue.  out_rgb_dev receives the views back to back, each planar
 * [3][H_v][W_v]; out_offsets (host, views+1 entries, may be NULL) receives the
 * float offset of each view's image and the total.  out_rgb_dev == NULL only
 * fills out_offsets.  use_graph = 1 captures the frame as a hipGraph after the
 * first eager run of a batch shape and replays it while the shape (splat
 * count, parameter and output pointers, view resolutions, options) is
 * unchanged; cameras may change between replays.  When tile instances
 * overflow the capacity (opts->instance_capacity, else a planner that starts
 * at 16 per splat-view and grows) the frame is re-run once with room for the
 * measured count.  Blocks until the frame is done.  opts->save_state = 1
 * keeps every view's per-pixel state for gsr_backward_batch.
 */
int gsr_render_batch(gsr_context* ctx, const float* params_dev, uint32_t count,
                     const gsr_camera* cams, uint32_t views, const gsr_options* opts,
                     int use_graph, float* out_rgb_dev, uint64_t* out_offsets);

/* Tile instances of the last batch frame (raw scan total over all views). */
int gsr_batch_instance_count(gsr_context* ctx, uint64_t* instances);

/* Instance capacity (key/value slots of the sort buffers) the last batch frame
 * ran with: opts->instance_capacity, else the planner's (16 per splat-view,
 * raised to 1.25x the measured count after an overflow).  The reference sizes
 * these buffers at 64 per splat (renderer.js:16, 536-577). */
int gsr_batch_capacity(gsr_context* ctx, uint64_t* slots);

/*
 * Autotuner for a batch shape: times tile size {8, 16, 32} x persistent raster
 * waves per CU {8, 16, 32} x onesweep blocks per CU {1, 2, 4} x radix digit
 * width {8, 9, 10, 11} bits (three timed
 * frames each, HIP events) and writes the fastest into *inout (other fields
 * kept).  out_rgb_dev is used as the scratch output.  The tile size changes
 * which splats reach a pixel's tile at the fringe of their 3-sigma square, so
 * images are comparable only between renders with the same tile size.
 */
int gsr_autotune_batch(gsr_context* ctx, const float* params_dev, uint32_t count,
                       const gsr_camera* cams, uint32_t views, gsr_options* inout,
                       float* out_rgb_dev, double* best_ms);

/*
 * Gradients of the last gsr_render_batch (which must have had save_state = 1
 * and the same params, count and cameras), summed over its views.
 * dL_drgb_dev: in the batch output layout (out_offsets).  grad_params_dev:
 * f32[59][N], overwritten.  grad2d_dev (may be NULL): the 2-D gradient of
 * every (view, splat), f32[views][N][12] in gsr_backward's layout.  Enqueues
 * on the context stream and returns without waiting.  Gradients accumulate
 * with float atomics, so the last bits may differ between runs.
 */
int gsr_backward_batch(gsr_context* ctx, const float* params_dev, uint32_t count,
                       const gsr_camera* cams, uint32_t views, const float* dL_drgb_dev,
                       float* grad_params_dev, float* grad2d_dev);

#ifdef __cplusplus
}
#endif
#endif /* GSR_H */
