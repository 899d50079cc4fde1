// pipeline.hip — the per-frame pass graph and the stage entry points.
//
// Pass order and sizing follow src/renderer/renderer.js:531-604 and 822-932:
// preprocess -> offset scan -> emit -> sort -> tile ranges -> raster.  The
// reference's indirect-dispatch step (indirect_arg.wgsl:14-28) becomes a
// device-side count (FrameCounters::n = min(I, capacity)) that the sort and
// range kernels read, so a frame never waits on the host mid-way.
#include <cstring>

#include "gsr_internal.h"

namespace gsr {

struct FramePlan {
    uint32_t W, H, tile, tiles_x, tiles_y, num_tiles, key_bytes, end_bit, sort_blocks, sort_digits;
    uint64_t capacity;
};

int plan_frame(Context* c, uint32_t count, const gsr_camera* cam, const gsr_options* opts, FramePlan* p) {
    if (!cam) return fail(c, GSR_ERR_ARGUMENT, "null camera");
    const float fw = cam->texture_size[0], fh = cam->texture_size[1];
    if (!(fw >= 1.0f && fh >= 1.0f && fw <= 65536.0f && fh <= 65536.0f))
        return fail(c, GSR_ERR_ARGUMENT, "texture size %gx%g out of range", fw, fh);
    p->W = (uint32_t)fw;
    p->H = (uint32_t)fh;
    p->tile = (opts && opts->tile_size) ? opts->tile_size : 16;
    if (p->tile != 8 && p->tile != 16 && p->tile != 32)
        return fail(c, GSR_ERR_ARGUMENT, "tile size %u not in {8, 16, 32}", p->tile);
    p->tiles_x = ceil_div(p->W, p->tile);
    p->tiles_y = ceil_div(p->H, p->tile);
    p->num_tiles = p->tiles_x * p->tiles_y;
    uint32_t tbits = 0;
    while ((1ull << tbits) < (uint64_t)p->num_tiles) ++tbits;
    p->end_bit = 16 + tbits;
    const uint32_t kb = opts ? opts->key_bytes : GSR_KEY_AUTO;
    if (kb == GSR_KEY_AUTO) p->key_bytes = p->end_bit <= 32 ? 4 : 8;
    else if (kb == GSR_KEY_32 || kb == GSR_KEY_64) p->key_bytes = kb;
    else return fail(c, GSR_ERR_ARGUMENT, "key_bytes %u not in {0, 4, 8}", kb);
    if (p->key_bytes == 4 && p->end_bit > 32)
        return fail(c, GSR_ERR_ARGUMENT, "%u tiles do not fit a 32-bit (tile<<16)|depth key", p->num_tiles);
    uint64_t cap = (opts && opts->instance_capacity) ? opts->instance_capacity : (uint64_t)count * kInstanceFactor;
    if (cap == 0) cap = 1;
    if (cap > kMaxSortKeys) {
        if (opts && opts->instance_capacity)
            return fail(c, GSR_ERR_ARGUMENT, "instance capacity %llu exceeds the sort limit %llu",
                        (unsigned long long)cap, (unsigned long long)kMaxSortKeys);
        cap = kMaxSortKeys;
    }
    p->capacity = cap;
    p->sort_blocks = (opts && opts->sort_blocks_per_cu) ? opts->sort_blocks_per_cu : 1;
    if (p->sort_blocks != 1 && p->sort_blocks != 2 && p->sort_blocks != 4)
        return fail(c, GSR_ERR_ARGUMENT, "sort blocks per CU %u not in {1, 2, 4}", p->sort_blocks);
    p->sort_digits = (opts && opts->sort_digit_bits) ? opts->sort_digit_bits
                                                      : default_digit_bits(p->key_bytes, p->end_bit);
    if (p->sort_digits < 8 || p->sort_digits > 11)
        return fail(c, GSR_ERR_ARGUMENT, "sort digit width %u not in 8..11", p->sort_digits);
    return GSR_OK;
}

// Enqueues one full forward frame; outputs stay in the context buffers
// ("image", "tfinal", "last", sorted pairs at index c->state_sorted_index).
int enqueue_forward(Context* c, const float* params, uint32_t count, const gsr_camera* cam, const FramePlan& p,
                    bool save_state, float* image_override) {
    const size_t N = count;
    uint32_t* tiles = (uint32_t*)buffer(c, "tiles", N * 4);
    uint32_t* offsets = (uint32_t*)buffer(c, "offsets", (N + 1) * 4);
    uint4* meta = (uint4*)buffer(c, "meta", N * 16);
    SplatRec* rec = (SplatRec*)buffer(c, "splat_rec", N * sizeof(SplatRec));
    void* k0 = buffer(c, "keys0", p.capacity * p.key_bytes);
    void* k1 = buffer(c, "keys1", p.capacity * p.key_bytes);
    uint32_t* v0 = (uint32_t*)buffer(c, "vals0", p.capacity * 4);
    uint32_t* v1 = (uint32_t*)buffer(c, "vals1", p.capacity * 4);
    uint32_t* ranges = (uint32_t*)buffer(c, "ranges", (size_t)p.num_tiles * 8);
    uint32_t* order = (uint32_t*)buffer(c, "frame_tile_order", (size_t)p.num_tiles * 4);
    const size_t px = (size_t)p.W * p.H;
    float* image = image_override ? image_override : (float*)buffer(c, "image", px * 12);
    float* tfinal = save_state ? (float*)buffer(c, "tfinal", px * 4) : nullptr;
    uint32_t* last = save_state ? (uint32_t*)buffer(c, "last", px * 4) : nullptr;
    FrameCounters* counters = (FrameCounters*)buffer(c, "counters", sizeof(FrameCounters));
    if (!tiles || !offsets || !meta || !rec || !order || !k0 || !k1 || !v0 || !v1 || !ranges || !image ||
        (save_state && (!tfinal || !last)) || !counters)
        return GSR_ERR_ALLOC;
    int rc;
    if (count == 0) {
        // no splats: a black frame (T_final 1, no last contributor)
        if ((rc = launch_fill_u32(c, (uint32_t*)counters, 0u, sizeof(FrameCounters) / 4))) return rc;
        if ((rc = launch_fill_u32(c, (uint32_t*)image, 0u, px * 3))) return rc;
        if (save_state) {
            if ((rc = launch_fill_u32(c, (uint32_t*)tfinal, 0x3F800000u, px))) return rc;
            if ((rc = launch_fill_u32(c, last, 0u, px))) return rc;
        }
        if (!image_override) {
            c->last_frame_w = p.W;
            c->last_frame_h = p.H;
        }
        c->have_state = save_state;
        c->state_params = params;
        std::memcpy(&c->state_cam, cam, sizeof(CamK));
        c->state_cam.count = 0;
        c->state_count = 0;
        c->state_w = p.W;
        c->state_h = p.H;
        c->state_tile = p.tile;
        c->state_key_bytes = p.key_bytes;
        c->state_capacity = p.capacity;
        c->state_sorted_index = 0;
        return GSR_OK;
    }
    CamK ck;
    std::memcpy(&ck, cam, sizeof(ck));
    ck.count = count;
    uint32_t* scan_status = scan_status_buffer(c, count);
    if (!scan_status) return GSR_ERR_ALLOC;
    const FillRegion sort_hist = sort_hist_region(c, 0, p.end_bit, p.sort_digits);
    if (!sort_hist.count) return GSR_ERR_ALLOC;
    // the preprocess grid also clears the counters, the scan's look-back
    // words, the tile ranges and the sort's histograms (all read by later
    // kernels only)
    const FillRegion clear[4] = {{(uint32_t*)counters, sizeof(FrameCounters) / 4},
                                 {scan_status, scan_status_words(count)},
                                 {ranges, (size_t)p.num_tiles * 2},
                                 sort_hist};
    mark(c, "preprocess");
    if ((rc = launch_preprocess(c, params, count, ck, p.tile, tiles, meta, nullptr, nullptr, nullptr, rec, clear, 4)))
        return rc;
    mark(c, "offset scan");
    if ((rc = launch_scan(c, tiles, count, offsets, counters, p.capacity, true))) return rc;
    mark(c, "emit");
    if ((rc = launch_emit(c, offsets, meta, count, p.tiles_x, p.capacity, p.key_bytes, k0, v0))) return rc;
    mark(c, "sort");
    int idx = 0;
    if ((rc = launch_sort(c, p.key_bytes, k0, v0, k1, v1, &counters->n, 0, p.capacity, 0, p.end_bit,
                          p.sort_blocks, counters, &idx, p.sort_digits, true)))
        return rc;
    mark(c, "tile ranges");
    void* ks = idx ? k1 : k0;
    uint32_t* vs = idx ? v1 : v0;
    if ((rc = launch_ranges(c, ks, p.key_bytes, &counters->n, 0, p.num_tiles, ranges, false))) return rc;
    if ((rc = launch_tile_order(c, (const uint2*)ranges, p.num_tiles, order))) return rc;
    mark(c, "raster");
    if ((rc = launch_raster(c, p.W, p.H, p.tile, count, &counters->n, 0, ranges, vs, rec, image, tfinal, last,
                             order)))
        return rc;
    mark_end(c);
    if (!image_override) {
        c->last_frame_w = p.W;
        c->last_frame_h = p.H;
    }
    c->have_state = save_state;
    c->state_params = params;
    c->state_cam = ck;
    c->state_count = count;
    c->state_w = p.W;
    c->state_h = p.H;
    c->state_tile = p.tile;
    c->state_key_bytes = p.key_bytes;
    c->state_capacity = p.capacity;
    c->state_sorted_index = idx;
    return GSR_OK;
}

int check_frame(Context* c, uint64_t capacity) {
    FrameCounters* counters = (FrameCounters*)buffer(c, "counters", sizeof(FrameCounters));
    if (!counters) return GSR_ERR_ALLOC;
    c->frame_unchecked = false;
    GSR_HIP(c, hipMemcpyAsync(c->host_counters, counters, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream));
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    finish_frame_timings(c);
    if (c->host_counters->error) {
        c->have_state = false;
        return fail(c, GSR_ERR_KERNEL, "a look-back wait exceeded its bound (code %u)", c->host_counters->error);
    }
    if (c->host_counters->instances > capacity) {
        c->have_state = false;
        return fail(c, GSR_ERR_CAPACITY, "%llu tile instances exceed the instance capacity %llu",
                    (unsigned long long)c->host_counters->instances, (unsigned long long)capacity);
    }
    return GSR_OK;
}

}  // namespace gsr

using namespace gsr;

static Context* ctx_of(gsr_context* ctx) { return reinterpret_cast<Context*>(ctx); }

extern "C" {

int gsr_forward(gsr_context* ctx, const float* params_dev, uint32_t count, const gsr_camera* cam,
                const gsr_options* opts, float* out_rgb_dev) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!params_dev && count > 0) return fail(c, GSR_ERR_ARGUMENT, "gsr_forward: no parameters");
    hipSetDevice(c->device);
    FramePlan p;
    int rc = plan_frame(c, count, cam, opts, &p);
    if (rc) return rc;
    c->timing_on = opts ? opts->timing : 0;
    if (c->marks.size() > 4096) {  // bound the event backlog of long timed runs
        GSR_HIP(c, hipStreamSynchronize(c->stream));
        finish_frame_timings(c);
    }
    c->have_state = false;
    rc = enqueue_forward(c, params_dev, count, cam, p, opts && opts->save_state, nullptr);
    if (rc) return rc;
    // counters are read by the next gsr_synchronize (or instance-count query);
    // a backward enqueued before that checks them on the device
    c->frame_unchecked = true;
    if (out_rgb_dev) {
        float* image = (float*)buffer(c, "image", (size_t)p.W * p.H * 12);
        GSR_HIP(c, hipMemcpyAsync(out_rgb_dev, image, (size_t)p.W * p.H * 12, hipMemcpyDeviceToDevice, c->stream));
    }
    return GSR_OK;
}

int gsr_last_instance_count(gsr_context* ctx, uint64_t* instances) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!instances) return fail(c, GSR_ERR_ARGUMENT, "null output");
    int rc = check_frame(c, c->state_capacity ? c->state_capacity : ~0ull);
    *instances = c->host_counters->instances;
    return rc;
}

int gsr_render(gsr_context* ctx, const gsr_camera* cam, const gsr_options* opts, float* out_rgb_host,
               float* out_rgb_dev) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    auto it = c->scenes.find(c->current);
    if (c->current.empty() || it == c->scenes.end()) return fail(c, GSR_ERR_NOT_FOUND, "no scene selected");
    hipSetDevice(c->device);
    FramePlan p;
    int rc = plan_frame(c, it->second.count, cam, opts, &p);
    if (rc) return rc;
    c->timing_on = opts ? opts->timing : 0;
    c->have_state = false;
    rc = enqueue_forward(c, it->second.params, it->second.count, cam, p, opts && opts->save_state, nullptr);
    if (rc) return rc;
    const size_t bytes = (size_t)p.W * p.H * 12;
    float* image = (float*)buffer(c, "image", bytes);
    if (out_rgb_dev) GSR_HIP(c, hipMemcpyAsync(out_rgb_dev, image, bytes, hipMemcpyDeviceToDevice, c->stream));
    if ((rc = check_frame(c, p.capacity))) return rc;
    if (out_rgb_host) GSR_HIP(c, hipMemcpy(out_rgb_host, image, bytes, hipMemcpyDeviceToHost));
    return GSR_OK;
}

int gsr_preprocess(gsr_context* ctx, const float* params_dev, uint32_t count, const gsr_camera* cam,
                   uint32_t tile_size, uint32_t* tiles_touched_dev, uint32_t* meta_dev, float* conic_opacity_dev,
                   float* xy_dev, float* rgb_dev) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!params_dev || !cam || !tiles_touched_dev || !meta_dev || !conic_opacity_dev || !xy_dev || !rgb_dev)
        return fail(c, GSR_ERR_ARGUMENT, "gsr_preprocess: null argument");
    if (tile_size != 8 && tile_size != 16 && tile_size != 32)
        return fail(c, GSR_ERR_ARGUMENT, "tile size %u not in {8, 16, 32}", tile_size);
    hipSetDevice(c->device);
    CamK ck;
    std::memcpy(&ck, cam, sizeof(ck));
    ck.count = count;
    int rc = launch_preprocess(c, params_dev, count, ck, tile_size, tiles_touched_dev, (uint4*)meta_dev,
                               (float4*)conic_opacity_dev, (float2*)xy_dev, (float4*)rgb_dev, nullptr);
    if (rc) return rc;
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    return GSR_OK;
}

int gsr_scan_offsets(gsr_context* ctx, const uint32_t* counts_dev, uint32_t n, uint32_t* offsets_dev,
                     uint32_t* total_dev) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!counts_dev || !offsets_dev) return fail(c, GSR_ERR_ARGUMENT, "gsr_scan_offsets: null argument");
    hipSetDevice(c->device);
    FrameCounters* counters = (FrameCounters*)buffer(c, "stage_counters", sizeof(FrameCounters));
    if (!counters) return GSR_ERR_ALLOC;
    int rc = launch_fill_u32(c, (uint32_t*)counters, 0u, sizeof(FrameCounters) / 4);
    if (!rc) rc = launch_scan(c, counts_dev, n, offsets_dev, counters, ~0ull);
    if (rc) return rc;
    if (total_dev)
        GSR_HIP(c, hipMemcpyAsync(total_dev, &counters->instances, 4, hipMemcpyDeviceToDevice, c->stream));
    GSR_HIP(c, hipMemcpyAsync(c->host_counters, counters, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream));
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    if (c->host_counters->error) return fail(c, GSR_ERR_KERNEL, "scan look-back wait exceeded its bound");
    return GSR_OK;
}

int gsr_emit(gsr_context* ctx, const uint32_t* offsets_dev, const uint32_t* meta_dev, uint32_t count,
             uint32_t width, uint32_t tile_size, uint64_t capacity, uint32_t key_bytes, void* keys_dev,
             uint32_t* values_dev) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!offsets_dev || !meta_dev || !keys_dev || !values_dev) return fail(c, GSR_ERR_ARGUMENT, "gsr_emit: null argument");
    if (key_bytes != 4 && key_bytes != 8) return fail(c, GSR_ERR_ARGUMENT, "key_bytes must be 4 or 8");
    if (tile_size != 8 && tile_size != 16 && tile_size != 32)
        return fail(c, GSR_ERR_ARGUMENT, "tile size %u not in {8, 16, 32}", tile_size);
    hipSetDevice(c->device);
    int rc = launch_emit(c, offsets_dev, (const uint4*)meta_dev, count, ceil_div(width, tile_size), capacity, key_bytes,
                         keys_dev, values_dev);
    if (rc) return rc;
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    return GSR_OK;
}

int gsr_sort_pairs_geometry(gsr_context* ctx, uint32_t key_bytes, void* keys0_dev, uint32_t* values0_dev,
                            void* keys1_dev, uint32_t* values1_dev, uint64_t n, uint32_t begin_bit, uint32_t end_bit,
                            uint32_t blocks_per_cu, uint32_t digit_bits, int* result_index) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!keys0_dev || !values0_dev || !keys1_dev || !values1_dev || !result_index)
        return fail(c, GSR_ERR_ARGUMENT, "gsr_sort_pairs: null argument");
    if (key_bytes != 4 && key_bytes != 8) return fail(c, GSR_ERR_ARGUMENT, "key_bytes must be 4 or 8");
    if (n > kMaxSortKeys) return fail(c, GSR_ERR_ARGUMENT, "%llu keys exceed the sort limit", (unsigned long long)n);
    hipSetDevice(c->device);
    FrameCounters* counters = (FrameCounters*)buffer(c, "stage_counters", sizeof(FrameCounters));
    if (!counters) return GSR_ERR_ALLOC;
    int rc = launch_fill_u32(c, (uint32_t*)counters, 0u, sizeof(FrameCounters) / 4);
    if (!rc) rc = launch_sort(c, key_bytes, keys0_dev, values0_dev, keys1_dev, values1_dev, nullptr, n,
                              n ? n : 1, begin_bit, end_bit, blocks_per_cu, counters, result_index, digit_bits);
    if (rc) return rc;
    GSR_HIP(c, hipMemcpyAsync(c->host_counters, counters, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream));
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    if (c->host_counters->error) return fail(c, GSR_ERR_KERNEL, "sort look-back wait exceeded its bound");
    return GSR_OK;
}

int gsr_sort_pairs(gsr_context* ctx, uint32_t key_bytes, void* keys0_dev, uint32_t* values0_dev, void* keys1_dev,
                   uint32_t* values1_dev, uint64_t n, uint32_t begin_bit, uint32_t end_bit, int* result_index) {
    return gsr_sort_pairs_geometry(ctx, key_bytes, keys0_dev, values0_dev, keys1_dev, values1_dev, n, begin_bit,
                                   end_bit, 1, 8, result_index);
}

int gsr_tile_ranges(gsr_context* ctx, const void* keys_dev, uint32_t key_bytes, uint64_t n, uint32_t num_tiles,
                    uint32_t* ranges_dev) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!keys_dev || !ranges_dev) return fail(c, GSR_ERR_ARGUMENT, "gsr_tile_ranges: null argument");
    if (key_bytes != 4 && key_bytes != 8) return fail(c, GSR_ERR_ARGUMENT, "key_bytes must be 4 or 8");
    if (n > 0xFFFFFFFFull) return fail(c, GSR_ERR_ARGUMENT, "too many keys");
    hipSetDevice(c->device);
    int rc = launch_ranges(c, keys_dev, key_bytes, nullptr, n, num_tiles, ranges_dev, true);
    if (rc) return rc;
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    return GSR_OK;
}

int gsr_rasterize(gsr_context* ctx, uint32_t width, uint32_t height, uint32_t tile_size, uint32_t count,
                  uint64_t instances, const uint32_t* ranges_dev, const uint32_t* values_dev,
                  const float* conic_opacity_dev, const float* xy_dev, const float* rgb_dev, float* out_rgb_dev,
                  float* T_final_dev, uint32_t* last_dev) {
    Context* c = ctx_of(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!ranges_dev || !values_dev || !conic_opacity_dev || !xy_dev || !rgb_dev || !out_rgb_dev)
        return fail(c, GSR_ERR_ARGUMENT, "gsr_rasterize: null argument");
    if (width == 0 || height == 0) return fail(c, GSR_ERR_ARGUMENT, "empty image");
    hipSetDevice(c->device);
    if (instances > 0xFFFFFFFFull) return fail(c, GSR_ERR_ARGUMENT, "too many instances");
    SplatRec* rec = (SplatRec*)scratch(c, "stage_splat_rec", (size_t)std::max<uint32_t>(count, 1) * sizeof(SplatRec));
    if (!rec) return GSR_ERR_ALLOC;
    int rc = launch_pack_splats(c, (const float4*)conic_opacity_dev, (const float2*)xy_dev, (const float4*)rgb_dev,
                                count, rec);
    if (!rc)
        rc = launch_raster(c, width, height, tile_size, count, nullptr, (uint32_t)instances, ranges_dev, values_dev,
                           rec, out_rgb_dev, T_final_dev, last_dev);
    if (rc) return rc;
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    return GSR_OK;
}

}  // extern "C"
