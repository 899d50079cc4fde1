// batch.hip — many cameras of uneven resolution in one launch sequence (the
// reference renders one view per frame, renderer.js:822-932; this is the
// same pass graph widened to a batch).
//
//   * one preprocess launch walks every camera per splat (parameters are
//     read once), writing per-(view, splat) outputs at v * N + g;
//   * one scan and one emit over all (view, splat) entries; the key is
//     (tile_base[v] + tile) << 16 | depth16, i.e. view|tile|depth, so a single
//     sort orders every view's tiles at once;
//   * one range pass over the global tile ids;
//   * a persistent raster kernel whose waves pull global tile ids from a
//     ticket counter and map them back to (view, tile);
//   * per-view cameras and tables live in device memory, filled from pinned
//     host staging by a copy that is part of the launch sequence, so the whole
//     sequence can be captured once as a hipGraph and replayed for new cameras
//     of the same shape.
#include <algorithm>
#include <cstddef>
#include <cstring>

#include "gsr_internal.h"

namespace gsr {

struct BatchPlan {
    uint32_t views = 0, tile = 16, total_tiles = 0, key_bytes = 8, end_bit = 32;
    uint32_t raster_waves = 32, sort_blocks = 2, sort_digits = 8;
    bool save = false;             // keep T_final / last contributor for gsr_backward_batch
    uint64_t capacity = 0, floats = 0;
    std::vector<uint4> info;       // {W, H, tiles_x, tile_base}
    std::vector<uint64_t> pix;     // float offset of each view's image
};

static int plan_batch(Context* c, uint32_t count, const gsr_camera* cams, uint32_t views, const gsr_options* opts,
                      BatchPlan* p) {
    if (!cams || views == 0) return fail(c, GSR_ERR_ARGUMENT, "gsr_render_batch: no cameras");
    if ((uint64_t)count * views >= 0xFFFFFFFFull)
        return fail(c, GSR_ERR_ARGUMENT, "%u splats x %u views exceed 2^32 entries", count, views);
    p->views = views;
    p->save = opts && opts->save_state;
    p->tile = (opts && opts->tile_size) ? opts->tile_size : 16;
    if (p->tile != 8 && p->tile != 16 && p->tile != 32)
        return fail(c, GSR_ERR_ARGUMENT, "tile size %u not in {8, 16, 32}", p->tile);
    // 32 persistent raster waves per CU (8 per SIMD) unless asked otherwise:
    // at 16 the queue raster of 6M splats x 8 views at 3840x2160 ran at half
    // occupancy, 6.3 ms against 5.0 ms
    p->raster_waves = (opts && opts->raster_waves_per_cu) ? opts->raster_waves_per_cu : 32;
    p->sort_blocks = (opts && opts->sort_blocks_per_cu) ? opts->sort_blocks_per_cu : 1;
    if (p->sort_blocks != 1 && p->sort_blocks != 2 && p->sort_blocks != 4)
        return fail(c, GSR_ERR_ARGUMENT, "sort blocks per CU %u not in {1, 2, 4}", p->sort_blocks);
    p->info.resize(views);
    p->pix.resize(views);
    uint64_t tiles = 0, floats = 0;
    for (uint32_t v = 0; v < views; ++v) {
        const float fw = cams[v].texture_size[0], fh = cams[v].texture_size[1];
        if (!(fw >= 1.0f && fh >= 1.0f && fw <= 65536.0f && fh <= 65536.0f))
            return fail(c, GSR_ERR_ARGUMENT, "view %u: texture size %gx%g out of range", v, fw, fh);
        const uint32_t W = (uint32_t)fw, H = (uint32_t)fh;
        const uint32_t tx = ceil_div(W, p->tile), ty = ceil_div(H, p->tile);
        p->info[v] = make_uint4(W, H, tx, (uint32_t)tiles);
        p->pix[v] = floats;
        tiles += (uint64_t)tx * ty;
        floats += (uint64_t)3 * W * H;
    }
    if (tiles >= (1ull << 31)) return fail(c, GSR_ERR_ARGUMENT, "too many tiles in the batch");
    p->total_tiles = (uint32_t)tiles;
    p->floats = floats;
    uint32_t tbits = 0;
    while ((1ull << tbits) < tiles) ++tbits;
    p->end_bit = 16 + tbits;
    const uint32_t kb = opts ? opts->key_bytes : GSR_KEY_AUTO;
    if (kb == GSR_KEY_AUTO) p->key_bytes = p->end_bit <= 32 ? 4 : 8;
    else if (kb == GSR_KEY_32 || kb == GSR_KEY_64) p->key_bytes = kb;
    else return fail(c, GSR_ERR_ARGUMENT, "key_bytes %u not in {0, 4, 8}", kb);
    if (p->key_bytes == 4 && p->end_bit > 32)
        return fail(c, GSR_ERR_ARGUMENT, "%llu tiles do not fit a 32-bit view|tile|depth key",
                    (unsigned long long)tiles);
    p->sort_digits = (opts && opts->sort_digit_bits) ? opts->sort_digit_bits
                                                      : default_digit_bits(p->key_bytes, p->end_bit);
    if (p->sort_digits < 8 || p->sort_digits > 11)
        return fail(c, GSR_ERR_ARGUMENT, "sort digit width %u not in 8..11", p->sort_digits);
    uint64_t cap = (opts && opts->instance_capacity) ? opts->instance_capacity : c->batch_capacity_hint;
    if (cap == 0) cap = std::min<uint64_t>(kMaxSortKeys, (uint64_t)16 * count * views);
    if (cap > kMaxSortKeys) {
        if (opts && opts->instance_capacity)
            return fail(c, GSR_ERR_ARGUMENT, "instance capacity %llu exceeds the sort limit %llu",
                        (unsigned long long)cap, (unsigned long long)kMaxSortKeys);
        cap = kMaxSortKeys;
    }
    p->capacity = std::max<uint64_t>(cap, 1);
    return GSR_OK;
}

// Stages the per-view tables in pinned host memory (read by the copy at the
// head of the launch sequence).
static int stage_tables(Context* c, const gsr_camera* cams, const BatchPlan& p, uint32_t count) {
    const size_t bytes = p.views * (sizeof(CamK) + sizeof(uint4) + sizeof(uint64_t));
    if (c->staging_bytes < bytes) {
        if (c->staging) {
            hipStreamSynchronize(c->stream);
            hipHostFree(c->staging);
            c->staging = nullptr;
        }
        if (hipHostMalloc(&c->staging, bytes, hipHostMallocDefault) != hipSuccess) {
            c->staging_bytes = 0;
            return fail(c, GSR_ERR_ALLOC, "pinned staging of %zu bytes failed", bytes);
        }
        c->staging_bytes = bytes;
    }
    uint8_t* s = (uint8_t*)c->staging;
    for (uint32_t v = 0; v < p.views; ++v) {
        CamK k;
        std::memcpy(&k, &cams[v], sizeof(k));
        k.count = count;
        std::memcpy(s + v * sizeof(CamK), &k, sizeof(k));
    }
    std::memcpy(s + p.views * sizeof(CamK), p.info.data(), p.views * sizeof(uint4));
    std::memcpy(s + p.views * (sizeof(CamK) + sizeof(uint4)), p.pix.data(), p.views * sizeof(uint64_t));
    return GSR_OK;
}

// Enqueues the batch frame (table copy included).  Allocation happens here
// only when buffers must grow, which never happens under graph capture
// because an eager frame of the same shape runs first.
static int enqueue_batch(Context* c, const float* params, uint32_t count, const BatchPlan& p, float* out) {
    const size_t NV = (size_t)count * p.views;
    const size_t tbytes = p.views * (sizeof(CamK) + sizeof(uint4) + sizeof(uint64_t));
    uint8_t* tables = (uint8_t*)buffer(c, "b_tables", tbytes);
    uint32_t* tiles = (uint32_t*)buffer(c, "b_tiles", NV * 4);
    uint32_t* offsets = (uint32_t*)buffer(c, "b_offsets", (NV + 1) * 4);
    uint4* meta = (uint4*)buffer(c, "b_meta", NV * 16);
    SplatRec* rec = (SplatRec*)buffer(c, "b_splat_rec", NV * sizeof(SplatRec));
    void* k0 = buffer(c, "b_keys0", p.capacity * p.key_bytes);
    void* k1 = buffer(c, "b_keys1", p.capacity * p.key_bytes);
    uint32_t* v0 = (uint32_t*)buffer(c, "b_vals0", p.capacity * 4);
    uint32_t* v1 = (uint32_t*)buffer(c, "b_vals1", p.capacity * 4);
    uint32_t* ranges = (uint32_t*)buffer(c, "b_ranges", (size_t)p.total_tiles * 8);
    FrameCounters* counters = (FrameCounters*)buffer(c, "b_counters", sizeof(FrameCounters));
    if (!tables || !tiles || !offsets || !meta || !rec || !k0 || !k1 || !v0 || !v1 || !ranges ||
        !counters)
        return GSR_ERR_ALLOC;
    float* tfinal = nullptr;
    uint32_t* last = nullptr;
    if (p.save) {
        tfinal = (float*)buffer(c, "b_tfinal", p.floats / 3 * 4);
        last = (uint32_t*)buffer(c, "b_last", p.floats / 3 * 4);
        if (!tfinal || !last) return GSR_ERR_ALLOC;
    }
    if (count == 0) {  // no splats: black views (T_final 1, no last contributor)
        GSR_HIP(c, hipMemcpyAsync(tables, c->staging, tbytes, hipMemcpyHostToDevice, c->stream));
        int rc = launch_fill_u32(c, (uint32_t*)counters, 0u, sizeof(FrameCounters) / 4);
        if (!rc) rc = launch_fill_u32(c, (uint32_t*)out, 0u, p.floats);
        if (!rc && p.save) rc = launch_fill_u32(c, (uint32_t*)tfinal, 0x3F800000u, p.floats / 3);
        if (!rc && p.save) rc = launch_fill_u32(c, last, 0u, p.floats / 3);
        return rc;
    }
    c->scratch_prefix = "b_";
    struct Restore {
        Context* c;
        ~Restore() { c->scratch_prefix = ""; }
    } restore{c};
    const CamK* cams_dev = (const CamK*)tables;
    const uint4* info_dev = (const uint4*)(tables + p.views * sizeof(CamK));
    const uint64_t* pix_dev = (const uint64_t*)(tables + p.views * (sizeof(CamK) + sizeof(uint4)));
    GSR_HIP(c, hipMemcpyAsync(tables, c->staging, tbytes, hipMemcpyHostToDevice, c->stream));
    int rc;
    uint32_t* scan_status = scan_status_buffer(c, (uint32_t)NV);
    if (!scan_status) return GSR_ERR_ALLOC;
    // one launch clears the counters, the scan's look-back words and the tile ranges
    const FillRegion clear[3] = {{(uint32_t*)counters, sizeof(FrameCounters) / 4},
                                 {scan_status, scan_status_words((uint32_t)NV)},
                                 {ranges, (size_t)p.total_tiles * 2}};
    if ((rc = launch_zero_regions(c, clear, 3))) return rc;
    mark(c, "preprocess");
    if ((rc = launch_preprocess_batch(c, params, count, cams_dev, p.views, p.tile, tiles, meta, rec)))
        return rc;
    mark(c, "offset scan");
    if ((rc = launch_scan(c, tiles, (uint32_t)NV, offsets, counters, p.capacity, true))) return rc;
    mark(c, "emit");
    if ((rc = launch_emit_batch(c, offsets, meta, count, p.views, info_dev, p.capacity, p.key_bytes, k0, v0)))
        return rc;
    mark(c, "sort");
    int idx = 0;
    if ((rc = launch_sort(c, p.key_bytes, k0, v0, k1, v1, &counters->n, 0, p.capacity, 0, p.end_bit,
                          p.sort_blocks, counters, &idx, p.sort_digits)))
        return rc;
    mark(c, "tile ranges");
    void* ks = idx ? k1 : k0;
    uint32_t* vs = idx ? v1 : v0;
    if ((rc = launch_ranges(c, ks, p.key_bytes, &counters->n, 0, p.total_tiles, ranges, false))) return rc;
    mark(c, "raster");
    if ((rc = launch_raster_queue(c, p.tile, info_dev, pix_dev, p.views, p.total_tiles, (uint32_t)NV, &counters->n,
                                  &counters->raster_ticket, ranges, vs, rec, out, p.raster_waves, tfinal, last)))
        return rc;
    mark_end(c);
    return GSR_OK;
}

static bool same_shape(const BatchGraph& g, const BatchPlan& p, uint32_t count, const float* params, const float* out) {
    if (!g.exec || g.count != count || g.params != params || g.out != out || g.views != p.views || g.tile != p.tile ||
        g.key_bytes != p.key_bytes || g.capacity != p.capacity || g.raster_waves != p.raster_waves ||
        g.sort_blocks != p.sort_blocks || g.sort_digits != p.sort_digits || g.save != p.save ||
        g.dims.size() != p.views)
        return false;
    for (uint32_t v = 0; v < p.views; ++v)
        if (g.dims[v] != ((uint64_t)p.info[v].x << 32 | p.info[v].y)) return false;
    return true;
}

static void drop_graph(Context* c) {
    if (c->graph.exec) hipGraphExecDestroy(c->graph.exec);
    c->graph = BatchGraph();
}

static int capture_batch(Context* c, const float* params, uint32_t count, const BatchPlan& p, float* out) {
    drop_graph(c);
    const uint32_t timing = c->timing_on;
    c->timing_on = 0;
    hipGraph_t graph = nullptr;
    GSR_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch(c, params, count, p, out);
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    c->timing_on = timing;
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(c, GSR_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(c, GSR_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    BatchGraph& g = c->graph;
    g.exec = exec;
    g.count = count;
    g.params = params;
    g.out = out;
    g.views = p.views;
    g.tile = p.tile;
    g.key_bytes = p.key_bytes;
    g.capacity = p.capacity;
    g.raster_waves = p.raster_waves;
    g.sort_blocks = p.sort_blocks;
    g.sort_digits = p.sort_digits;
    g.save = p.save;
    g.dims.resize(p.views);
    for (uint32_t v = 0; v < p.views; ++v) g.dims[v] = (uint64_t)p.info[v].x << 32 | p.info[v].y;
    return GSR_OK;
}

static void save_batch_state(Context* c, const BatchPlan& p, uint32_t count, const gsr_camera* cams) {
    BatchState& s = c->bstate;
    s.valid = true;
    s.count = count;
    s.views = p.views;
    s.tile = p.tile;
    s.total_tiles = p.total_tiles;
    s.sorted_index = (int)sort_result_index(0, p.end_bit, p.sort_digits);
    s.cams.resize(p.views);
    for (uint32_t v = 0; v < p.views; ++v) {
        std::memcpy(&s.cams[v], &cams[v], sizeof(CamK));
        s.cams[v].count = count;
    }
}

// Runs one batch frame (graph replay when the shape matches) and checks it;
// on instance overflow the capacity is raised and the frame re-run once.
static int run_batch(Context* c, const float* params, uint32_t count, const gsr_camera* cams, uint32_t views,
                     const gsr_options* opts, int use_graph, float* out, uint64_t* out_offsets) {
    if (out) c->bstate.valid = false;  // the batch buffers are about to be overwritten
    for (int attempt = 0; attempt < 2; ++attempt) {
        BatchPlan p;
        int rc = plan_batch(c, count, cams, views, opts, &p);
        if (rc) return rc;
        if (out_offsets) {
            for (uint32_t v = 0; v < views; ++v) out_offsets[v] = p.pix[v];
            out_offsets[views] = p.floats;
        }
        if (!out) return GSR_OK;  // layout query only
        if ((rc = stage_tables(c, cams, p, count))) return rc;
        const bool replay = use_graph && same_shape(c->graph, p, count, params, out);
        if (replay) {
            mark(c, "batch graph");
            GSR_HIP(c, hipGraphLaunch(c->graph.exec, c->stream));
            mark_end(c);
        } else {
            if ((rc = enqueue_batch(c, params, count, p, out))) return rc;
        }
        FrameCounters* counters = (FrameCounters*)buffer(c, "b_counters", sizeof(FrameCounters));
        GSR_HIP(c, hipMemcpyAsync(c->host_counters, counters, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream));
        GSR_HIP(c, hipStreamSynchronize(c->stream));
        // with saved state the backward's stages join this frame's rows
        // (collected by gsr_get_timings)
        if (!p.save) finish_frame_timings(c);
        if (c->host_counters->error)
            return fail(c, GSR_ERR_KERNEL, "a look-back wait exceeded its bound (code %u)", c->host_counters->error);
        const uint64_t inst = c->host_counters->instances;
        c->batch_last_instances = inst;
        c->batch_last_capacity = p.capacity;
        if (inst <= p.capacity) {
            if (use_graph && !replay) {
                if ((rc = capture_batch(c, params, count, p, out))) return rc;
            }
            if (p.save) save_batch_state(c, p, count, cams);
            return GSR_OK;
        }
        // overflow: grow (unless the caller fixed the capacity) and re-run
        if ((opts && opts->instance_capacity) || attempt == 1 || inst > kMaxSortKeys)
            return fail(c, GSR_ERR_CAPACITY, "%llu tile instances exceed the instance capacity %llu",
                        (unsigned long long)inst, (unsigned long long)p.capacity);
        c->batch_capacity_hint = std::min<uint64_t>(kMaxSortKeys, inst + inst / 4);
        drop_graph(c);
    }
    return GSR_OK;
}

void destroy_batch_graph(Context* c) { drop_graph(c); }

}  // namespace gsr

using namespace gsr;

extern "C" {

int gsr_render_batch(gsr_context* ctx, const float* params_dev, uint32_t count, const gsr_camera* cams, uint32_t views,
                     const gsr_options* opts, int use_graph, float* out_rgb_dev, uint64_t* out_offsets) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!params_dev && count > 0) return fail(c, GSR_ERR_ARGUMENT, "gsr_render_batch: no parameters");
    hipSetDevice(c->device);
    c->timing_on = opts ? opts->timing : 0;
    return run_batch(c, params_dev, count, cams, views, opts, use_graph, out_rgb_dev, out_offsets);
}

int gsr_batch_instance_count(gsr_context* ctx, uint64_t* instances) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!instances) return fail(c, GSR_ERR_ARGUMENT, "null output");
    *instances = c->batch_last_instances;
    return GSR_OK;
}

int gsr_batch_capacity(gsr_context* ctx, uint64_t* slots) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!slots) return fail(c, GSR_ERR_ARGUMENT, "null output");
    *slots = c->batch_last_capacity;
    return GSR_OK;
}

int gsr_autotune_batch(gsr_context* ctx, const float* params_dev, uint32_t count, const gsr_camera* cams,
                       uint32_t views, gsr_options* inout, float* out_rgb_dev, double* best_ms) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!inout || !out_rgb_dev) return fail(c, GSR_ERR_ARGUMENT, "gsr_autotune_batch: null argument");
    hipSetDevice(c->device);
    c->bstate.valid = false;
    const uint32_t tiles[3] = {8, 16, 32};
    const uint32_t waves[3] = {8, 16, 32};
    const uint32_t blocks[3] = {1, 2, 4};
    const uint32_t digits[4] = {8, 9, 10, 11};
    hipEvent_t e0, e1;
    GSR_HIP(c, hipEventCreate(&e0));
    GSR_HIP(c, hipEventCreate(&e1));
    double best = 1e30;
    gsr_options best_opts = *inout;
    const uint32_t timing = c->timing_on;
    c->timing_on = 0;
    int rc = GSR_OK;
    for (uint32_t ti = 0; ti < 3 && rc == GSR_OK; ++ti)
        for (uint32_t wi = 0; wi < 3 && rc == GSR_OK; ++wi)
            for (uint32_t bd = 0; bd < 12 && rc == GSR_OK; ++bd) {
                gsr_options o = *inout;
                o.tile_size = tiles[ti];
                o.raster_waves_per_cu = waves[wi];
                o.sort_blocks_per_cu = blocks[bd % 3];
                o.sort_digit_bits = digits[bd / 3];
                o.timing = 0;
                // a tile size whose view|tile|depth key does not fit the requested key width is skipped
                BatchPlan p;
                if (plan_batch(c, count, cams, views, &o, &p) != GSR_OK) continue;
                if ((rc = run_batch(c, params_dev, count, cams, views, &o, 0, out_rgb_dev, nullptr))) break;
                float ms_sum = 0.0f;
                const int reps = 3;
                for (int r = 0; r < reps && rc == GSR_OK; ++r) {
                    if ((rc = stage_tables(c, cams, p, count))) break;
                    hipEventRecord(e0, c->stream);
                    if ((rc = enqueue_batch(c, params_dev, count, p, out_rgb_dev))) break;
                    hipEventRecord(e1, c->stream);
                    hipEventSynchronize(e1);
                    float ms = 0.0f;
                    hipEventElapsedTime(&ms, e0, e1);
                    ms_sum += ms;
                }
                const double avg = ms_sum / reps;
                if (rc == GSR_OK && avg < best) {
                    best = avg;
                    best_opts = o;
                }
            }
    c->timing_on = timing;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    best_opts.timing = inout->timing;
    *inout = best_opts;
    if (best_ms) *best_ms = best;
    return GSR_OK;
}

int gsr_backward_batch(gsr_context* ctx, const float* params_dev, uint32_t count, const gsr_camera* cams,
                       uint32_t views, const float* dL_drgb_dev, float* grad_params_dev, float* grad2d_dev) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!cams || ((!params_dev || !dL_drgb_dev || !grad_params_dev) && count > 0))
        return fail(c, GSR_ERR_ARGUMENT, "gsr_backward_batch: null argument");
    const BatchState& s = c->bstate;
    if (!s.valid || s.count != count || s.views != views)
        return fail(c, GSR_ERR_STATE, "gsr_backward_batch needs the matching gsr_render_batch with save_state = 1 first");
    for (uint32_t v = 0; v < views; ++v) {
        CamK k;
        std::memcpy(&k, &cams[v], sizeof(k));
        k.count = count;
        if (std::memcmp(&k, &s.cams[v], offsetof(CamK, pad)) != 0)
            return fail(c, GSR_ERR_STATE, "view %u: camera differs from the saved batch forward", v);
    }
    hipSetDevice(c->device);
    uint8_t* tables = (uint8_t*)buffer(c, "b_tables", 0);
    if (!tables) return GSR_ERR_ALLOC;
    const CamK* cams_dev = (const CamK*)tables;
    const uint4* info_dev = (const uint4*)(tables + views * sizeof(CamK));
    const uint64_t* pix_dev = (const uint64_t*)(tables + views * (sizeof(CamK) + sizeof(uint4)));
    return launch_backward_batch(c, params_dev, count, views, s.tile, s.total_tiles, cams_dev, info_dev, pix_dev,
                                 s.sorted_index, dL_drgb_dev, grad_params_dev, grad2d_dev);
}

}  // extern "C"
