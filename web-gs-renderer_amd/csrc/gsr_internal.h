// gsr_internal.h — shared definitions of the gfx950 renderer library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/gsr.h"

namespace gsr {

// ---- device-side constants of the reference pipeline --------------------
constexpr float kNearView = 0.2f;        // preprocess.wgsl:34
constexpr float kFrustumGuard = 1.3f;    // preprocess.wgsl:35
constexpr float kSmall = 1e-7f;          // preprocess.wgsl:32
constexpr float kAlphaClamp = 0.99f;     // raseter.wgsl:25
constexpr float kAlphaMin = 1.0f / 255.0f;  // raseter.wgsl:26
constexpr float kTransmitMin = 0.0001f;  // raseter.wgsl:27
constexpr uint32_t kParams = 59;         // floats per splat (loader.js:4-11)
constexpr uint32_t kInstanceFactor = 64; // renderer.js:16

// Planar channel offsets of the [59][N] device layout.
constexpr int CH_POS = 0, CH_SCALE = 3, CH_ROT = 6, CH_DC = 10, CH_OPACITY = 13, CH_REST = 14;

// exp for FP32, the same operation sequence as the oracle's exp_f32:
// Cody-Waite reduction x = n ln2 + r, degree-7 Taylor polynomial of e^r in
// Horner form with fused multiply-adds, scaled by 2^n (v_ldexp_f32).
// exp_core is valid for x in [-87, 88.7]; gsr_exp adds the range handling.
__device__ __forceinline__ float exp_core(float x) {
    const float n = __builtin_rintf(x * 0x1.715476p+0f);
    float r = __builtin_fmaf(n, -0x1.62e430p-1f, x);
    r = __builtin_fmaf(n, 0x1.05c610p-29f, r);
    float p = 0x1.a01a02p-13f;
    p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
    p = __builtin_fmaf(p, r, 0x1.111112p-7f);
    p = __builtin_fmaf(p, r, 0x1.555556p-5f);
    p = __builtin_fmaf(p, r, 0x1.555556p-3f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_ldexpf(p, (int)n);
}

// e^p for the alpha of a splat at a pixel, p = the power in [-5.55, 0] (the
// range where alpha can reach 1/255; results outside it are never used):
// 2^(p log2 e) with the integer part k rounded off against 1.5 * 2^23, the
// fraction f = p log2 e - k (|f| <= 1/2) from one fused multiply-add, 2^f by a
// degree-5 polynomial fitted for relative error on [-1/2, 1/2], scaled by 2^k.
// Within 2.7 ulp of e^p for every float in [-5.55, 0].  The oracle's
// exp_blend_f32 is the same sequence.  The scaling adds k to the exponent
// field in one instruction: t's bits are those of 1.5 * 2^23 plus k, and
// 1.5 * 2^23's bits shifted left by 23 wrap to 0, so (bits(t) << 23) is
// k << 23.  On [-5.55, 0] (k in [-9, 0], 2^f in [0.7, 1.42]) the result is a
// normal float and equals ldexp(2^f, k) bit for bit; elsewhere it may not,
// and those lanes' values are never used (a NaN power stays NaN).
__device__ __forceinline__ float exp_blend(float p) {
    constexpr float kLog2e = 0x1.715476p+0f, kRound = 0x1.8p+23f;
    const float t = __builtin_fmaf(p, kLog2e, kRound);
    const float f = __builtin_fmaf(p, kLog2e, -(t - kRound));
    float q = 0x1.5bba1p-10f;
    q = __builtin_fmaf(q, f, 0x1.3cea88p-7f);
    q = __builtin_fmaf(q, f, 0x1.c6b752p-5f);
    q = __builtin_fmaf(q, f, 0x1.ebf9bcp-3f);
    q = __builtin_fmaf(q, f, 0x1.62e42ap-1f);
    q = __builtin_fmaf(q, f, 1.0f);
    static_assert((0x4B400000u << 23) == 0u, "the rounding constant's bits vanish in the shift");
    return __uint_as_float(__float_as_uint(q) + (__float_as_uint(t) << 23));
}

__device__ __forceinline__ float gsr_exp(float x) {
    if (x != x) return x;
    if (x > 88.8f) return __builtin_inff();
    if (x < -87.0f) return 0.0f;
    return exp_core(x);
}

// What the raster kernels gather per splat: 48 bytes of data padded to a
// 64-byte aligned record, so that one instance reads one 64-byte sector
// (a 48-byte stride put every other record across two sectors: the raster
// fetched 39% more per frame) instead of three separate arrays' lines.
//   co = (A, B, C, opacity): conic and opacity
//   pq = (pixel x, pixel y, q_max, 0): q_max the widened reach threshold of
//        cull.h, computed once per splat
//   cl = (r, g, b, 0)
//   pad = 0 (left unwritten in view batches; no kernel reads it)
// The record of a splat that touches no tile is never read.
struct alignas(64) SplatRec {
    float4 co, pq, cl, pad;
};
static_assert(sizeof(SplatRec) == 64, "splat record must be 64 bytes");

// Camera block as the kernels read it (same bytes as gsr_camera).
struct CamK {
    float view[16];
    float view_proj[16];
    float eye[3];
    uint32_t count;
    float focal[2];
    float tan_fov[2];
    float texture_size[2];
    uint32_t pad[2];
};
static_assert(sizeof(CamK) == 176, "camera block must be 176 bytes");
static_assert(sizeof(gsr_camera) == 176, "gsr_camera must be 176 bytes");

// ---- sort tuning ---------------------------------------------------------
constexpr int kMaxPasses = 8;                           // 64-bit keys at 8-bit digits
constexpr uint32_t kStatusValueMask = 0x3FFFFFFFu;    // 30-bit counts in look-back words
constexpr uint64_t kMaxSortKeys = kStatusValueMask;   // sort capacity limit

// Device-resident per-frame counters (one 256-byte block).
struct FrameCounters {
    uint64_t instances;       // raw scan total I
    uint32_t n;               // min(I, capacity)
    uint32_t error;           // nonzero: a kernel hit a bounded-wait limit
    uint32_t raster_ticket;   // next (view, tile) work item of the batch raster queue
    uint32_t pad[64 - 5];
};

// ---- host-side context -----------------------------------------------------
struct Scene {
    float* params = nullptr;   // planar [59][N]
    uint32_t count = 0;
};

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct TimingAcc {
    std::vector<std::string> order;
    std::map<std::string, double> sums;
    int frames = 0;
};

// A captured batch frame and the shape it was captured for.
struct BatchGraph {
    hipGraphExec_t exec = nullptr;
    uint32_t count = 0, views = 0, tile = 0, key_bytes = 0, raster_waves = 0, sort_blocks = 0, sort_digits = 0;
    uint64_t capacity = 0;
    bool save = false;
    const float* params = nullptr;
    const float* out = nullptr;
    std::vector<uint64_t> dims;  // (W << 32) | H per view
};

// What a batch frame with save_state left for gsr_backward_batch.
struct BatchState {
    bool valid = false;
    uint32_t count = 0, views = 0, tile = 16, total_tiles = 0;
    int sorted_index = 0;          // which of the b_vals buffers holds the sorted entries
    std::vector<CamK> cams;        // the frame's cameras (device copy at the head of b_tables)
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    std::map<std::string, Scene> scenes;
    std::string current;
    std::map<std::string, DevBuf> bufs;   // grow-only named scratch
    int cu_count = 256;
    // timing
    std::vector<std::pair<std::string, hipEvent_t>> marks;  // label, start event (end = next start)
    std::vector<hipEvent_t> event_pool;
    size_t event_used = 0;
    TimingAcc timing;
    std::vector<gsr_timing_row> last_rows;
    int last_frames = 0;
    uint32_t timing_on = 0;  // gsr_options.timing of the current call
    // saved forward state for backward: the forward's camera (count filled
    // in) and parameters, which gsr_backward must be called with
    bool have_state = false;
    uint32_t state_count = 0, state_w = 0, state_h = 0, state_tile = 16, state_key_bytes = 4;
    uint64_t state_capacity = 0;
    int state_sorted_index = 0;
    CamK state_cam{};
    const float* state_params = nullptr;
    // a gsr_forward whose counters (instance overflow, look-back failure) no
    // host call has read yet; gsr_synchronize reads them
    bool frame_unchecked = false;
    // entries of b_grad2d that the last batched backward left zero (it
    // clears the sums it has read; 0: unknown, fill before use)
    size_t b_grad2d_clean = 0;
    size_t grad2d_clean = 0;     // the same for the plain backward's sums
    uint32_t last_frame_w = 0, last_frame_h = 0;  // size of the frame in "image"
    // prefix of the scan/sort scratch buffer names ("" plain frames, "b_" batches)
    const char* scratch_prefix = "";
    // batches
    void* staging = nullptr;   // pinned per-view tables
    size_t staging_bytes = 0;
    uint64_t batch_capacity_hint = 0, batch_last_instances = 0, batch_last_capacity = 0;
    BatchGraph graph;
    BatchState bstate;
    // host pinned staging for small readbacks
    FrameCounters* host_counters = nullptr;
    // device address of this device's fixed no-allocation scratch (null when
    // its symbol lookup failed; noalloc_message says why)
    uint8_t* noalloc_scratch = nullptr;
    std::string noalloc_message;
};

// Looks up (once per device) the device address of the fixed scratch array
// that gsr_sort_pairs_noalloc uses, with c's device current.
void noalloc_attach(Context* c);

// Reads the frame counters back (stream synced) and reports overflow of
// `capacity` (GSR_ERR_CAPACITY) or a failed bounded wait (GSR_ERR_KERNEL); a
// failed frame's saved state is dropped.
int check_frame(Context* c, uint64_t capacity);

// Named device buffer, grown to at least `bytes` (contents not preserved).
// Growing a buffer drops any captured batch graph (it may point at it).
void* buffer(Context* c, const char* name, size_t bytes);
// Scratch buffer named scratch_prefix + name.
void* scratch(Context* c, const char* name, size_t bytes);
void destroy_batch_graph(Context* c);

int fail(Context* c, int code, const char* fmt, ...);
#define GSR_HIP(ctx, call)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::gsr::fail((ctx), GSR_ERR_DEVICE, "%s failed: %s", #call,              \
                               hipGetErrorString(e_));                                     \
    } while (0)

// Timing marks on the context stream (no-ops unless timing is on).
void mark(Context* c, const char* label);
// Closes the last stage; frame_end = false for a stage group (e.g. the
// backward) that belongs to the same frame as the preceding forward.
void mark_end(Context* c, bool frame_end = true);
// After the stream has drained: turn the recorded frames' marks into timing
// rows (per-frame averages over every frame since the last collection).
void finish_frame_timings(Context* c);

// ---- kernel launchers (each enqueues on c->stream) -------------------------
// A device region of u32 words (zeroed by launch_zero_regions and friends).
struct FillRegion {
    uint32_t* p;
    size_t count;
};
// Preprocess writes either the three separate arrays (conic, xy, rgb; the
// stage entry point) or, when rec is non-null, the packed records.
// clear (nullable): regions the preprocess grid zeroes on the side.
int launch_preprocess(Context* c, const float* params, uint32_t n, const CamK& cam, uint32_t tile,
                      uint32_t* tiles, uint4* meta, float4* conic, float2* xy, float4* rgb, SplatRec* rec,
                      const FillRegion* clear = nullptr, int nclear = 0);
int launch_preprocess_batch(Context* c, const float* params, uint32_t n, const CamK* cams, uint32_t views,
                            uint32_t tile, uint32_t* tiles, uint4* meta, SplatRec* rec);
// Packs separate (conic, xy, rgb) arrays into records.
int launch_pack_splats(Context* c, const float4* conic, const float2* xy, const float4* rgb, uint32_t n,
                       SplatRec* rec);
int launch_emit_batch(Context* c, const uint32_t* offsets, const uint4* meta, uint32_t n, uint32_t num_views,
                      const uint4* views, uint64_t capacity, uint32_t key_bytes, void* keys, uint32_t* vals);
// Scan of n counts; its look-back status buffer (scan_status_buffer, of
// scan_status_words(n) words) must be zero, or is cleared first when
// status_cleared is false.
int launch_scan(Context* c, const uint32_t* in, uint32_t n, uint32_t* out, FrameCounters* counters,
                uint64_t capacity, bool status_cleared = false);
size_t scan_status_words(uint32_t n);
uint32_t* scan_status_buffer(Context* c, uint32_t n);
int launch_emit(Context* c, const uint32_t* offsets, const uint4* meta, uint32_t n, uint32_t tiles_x,
                uint64_t capacity, uint32_t key_bytes, void* keys, uint32_t* vals);
// Sort of n (device counter n_dev if non-null, else n_host) pairs with
// onesweep blocks_per_cu in {1, 2, 4} and digit_bits in 8..11; returns the
// result index (which ping-pong buffer holds the output).  hist_cleared: an
// earlier kernel of the stream zeroed sort_hist_region (for the same bit
// range and digit width), so the sort skips its own clearing launch.
// mem (nullable): caller-provided scratch for the histograms and look-back
// words instead of the context's buffers (no allocation); it must hold
// sort_scratch_words() of each kind (one histogram copy; the context's
// buffers hold several, see sort.hip).
struct SortMem {
    uint32_t* hist;
    size_t hist_words;
    uint32_t* status;
    size_t status_words;
};
struct SortWords {
    size_t hist, status;
};
int launch_sort(Context* c, uint32_t key_bytes, void* k0, uint32_t* v0, void* k1, uint32_t* v1,
                const uint32_t* n_dev, uint64_t n_host, uint64_t capacity, uint32_t begin_bit,
                uint32_t end_bit, uint32_t blocks_per_cu, FrameCounters* counters, int* result_index,
                uint32_t digit_bits = 8, bool hist_cleared = false, const SortMem* mem = nullptr);
// Scratch words launch_sort needs in caller-provided memory (mem) for
// `capacity` keys with 1 block per CU and 8-bit digits (other geometries:
// SIZE_MAX).
SortWords sort_scratch_words(const Context* c, uint32_t key_bytes, uint64_t capacity, uint32_t begin_bit,
                             uint32_t end_bit, uint32_t blocks_per_cu, uint32_t digit_bits);
// Keys per full onesweep tile of that geometry.
uint32_t sort_default_tile_keys(uint32_t key_bytes);
// The sort's histogram and ticket words for this bit range and digit width
// (count 0 when the buffer cannot be allocated).
FillRegion sort_hist_region(Context* c, uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits);
// The result index launch_sort reports for this bit range and digit width.
uint32_t sort_result_index(uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits);
// Digit width used when the options leave it open: 8 bits for 32-bit keys;
// for 64-bit keys the narrowest of 8..10 bits with the fewest passes over
// [0, end_bit) (a 33- or 34-bit view|tile|depth key: 9 bits, 4 passes).
uint32_t default_digit_bits(uint32_t key_bytes, uint32_t end_bit);
int launch_ranges(Context* c, const void* keys, uint32_t key_bytes, const uint32_t* n_dev,
                  uint64_t n_host, uint32_t num_tiles, uint32_t* ranges, bool clear);
// splats bounds the per-splat records, the instance count (n_dev, else n_host) the ranges.
// order: tiles heaviest first (launch_tile_order) or null to compute it.
int launch_raster(Context* c, uint32_t W, uint32_t H, uint32_t tile, uint32_t splats, const uint32_t* n_dev,
                  uint32_t n_host, const uint32_t* ranges, const uint32_t* vals, const SplatRec* rec, float* out,
                  float* tfinal, uint32_t* last, const uint32_t* order = nullptr);
int launch_fill_u32(Context* c, uint32_t* p, uint32_t value, size_t count);
// order[0..tiles) = the tiles, heaviest (longest instance range) first; with
// views (device table, views[v].w = first tile of view v) heaviest first
// within each view, views in turn.
int launch_tile_order(Context* c, const uint2* ranges, uint32_t tiles, uint32_t* order,
                      const uint4* views = nullptr, uint32_t num_views = 0);
// Zeroes up to 4 regions (pointer, u32 count) with one launch.
int launch_zero_regions(Context* c, const FillRegion* regions, int count);

// Up to 4 regions to zero, as a kernel argument: a kernel that runs anyway
// clears them on the side (zero_regions_part) instead of a launch of its own.
struct ZeroRegions {
    uint32_t* p[4];
    size_t end[4];  // running end of region r in the concatenated index space
    int count;
};
inline ZeroRegions make_zero_regions(const FillRegion* regions, int count) {
    ZeroRegions z{};
    size_t total = 0;
    z.count = count;
    for (int r = 0; r < count && r < 4; ++r) {
        z.p[r] = regions[r].p;
        total += regions[r].count;
        z.end[r] = total;
    }
    return z;
}
// Thread t of a grid of `threads` zeroes words t, t + threads, ...
__device__ __forceinline__ void zero_regions_part(const ZeroRegions& z, size_t t, size_t threads) {
    if (z.count <= 0) return;
    const size_t total = z.end[z.count - 1];
    for (size_t i = t; i < total; i += threads) {
        int r = 0;
        while (i >= z.end[r]) ++r;
        z.p[r][i - (r ? z.end[r - 1] : 0)] = 0u;
    }
}
// Batch raster: views[v] = {W, H, tiles_x, tile_base}, pix[v] = float offset of view v's image.
// tfinal / last (both or neither): per-pixel state, view v's at pix[v] / 3.
int launch_raster_queue(Context* c, uint32_t tile, const uint4* views, const uint64_t* pix, uint32_t num_views,
                        uint32_t total_tiles, uint32_t splats, const uint32_t* n_dev, uint32_t* ticket,
                        const uint32_t* ranges, const uint32_t* vals, const SplatRec* rec, float* out,
                        uint32_t waves_per_cu, float* tfinal = nullptr, uint32_t* last = nullptr);
// View batches after a batch forward with saved state: entry v * count + g
// of the batch buffers ("b_" prefix); dLdC in the batch output layout;
// gradients summed over views into grad; grad2d_out (nullable) gets the 2-D
// gradient of every entry.
int launch_backward_batch(Context* c, const float* params, uint32_t count, uint32_t views, uint32_t tile,
                          uint32_t total_tiles, const CamK* cams, const uint4* info, const uint64_t* pix,
                          int sorted_index, const float* dLdC, float* grad, float* grad2d_out);

__host__ __device__ inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace gsr
