// raster.hip — tile ranges (src/shaders/tile_ranges.wgsl:7-35) and the
// front-to-back alpha-blend rasterizer (src/shaders/raseter.wgsl:35-117).
//
// Raster layout: one wave64 per 8x8 pixel block of a tile (1, 4 or 16 waves
// for 8, 16, 32-pixel tiles), one pixel per lane; every wave of a tile walks
// the tile's whole instance range, so each pixel sees the reference's
// per-tile order.  Each batch of 64 instances is gathered one per lane
// (value -> conic/opacity, pixel, colour), staged in the wave's own LDS slots
// and read back per splat with same-address (broadcast) loads, which keeps
// the broadcast off the VALU; the batch loop stops once every pixel of the
// block is saturated, as the reference's per-batch vote does per tile.
// Per pixel the FP32 steps are the oracle's:
//   power = fma(-0.5, fma(A dx, dx, (C dy) dy), -(B dx) dy); skip if > 0;
//   alpha = min(0.99, opacity * exp(power)); skip if < 1/255;
//   T' = T (1 - alpha); stop before blending when T' < 1e-4;
//   rgb += colour * (alpha T); T = T'.
// A power below -5.55 is skipped too: exp(-5.55) < 1/255 and the opacity is
// at most 1, so the reference's alpha test would reject it as well.
// Per batch each lane also tests its splat against the 8x8 block (cull.h):
// splats that cannot reach any pixel of the block are skipped without
// changing any pixel's result.
#include "cull.h"
#include "gsr_internal.h"
#include "wave.h"

namespace gsr {
namespace {

// Blocks take kRangeChunk-key chunks in a grid-stride loop; each wave loads
// kRangeItems rows of 64 consecutive keys (coalesced, all in flight), and a key's
// predecessor comes from the neighbouring lane (DPP wave_shr:1), from lane 63
// of the previous row, or, for the wave's first key, from memory.
constexpr int kRangeItems = 8;  // 8 rows and 8 blocks per CU: 2.7 us less than 16 and 4 at 1080p
constexpr uint32_t kRangeChunk = 4 * 64 * kRangeItems;

template <typename K>
__global__ __launch_bounds__(256) void k_ranges(const K* __restrict__ keys, const uint32_t* n_dev,
                                                uint64_t n_host, uint32_t num_tiles,
                                                uint32_t* __restrict__ ranges) {
    constexpr int ITEMS = kRangeItems;
    const uint32_t n = n_dev ? *n_dev : (uint32_t)n_host;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t chunk = blockIdx.x; chunk * kRangeChunk < n; chunk += gridDim.x) {
        const uint64_t wbase = chunk * kRangeChunk + wave * (64 * ITEMS);
        uint32_t tile[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint64_t p = wbase + i * 64 + lane;
            tile[i] = p < n ? (uint32_t)(keys[p] >> 16) : 0u;
        }
        const uint32_t before = (wbase > 0 && wbase - 1 < n) ? (uint32_t)(keys[wbase - 1] >> 16) : 0u;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint64_t p = wbase + i * 64 + lane;
            uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)tile[i], 0x138, 0xF, 0xF, false);
            if (lane == 0) prev = i == 0 ? before : (uint32_t)__builtin_amdgcn_readlane((int)tile[i - 1], 63);
            const uint32_t t = tile[i];
            if (p >= n || t >= num_tiles) continue;
            if (p == 0) {
                ranges[2 * t] = 0;
            } else if (t != prev) {
                if (prev < num_tiles) ranges[2 * prev + 1] = (uint32_t)p;
                ranges[2 * t] = (uint32_t)p;
            }
            if (p == n - 1) ranges[2 * t + 1] = n;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t count) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

__global__ __launch_bounds__(256) void k_zero_regions(ZeroRegions z) {
    zero_regions_part(z, (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

// One gathered splat as staged in LDS: 48 bytes, read back with three
// same-address loads (the LDS broadcasts them to the wave).
struct SplatL {
    float4 co;  // -A/2, -B, -C/2 of the conic, and opacity
    float4 cl;  // colour (w unused)
    float2 p;   // pixel position
    float2 pad;
};

// One 8x8 block of one tile of one view, run by one wave64, one pixel per
// lane: lane l owns pixel (x0 + l % 8, y0 + l / 8).  start/end delimit the
// tile's instances; out/tfinal/last address the view's image ([3][H][W]).
// The caller's workgroup has 4 waves (wave w stages its batch in s_spl[w]).
// SAVE: also write T_final and the last contributor (tfinal/last non-null).
template <bool SAVE>
__device__ __forceinline__ void raster_block(uint32_t W, uint32_t H, uint32_t x0, uint32_t y0, uint32_t start,
                                            uint32_t end, uint32_t splats, const uint32_t* __restrict__ vals,
                                            const SplatRec* __restrict__ rec, float* __restrict__ out,
                                            float* __restrict__ tfinal, uint32_t* __restrict__ last) {
    __shared__ SplatL s_spl[4][64];
    const uint32_t lane = threadIdx.x & 63;
    // the wave's slot row as a wave-uniform (scalar) index, so that a
    // record's LDS address is scalar arithmetic
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t x = x0 + lane % 8, y = y0 + lane / 8;
    const bool inside = x < W && y < H;
    bool done = !inside;
    const float fx = (float)x, fy = (float)y;
    float T = 1.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    uint32_t lst = 0;
    const float rxa = (float)x0, rxb = (float)min(x0 + 7, W - 1);
    const float rya = (float)y0, ryb = (float)min(y0 + 7, H - 1);
    // Lanes past the range or holding a value that names no splat read a
    // valid record anyway (the range's last value, splat 0) and are not
    // live: no zero-filled registers and no branch around the loads.  With
    // no splats there is no record to read and no instance can be valid.
    const uint32_t bend = splats != 0 ? end : start;
    for (uint32_t b = start; b < bend; b += 64) {
        if (ballot(!done) == 0) break;
        const uint32_t idx = b + lane;
        const uint32_t g = vals[min(idx, end - 1)];
        const bool valid = idx < end && g < splats;
        const SplatRec sr = rec[valid ? g : 0u];
        const float4 co = sr.co, cl = sr.cl;
        const float2 p = make_float2(sr.pq.x, sr.pq.y);
        const bool live = valid && reaches_rect(p.x, p.y, co.x, co.y, co.z, sr.pq.z, rxa, rxb, rya, ryb);
        // Stage the batch in this wave's LDS slots.  Only this wave touches
        // them: a wave-level barrier orders the previous batch's reads before
        // these writes and these writes before the reads below.
        // The conic is staged as (-A/2, -B, -C/2) (exact scalings), so that
        // the power is dx (a dx + b dy) + c dy^2: two FMAs and three products
        __builtin_amdgcn_wave_barrier();
        s_spl[wv][lane].co = make_float4(-0.5f * co.x, -co.y, -0.5f * co.z, co.w);
        s_spl[wv][lane].cl = cl;
        s_spl[wv][lane].p = p;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Only splats that can reach the block, in range order.  The pixel
        // tests are evaluated as predicates and the blend update as selects,
        // so the loop body has no branch and reads the whole LDS record with
        // one wait; the tests are the skip/stop conditions above, in the same
        // order of effect.  colour += c * 0 where nothing is blended leaves
        // it unchanged bit for bit (colours are finite and >= 0).
        // 1 + the batch's position in the range, as a scalar: the last
        // contributor's index is then a scalar add per splat
        const uint32_t base = (uint32_t)__builtin_amdgcn_readfirstlane((int)(b - start + 1));
        for (uint64_t lm = ballot(live); lm != 0; lm &= lm - 1) {
            const int j = (int)__builtin_ctzll(lm);
            const SplatL s = s_spl[wv][j];
            const float dx = s.p.x - fx, dy = s.p.y - fy;
            const float power = __builtin_fmaf(dx, __builtin_fmaf(s.co.x, dx, s.co.y * dy), (s.co.z * dy) * dy);
            const float alpha = fminf(kAlphaClamp, s.co.w * exp_blend(power));
            const float tn = __builtin_fmaf(-alpha, T, T);  // T (1 - alpha) as T - alpha T
            const bool ok = !done & !(power > 0.0f) & !(power < -5.55f) & !(alpha < kAlphaMin);
            done |= ok & (tn < kTransmitMin);
            // ok lanes were not done before: blended where still not done
            // (one compare of tn, the other test a mask operation)
            const bool use = ok & !done;
            const float w = use ? alpha * T : 0.0f;
            c0 = __builtin_fmaf(s.cl.x, w, c0);
            c1 = __builtin_fmaf(s.cl.y, w, c1);
            c2 = __builtin_fmaf(s.cl.z, w, c2);
            T = use ? tn : T;
            if (SAVE) lst = use ? base + (uint32_t)j : lst;
        }
    }
    if (inside) {
        const size_t plane = (size_t)W * H;
        const size_t o = (size_t)y * W + x;
        out[o] = c0;
        out[plane + o] = c1;
        out[2 * plane + o] = c2;
        if (SAVE) {
            tfinal[o] = T;
            last[o] = lst;
        }
    }
}

// Heaviest tiles first: the raster kernels take tiles in the order written
// here, tiles bucketed by instance count on a log scale (4 classes per
// octave) and the classes in descending order, so the longest tiles start
// first and the short ones fill in behind them instead of leaving a tail.
// One workgroup per range of tiles: with views (nullable; views[v].w = first
// tile of view v) workgroup v orders view v's tiles within their own slots,
// so the order stays view by view and a pass over it touches one view's
// per-splat data at a time.
// Each class has kOrderCopies counters, lane l of a wave using copy l % 32:
// neighbouring tiles mostly share a class, and LDS atomics of a wave's lanes
// on one address are serialised, which with a single counter per class took
// most of this kernel's time.  The order within a class follows
// (copy, arrival) and is any permutation of the class's tiles.
constexpr uint32_t kOrderCopies = 32;
constexpr uint32_t kOrderItems = 8;
__global__ __launch_bounds__(1024) void k_tile_order(const uint2* __restrict__ ranges, uint32_t tiles,
                                                     const uint4* __restrict__ views, uint32_t num_views,
                                                     uint32_t* __restrict__ order) {
    constexpr uint32_t NC = 128 * kOrderCopies;  // counters, class-major (4 per thread)
    __shared__ uint32_t s_cnt[NC];
    __shared__ uint32_t s_wave[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t t0 = 0, t1 = tiles;
    if (views) {
        t0 = views[blockIdx.x].w;
        t1 = blockIdx.x + 1 < num_views ? views[blockIdx.x + 1].w : tiles;
    }
    for (uint32_t i = tid; i < NC; i += 1024) s_cnt[i] = 0;
    __syncthreads();
    auto cls = [](uint32_t c) -> uint32_t {  // without a branch (selects only)
        const uint32_t lg = 31 - __builtin_clz(max(c, 4u));
        const uint32_t big = min(127u, 4 * lg + ((c >> (lg - 2)) & 3u) - 4);
        return c < 4 ? c : big;
    };
    // counter of tile t for this lane: heaviest class first, then the copy
    auto slot_of = [&](uint2 r) -> uint32_t { return (127 - cls(r.y - r.x)) * kOrderCopies + (lane % kOrderCopies); };
    // Each pass loads kOrderItems ranges per thread before using any (one
    // memory round trip per 8192 tiles); with at most 8192 tiles the second
    // pass reuses the first one's ranges.
    uint2 r[kOrderItems];
    for (uint32_t c0 = t0; c0 < t1; c0 += 1024 * kOrderItems) {
#pragma unroll
        for (uint32_t i = 0; i < kOrderItems; ++i) r[i] = ranges[min(c0 + i * 1024 + tid, t1 - 1)];  // in range
#pragma unroll
        for (uint32_t i = 0; i < kOrderItems; ++i)
            if (c0 + i * 1024 + tid < t1) atomicAdd(&s_cnt[slot_of(r[i])], 1u);
    }
    const bool one_chunk = t1 - t0 <= 1024 * kOrderItems;
    __syncthreads();
    {   // exclusive scan of the counters in index order: 4 per thread
        const uint32_t c[4] = {s_cnt[4 * tid], s_cnt[4 * tid + 1], s_cnt[4 * tid + 2], s_cnt[4 * tid + 3]};
        const uint32_t sum = c[0] + c[1] + c[2] + c[3];
        const uint32_t incl = wave_inclusive_sum(sum);
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        uint32_t base = incl - sum;
        for (uint32_t w = 0; w < wave; ++w) base += s_wave[w];
        s_cnt[4 * tid] = base;
        s_cnt[4 * tid + 1] = base + c[0];
        s_cnt[4 * tid + 2] = base + c[0] + c[1];
        s_cnt[4 * tid + 3] = base + c[0] + c[1] + c[2];
    }
    __syncthreads();
    for (uint32_t c0 = t0; c0 < t1; c0 += 1024 * kOrderItems) {
        if (!one_chunk)
#pragma unroll
            for (uint32_t i = 0; i < kOrderItems; ++i) r[i] = ranges[min(c0 + i * 1024 + tid, t1 - 1)];  // in range
#pragma unroll
        for (uint32_t i = 0; i < kOrderItems; ++i) {
            const uint32_t t = c0 + i * 1024 + tid;
            if (t < t1) order[t0 + atomicAdd(&s_cnt[slot_of(r[i])], 1u)] = t;
        }
    }
}

// Workgroups of 4 waves.  Wave w of the grid takes 8x8 block w % BLOCKS of
// tile order[w / BLOCKS] (blocks of a tile in row order), so a tile's blocks
// share a CU.  An 8x8 block is the most compact 64-pixel footprint, which
// culls the most splats per wave.
template <int TS, bool SAVE>
__global__ __launch_bounds__(256) void k_raster(uint32_t W, uint32_t H, uint32_t tiles_x, uint32_t num_tiles,
                                                uint32_t splats, const uint32_t* n_dev, uint32_t n_host,
                                                const uint2* __restrict__ ranges,
                                                const uint32_t* __restrict__ vals,
                                                const SplatRec* __restrict__ rec, float* __restrict__ out,
                                                float* __restrict__ tfinal, uint32_t* __restrict__ last,
                                                const uint32_t* __restrict__ order) {
    constexpr uint32_t SIDE = TS / 8, BLOCKS = SIDE * SIDE;
    const uint32_t wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t slot = wv / BLOCKS, sub = wv % BLOCKS;
    if (slot >= num_tiles) return;
    const uint32_t tile = order[slot];
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const uint2 r = ranges[tile];
    const uint32_t n_inst = n_dev ? *n_dev : n_host;
    const uint32_t end = min(r.y, n_inst);  // ranges never exceed n; bounds the reads
    raster_block<SAVE>(W, H, tx * TS + (sub % SIDE) * 8, ty * TS + (sub / SIDE) * 8, r.x, end, splats, vals, rec,
                       out, tfinal, last);
}

// Persistent form for view batches.  The work items are the 8x8 blocks of
// every (view, tile), tile-major; a workgroup of 4 waves takes 4 consecutive
// items per ticket (one 16-pixel tile, four 8-pixel tiles or a quarter of a
// 32-pixel tile), one per wave, until all are done.  views[v] = {W, H,
// tiles_x, tile_base} and pix[v] the float offset of view v's image.  SAVE:
// T_final and the last contributor too, view v's at pixel offset pix[v] / 3.
template <int TS, bool SAVE>
__global__ __launch_bounds__(256) void k_raster_queue(const uint4* __restrict__ views,
                                                      const uint64_t* __restrict__ pix, uint32_t num_views,
                                                      uint32_t total_tiles, uint32_t splats, const uint32_t* n_dev,
                                                      uint32_t* ticket, const uint2* __restrict__ ranges,
                                                      const uint32_t* __restrict__ vals,
                                                      const SplatRec* __restrict__ rec, float* __restrict__ out,
                                                      float* __restrict__ tfinal, uint32_t* __restrict__ last) {
    constexpr uint32_t SIDE = TS / 8, BLOCKS = SIDE * SIDE;
    __shared__ uint32_t s_ticket[2];
    const uint32_t n_inst = *n_dev;
    const uint32_t items = total_tiles * BLOCKS;
    const uint32_t tickets = ceil_div(items, 4u);
    for (uint32_t round = 0;; ++round) {
        // Double-buffered slot: one barrier per ticket.
        if (threadIdx.x == 0) s_ticket[round & 1] = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t k = s_ticket[round & 1];
        if (k >= tickets) break;  // uniform per workgroup
        const uint32_t it = k * 4 + (threadIdx.x >> 6);
        if (it >= items) continue;
        const uint32_t t = it / BLOCKS, sub = it % BLOCKS;
        // view holding global tile t: last v with tile_base <= t
        uint32_t lo = 0, hi = num_views;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (views[mid].w <= t) lo = mid;
            else hi = mid;
        }
        const uint4 vi = views[lo];
        const uint32_t local = t - vi.w;
        const uint32_t tx = local % vi.z, ty = local / vi.z;
        const uint2 r = ranges[t];
        const uint64_t po = pix[lo];
        raster_block<SAVE>(vi.x, vi.y, tx * TS + (sub % SIDE) * 8, ty * TS + (sub / SIDE) * 8, r.x,
                           min(r.y, n_inst), splats, vals, rec, out + po, SAVE ? tfinal + po / 3 : nullptr,
                           SAVE ? last + po / 3 : nullptr);
    }
}

}  // namespace

int launch_fill_u32(Context* c, uint32_t* p, uint32_t value, size_t count) {
    if (count == 0) return GSR_OK;
    const uint32_t blocks = (uint32_t)std::min<size_t>((count + 255) / 256, 2048);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, c->stream, p, value, count);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

int launch_tile_order(Context* c, const uint2* ranges, uint32_t tiles, uint32_t* order, const uint4* views,
                      uint32_t num_views) {
    if (tiles == 0) return GSR_OK;
    hipLaunchKernelGGL(k_tile_order, dim3(views ? num_views : 1), dim3(1024), 0, c->stream, ranges, tiles, views,
                       num_views, order);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

int launch_zero_regions(Context* c, const FillRegion* regions, int count) {
    if (count < 1 || count > 4) return fail(c, GSR_ERR_ARGUMENT, "zero regions: %d regions", count);
    const ZeroRegions z = make_zero_regions(regions, count);
    const size_t total = z.end[count - 1];
    if (total == 0) return GSR_OK;
    const uint32_t blocks = (uint32_t)std::min<size_t>((total + 255) / 256, 2048);
    hipLaunchKernelGGL(k_zero_regions, dim3(blocks), dim3(256), 0, c->stream, z);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

int launch_ranges(Context* c, const void* keys, uint32_t key_bytes, const uint32_t* n_dev, uint64_t n_host,
                  uint32_t num_tiles, uint32_t* ranges, bool clear) {
    if (clear) {
        int rc = launch_fill_u32(c, ranges, 0u, (size_t)num_tiles * 2);
        if (rc) return rc;
    }
    uint32_t blocks = (uint32_t)c->cu_count * 8;  // device count: sized for any n
    if (!n_dev) blocks = (uint32_t)std::min<uint64_t>(blocks, std::max<uint64_t>(1, ceil_div(n_host, (uint64_t)kRangeChunk)));
    if (key_bytes == 8)
        hipLaunchKernelGGL(k_ranges<uint64_t>, dim3(blocks), dim3(256), 0, c->stream, (const uint64_t*)keys,
                           n_dev, n_host, num_tiles, ranges);
    else
        hipLaunchKernelGGL(k_ranges<uint32_t>, dim3(blocks), dim3(256), 0, c->stream, (const uint32_t*)keys,
                           n_dev, n_host, num_tiles, ranges);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

template <int TS>
static void launch_queue_tile(Context* c, uint32_t groups, const uint4* views, const uint64_t* pix,
                              uint32_t num_views, uint32_t total_tiles, uint32_t splats, const uint32_t* n_dev,
                              uint32_t* ticket, const uint2* rg, const uint32_t* vals, const SplatRec* rec, float* out,
                              float* tfinal, uint32_t* last) {
    if (tfinal && last)
        hipLaunchKernelGGL((k_raster_queue<TS, true>), dim3(groups), dim3(256), 0, c->stream, views, pix, num_views,
                           total_tiles, splats, n_dev, ticket, rg, vals, rec, out, tfinal, last);
    else
        hipLaunchKernelGGL((k_raster_queue<TS, false>), dim3(groups), dim3(256), 0, c->stream, views, pix, num_views,
                           total_tiles, splats, n_dev, ticket, rg, vals, rec, out, nullptr, nullptr);
}

int launch_raster_queue(Context* c, uint32_t tile, const uint4* views, const uint64_t* pix, uint32_t num_views,
                        uint32_t total_tiles, uint32_t splats, const uint32_t* n_dev, uint32_t* ticket,
                        const uint32_t* ranges, const uint32_t* vals, const SplatRec* rec, float* out,
                        uint32_t waves_per_cu, float* tfinal, uint32_t* last) {
    if (total_tiles == 0) return GSR_OK;
    if ((tfinal != nullptr) != (last != nullptr))
        return fail(c, GSR_ERR_ARGUMENT, "batch raster: T_final and last contributor go together");
    const uint32_t tickets = ceil_div(total_tiles * (tile * tile / 64), 4u);
    const uint32_t groups =
        std::max<uint32_t>(1, std::min<uint32_t>(tickets, (uint32_t)c->cu_count * (waves_per_cu ? waves_per_cu : 32) / 4));
    const uint2* rg = (const uint2*)ranges;
    switch (tile) {
        case 8:
            launch_queue_tile<8>(c, groups, views, pix, num_views, total_tiles, splats, n_dev, ticket, rg, vals, rec,
                                 out, tfinal, last);
            break;
        case 16:
            launch_queue_tile<16>(c, groups, views, pix, num_views, total_tiles, splats, n_dev, ticket, rg, vals, rec,
                                  out, tfinal, last);
            break;
        case 32:
            launch_queue_tile<32>(c, groups, views, pix, num_views, total_tiles, splats, n_dev, ticket, rg, vals, rec,
                                  out, tfinal, last);
            break;
        default:
            return fail(c, GSR_ERR_ARGUMENT, "tile size %u not in {8, 16, 32}", tile);
    }
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

template <int TS>
static void launch_raster_tile(Context* c, uint32_t blocks, uint32_t W, uint32_t H, uint32_t tx, uint32_t tiles,
                               uint32_t splats, const uint32_t* n_dev, uint32_t n_host, const uint2* rg,
                               const uint32_t* vals, const SplatRec* rec, float* out, float* tfinal, uint32_t* last,
                               const uint32_t* order) {
    if (tfinal && last)
        hipLaunchKernelGGL((k_raster<TS, true>), dim3(blocks), dim3(256), 0, c->stream, W, H, tx, tiles, splats, n_dev,
                           n_host, rg, vals, rec, out, tfinal, last, order);
    else
        hipLaunchKernelGGL((k_raster<TS, false>), dim3(blocks), dim3(256), 0, c->stream, W, H, tx, tiles, splats, n_dev,
                           n_host, rg, vals, rec, out, nullptr, nullptr, order);
}

// T_final and the last contributor: when only one is requested the other goes
// to a scratch buffer (one kernel writes both or neither).  order: the tiles
// heaviest first (launch_tile_order), computed here when null.
int launch_raster(Context* c, uint32_t W, uint32_t H, uint32_t tile, uint32_t splats, const uint32_t* n_dev,
                  uint32_t n_host, const uint32_t* ranges, const uint32_t* vals, const SplatRec* rec, float* out,
                  float* tfinal, uint32_t* last, const uint32_t* order) {
    const uint32_t tx = ceil_div(W, tile), ty = ceil_div(H, tile);
    const uint32_t tiles = tx * ty;
    if (tiles == 0) return GSR_OK;
    if ((tfinal != nullptr) != (last != nullptr)) {
        void* tmp = scratch(c, "raster_state", (size_t)W * H * 4);
        if (!tmp) return GSR_ERR_ALLOC;
        if (!tfinal) tfinal = (float*)tmp;
        else last = (uint32_t*)tmp;
    }
    const uint2* rg = (const uint2*)ranges;
    const uint32_t blocks = ceil_div(tiles * (tile * tile / 64), 4u);
    if (!order) {
        uint32_t* o = (uint32_t*)scratch(c, "tile_order", (size_t)tiles * 4);
        if (!o) return GSR_ERR_ALLOC;
        const int rc = launch_tile_order(c, rg, tiles, o);
        if (rc) return rc;
        order = o;
    }
    switch (tile) {
        case 8:
            launch_raster_tile<8>(c, blocks, W, H, tx, tiles, splats, n_dev, n_host, rg, vals, rec, out, tfinal,
                                  last, order);
            break;
        case 16:
            launch_raster_tile<16>(c, blocks, W, H, tx, tiles, splats, n_dev, n_host, rg, vals, rec, out, tfinal,
                                   last, order);
            break;
        case 32:
            launch_raster_tile<32>(c, blocks, W, H, tx, tiles, splats, n_dev, n_host, rg, vals, rec, out, tfinal,
                                   last, order);
            break;
        default:
            return fail(c, GSR_ERR_ARGUMENT, "tile size %u not in {8, 16, 32}", tile);
    }
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

}  // namespace gsr
