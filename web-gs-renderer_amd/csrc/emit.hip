// emit.hip — tile instances (src/shaders/emit.wgsl:28-67).
//
// For splat g with a non-empty tile rectangle, slot offsets[g] + k receives
// key ((ty * tiles_x + tx) << 16) | depth16 and value g, for the covered
// tiles in row-major order; slots at or beyond the capacity are dropped (the
// frame reports the overflow instead of truncating silently).  Keys are u32
// or u64 (same value); u64 keys carry view-offset tile ids in batches.
// Offsets are the scan's exclusive prefix (entries + 1 values).
#include "gsr_internal.h"

namespace gsr {
namespace {

// One block per 256 consecutive entries (splats, or (view, splat) pairs in a
// batch).  The block stages each entry's start slot, first tile id, rect
// width, tile-row stride and depth in LDS, then writes its whole slot range
// [offsets[i0], offsets[i0 + 256]) cooperatively: thread t writes slots
// t, t + 256, ..., finding each slot's entry by binary search over the start
// slots, so consecutive threads write consecutive slots (coalesced) whatever
// the per-entry tile counts.  BATCH: entry i belongs to view i / n, whose
// views[v] = {W, H, tiles_x, tile_base}; otherwise tiles_x is given.
template <typename K, bool BATCH>
__global__ __launch_bounds__(256) void k_emit(const uint32_t* __restrict__ offsets, const uint4* __restrict__ meta,
                                              uint64_t entries, uint32_t n, uint32_t tiles_x,
                                              const uint4* __restrict__ views, uint64_t capacity,
                                              K* __restrict__ keys, uint32_t* __restrict__ vals) {
    __shared__ uint32_t s_start[257];  // slot of entry s relative to the block's first slot
    __shared__ uint32_t s_tile0[256];  // tile id of the rect's first tile
    __shared__ uint32_t s_width[256];  // rect width in tiles
    __shared__ uint32_t s_stride[256]; // tiles per row of the entry's view
    __shared__ uint32_t s_depth[256];
    const uint32_t tid = threadIdx.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * 256;
    const uint64_t i = i0 + tid;
    const uint64_t last = min(i0 + 256, entries);
    const uint32_t base = offsets[i0];
    if (i < entries) {
        const uint32_t st = offsets[i];
        const uint32_t cnt = offsets[i + 1] - st;
        s_start[tid] = st - base;
        uint32_t stride = tiles_x, tbase = 0;
        if (BATCH) {
            const uint4 vi = views[i / n];
            stride = vi.z;
            tbase = vi.w;
        }
        uint32_t t0 = 0, w = 1, d = 0;
        if (cnt) {
            const uint4 m = meta[i];
            const uint32_t x0 = m.x & 0xFFFFu, y0 = m.x >> 16, x1 = m.y & 0xFFFFu;
            t0 = tbase + y0 * stride + x0;
            w = x1 - x0 + 1;
            d = m.z & 0xFFFFu;
        }
        s_tile0[tid] = t0;
        s_width[tid] = w;
        s_stride[tid] = stride;
        s_depth[tid] = d;
    } else {
        s_start[tid] = offsets[entries] - base;
    }
    if (tid == 0) s_start[256] = offsets[last] - base;
    __syncthreads();
    const uint32_t total = s_start[256];
    for (uint32_t j = tid; j < total; j += 256) {
        const uint64_t slot = (uint64_t)base + j;
        if (slot >= capacity) break;  // slots only grow with j
        // entry owning slot j: the last s with s_start[s] <= j
        uint32_t lo = 0, hi = 256;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_start[mid] <= j) lo = mid;
            else hi = mid;
        }
        const uint32_t k = j - s_start[lo], w = s_width[lo];
        const uint32_t row = k / w, col = k - row * w;
        const uint32_t tile = s_tile0[lo] + row * s_stride[lo] + col;
        keys[slot] = ((K)tile << 16) | (K)s_depth[lo];
        vals[slot] = (uint32_t)(i0 + lo);
    }
}

}  // namespace

int launch_emit_batch(Context* c, const uint32_t* offsets, const uint4* meta, uint32_t n, uint32_t num_views,
                      const uint4* views, uint64_t capacity, uint32_t key_bytes, void* keys, uint32_t* vals) {
    const uint64_t total = (uint64_t)n * num_views;
    if (total == 0) return GSR_OK;
    const uint32_t blocks = (uint32_t)((total + 255) / 256);
    if (key_bytes == 8)
        hipLaunchKernelGGL((k_emit<uint64_t, true>), dim3(blocks), dim3(256), 0, c->stream, offsets, meta, total, n, 0u,
                           views, capacity, (uint64_t*)keys, vals);
    else
        hipLaunchKernelGGL((k_emit<uint32_t, true>), dim3(blocks), dim3(256), 0, c->stream, offsets, meta, total, n, 0u,
                           views, capacity, (uint32_t*)keys, vals);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

int launch_emit(Context* c, const uint32_t* offsets, const uint4* meta, uint32_t n, uint32_t tiles_x,
                uint64_t capacity, uint32_t key_bytes, void* keys, uint32_t* vals) {
    if (n == 0) return GSR_OK;
    if (key_bytes == 8)
        hipLaunchKernelGGL((k_emit<uint64_t, false>), dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, offsets, meta,
                           (uint64_t)n, n, tiles_x, nullptr, capacity, (uint64_t*)keys, vals);
    else
        hipLaunchKernelGGL((k_emit<uint32_t, false>), dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, offsets, meta,
                           (uint64_t)n, n, tiles_x, nullptr, capacity, (uint32_t*)keys, vals);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

}  // namespace gsr
