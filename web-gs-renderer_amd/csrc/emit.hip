// emit.hip — tile instances (src/shaders/emit.wgsl:28-67).
//
// For splat g with a non-empty tile rectangle, slot offsets[g] + k receives
// key ((ty * tiles_x + tx) << 16) | depth16 and value g, for the covered
// tiles in row-major order; slots at or beyond the capacity are dropped (the
// frame reports the overflow instead of truncating silently).  Keys are u32
// or u64 (same value); u64 keys carry view-offset tile ids in batches.
#include "gsr_internal.h"

namespace gsr {
namespace {

template <typename K>
__global__ __launch_bounds__(256) void k_emit(const uint32_t* __restrict__ offsets,
                                              const uint4* __restrict__ meta, uint32_t n,
                                              uint32_t tiles_x, uint64_t capacity, K* __restrict__ keys,
                                              uint32_t* __restrict__ vals) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint32_t base = offsets[g];
    const uint32_t cnt = offsets[g + 1] - base;
    if (cnt == 0) return;
    const uint4 m = meta[g];
    const uint32_t x0 = m.x & 0xFFFFu, y0 = m.x >> 16, x1 = m.y & 0xFFFFu, y1 = m.y >> 16;
    const K d16 = (K)(m.z & 0xFFFFu);
    uint64_t slot = base;
    for (uint32_t ty = y0; ty <= y1; ++ty) {
        for (uint32_t tx = x0; tx <= x1; ++tx) {
            if (slot >= capacity) return;
            keys[slot] = ((K)(ty * tiles_x + tx) << 16) | d16;
            vals[slot] = g;
            ++slot;
        }
    }
}

// View batches: entry i = v * n + g of the batched preprocess outputs; the
// key carries the global tile id tile_base[v] + ty * tiles_x[v] + tx and the
// value is i.  views[v] = {W, H, tiles_x, tile_base}.
template <typename K>
__global__ __launch_bounds__(256) void k_emit_batch(const uint32_t* __restrict__ offsets,
                                                    const uint4* __restrict__ meta, uint32_t n, uint32_t num_views,
                                                    const uint4* __restrict__ views, uint64_t capacity,
                                                    K* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n * num_views) return;
    const uint32_t base = offsets[i];
    const uint32_t cnt = offsets[i + 1] - base;
    if (cnt == 0) return;
    const uint4 vi = views[i / n];
    const uint4 m = meta[i];
    const uint32_t x0 = m.x & 0xFFFFu, y0 = m.x >> 16, x1 = m.y & 0xFFFFu, y1 = m.y >> 16;
    const K d16 = (K)(m.z & 0xFFFFu);
    uint64_t slot = base;
    for (uint32_t ty = y0; ty <= y1; ++ty) {
        for (uint32_t tx = x0; tx <= x1; ++tx) {
            if (slot >= capacity) return;
            keys[slot] = ((K)(vi.w + ty * vi.z + tx) << 16) | d16;
            vals[slot] = (uint32_t)i;
            ++slot;
        }
    }
}

}  // namespace

int launch_emit_batch(Context* c, const uint32_t* offsets, const uint4* meta, uint32_t n, uint32_t num_views,
                      const uint4* views, uint64_t capacity, uint32_t key_bytes, void* keys, uint32_t* vals) {
    const uint64_t total = (uint64_t)n * num_views;
    if (total == 0) return GSR_OK;
    const uint32_t blocks = (uint32_t)((total + 255) / 256);
    if (key_bytes == 8)
        hipLaunchKernelGGL(k_emit_batch<uint64_t>, dim3(blocks), dim3(256), 0, c->stream, offsets, meta, n, num_views,
                           views, capacity, (uint64_t*)keys, vals);
    else
        hipLaunchKernelGGL(k_emit_batch<uint32_t>, dim3(blocks), dim3(256), 0, c->stream, offsets, meta, n, num_views,
                           views, capacity, (uint32_t*)keys, vals);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

int launch_emit(Context* c, const uint32_t* offsets, const uint4* meta, uint32_t n, uint32_t tiles_x,
                uint64_t capacity, uint32_t key_bytes, void* keys, uint32_t* vals) {
    if (n == 0) return GSR_OK;
    if (key_bytes == 8)
        hipLaunchKernelGGL(k_emit<uint64_t>, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, offsets, meta,
                           n, tiles_x, capacity, (uint64_t*)keys, vals);
    else
        hipLaunchKernelGGL(k_emit<uint32_t>, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, offsets, meta,
                           n, tiles_x, capacity, (uint32_t*)keys, vals);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

}  // namespace gsr
