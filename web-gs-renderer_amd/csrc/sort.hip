// sort.hip — stable LSD radix sort of (key, value) pairs, one sweep per digit.
//
// Replaces the reference's 5-kernel-per-pass sort (src/shaders/radix_parallel.wgsl:26-160,
// driven by src/renderer/renderer.js:869-917).  Same result: a stable sort on
// the key bits [begin_bit, end_bit), hence the (key, value)-lexicographic
// order of the emitted instances.  Structure per sort:
//   k_sort_init      zero digit histograms, tile tickets and look-back words
//                    (only for the tiles the device-side count needs);
//   k_sort_hist      one read of the keys builds every pass's 256-bin digit
//                    histogram (LDS bins, then one global atomic per bin);
//   k_onesweep x P   persistent blocks take 4096-key tiles by ticket, rank
//                    keys per wave64 with ballot digit matching (stable), chain
//                    per-digit tile prefixes with decoupled look-back, stage the
//                    tile digit-sorted in LDS and write it out in runs.
// The key count may live on the device (the frame pipeline), so every grid is
// sized from the capacity and blocks exit on the device count: no host
// readback, capturable in a hipGraph.
#include "gsr_internal.h"
#include "wave.h"

namespace gsr {
namespace {

constexpr uint32_t kAgg = 1u << 30, kInc = 2u << 30;
constexpr int kLookBack = 16;  // predecessor words fetched per look-back step

__device__ __forceinline__ uint32_t read_n(const uint32_t* n_dev, uint64_t n_host) {
    return n_dev ? *n_dev : (uint32_t)n_host;
}

// Exclusive prefix over the 256 threads of a block (4 waves); s4 holds 4 words.
__device__ __forceinline__ uint32_t block_exclusive_256(uint32_t v, uint32_t* s4, uint32_t* total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane == 63) s4[wave] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t t = s4[w];
        if (w < (int)wave) pre += t;
        tot += t;
    }
    __syncthreads();
    if (total) *total = tot;
    return pre + incl - v;
}

__global__ __launch_bounds__(256) void k_sort_init(uint32_t* __restrict__ hist_tickets,
                                                   uint32_t* __restrict__ status, const uint32_t* n_dev,
                                                   uint64_t n_host, uint32_t passes, uint32_t max_tiles) {
    const uint32_t n = read_n(n_dev, n_host);
    const uint32_t tiles = ceil_div(n, kSortTile);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = t0; i < (size_t)kMaxPasses * kRadix + kMaxPasses; i += stride) hist_tickets[i] = 0;
    const size_t per_pass = (size_t)tiles * kRadix;
    for (uint32_t p = 0; p < passes; ++p) {
        uint32_t* s = status + (size_t)p * max_tiles * kRadix;
        for (size_t i = t0; i < per_pass; i += stride) s[i] = 0;
    }
}

template <typename K>
__global__ __launch_bounds__(256) void k_sort_hist(const K* __restrict__ keys, const uint32_t* n_dev,
                                                   uint64_t n_host, uint32_t begin_bit, uint32_t end_bit,
                                                   uint32_t passes, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_hist[kMaxPasses * kRadix];
    const uint32_t n = read_n(n_dev, n_host);
    for (uint32_t i = threadIdx.x; i < passes * kRadix; i += blockDim.x) s_hist[i] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const K k = keys[i];
        for (uint32_t p = 0; p < passes; ++p) {
            const uint32_t shift = begin_bit + p * kRadixBits;
            const uint32_t bits = min((uint32_t)kRadixBits, end_bit - shift);
            const uint32_t d = (uint32_t)(k >> shift) & ((1u << bits) - 1u);
            atomicAdd(&s_hist[p * kRadix + d], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < passes * kRadix; i += blockDim.x) {
        const uint32_t v = s_hist[i];
        if (v) atomicAdd(&hist[i], v);
    }
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* n_dev, uint64_t n_host, uint32_t shift, uint32_t bits,
    const uint32_t* __restrict__ hist, uint32_t* ticket, uint32_t* status, uint32_t* error) {
    __shared__ uint32_t s_wave_cnt[4][kRadix];
    __shared__ uint32_t s_global[kRadix];
    __shared__ uint32_t s_local[kRadix];
    __shared__ uint32_t s_scatter[kRadix];
    __shared__ K s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s4[4];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n = read_n(n_dev, n_host);
    const uint32_t num_tiles = ceil_div(n, kSortTile);
    const uint32_t mask = (1u << bits) - 1u;
    if (blockIdx.x >= num_tiles) return;  // uniform per block
    s_global[tid] = block_exclusive_256(hist[tid], s4, nullptr);

    for (;;) {
        if (tid == 0) s_tile = atomicAdd(ticket, 1u);
#pragma unroll
        for (int w = 0; w < 4; ++w) s_wave_cnt[w][tid] = 0;
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= num_tiles) break;
        const uint32_t wbase = tile * kSortTile + wave * (64 * kSortItems);
        K key[kSortItems];
        uint32_t val[kSortItems];
        uint32_t rank[kSortItems];
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            const bool ok = p < n;
            key[i] = ok ? kin[p] : (K)0;
            val[i] = ok ? vin[p] : 0u;
        }
        // Stable per-wave ranks: items in position order, lanes in lane order.
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            const bool ok = p < n;
            const uint32_t d = (uint32_t)(key[i] >> shift) & mask;
            uint64_t m = ballot(ok);
            for (uint32_t b = 0; b < bits; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t bb = ballot(bit);
                m &= bit ? bb : ~bb;
            }
            const uint32_t below = count_below(m);
            uint32_t prior = 0;
            if (ok) prior = s_wave_cnt[wave][d];
            rank[i] = prior + below;
            if (ok && below == 0) s_wave_cnt[wave][d] = prior + (uint32_t)__popcll(m);
        }
        __syncthreads();
        // Per digit (thread = digit): wave prefixes, tile count, look-back.
        const uint32_t c0 = s_wave_cnt[0][tid], c1 = s_wave_cnt[1][tid];
        const uint32_t c2 = s_wave_cnt[2][tid], c3 = s_wave_cnt[3][tid];
        const uint32_t cnt = c0 + c1 + c2 + c3;
        uint32_t* my_status = status + (size_t)tile * kRadix + tid;
        st_release_u32(my_status, (tile == 0 ? kInc : kAgg) | cnt);
        s_wave_cnt[0][tid] = 0;
        s_wave_cnt[1][tid] = c0;
        s_wave_cnt[2][tid] = c0 + c1;
        s_wave_cnt[3][tid] = c0 + c1 + c2;
        const uint32_t local = block_exclusive_256(cnt, s4, nullptr);
        uint32_t excl = 0;
        if (tile > 0) {
            // Batched look-back: the words of up to kLookBack predecessors are
            // loaded together (independent loads, one round trip), then
            // consumed nearest first; only a word still not ready is re-polled.
            int64_t j = (int64_t)tile - 1;
            uint32_t spins = 0;
            bool found = false;
            while (!found && j >= 0) {
                uint32_t w[kLookBack];
#pragma unroll
                for (int d = 0; d < kLookBack; ++d)
                    w[d] = (j - d >= 0) ? ld_acquire_u32(status + (size_t)(j - d) * kRadix + tid) : kInc;
#pragma unroll
                for (int d = 0; d < kLookBack; ++d) {
                    if (found || j - d < 0) continue;
                    uint32_t s = w[d];
                    while ((s >> 30) == 0) {
                        if (++spins > kSpinLimit) {
                            atomicOr(error, 2u);
                            s = kInc;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        s = ld_acquire_u32(status + (size_t)(j - d) * kRadix + tid);
                    }
                    excl += s & kStatusValueMask;
                    if (s & kInc) found = true;
                }
                j -= kLookBack;
            }
            st_release_u32(my_status, kInc | ((excl + cnt) & kStatusValueMask));
        }
        s_local[tid] = local;
        s_scatter[tid] = s_global[tid] + excl - local;
        __syncthreads();
        // Digit-sorted staging of the tile in LDS.
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            if (p < n) {
                const uint32_t d = (uint32_t)(key[i] >> shift) & mask;
                const uint32_t lp = s_local[d] + s_wave_cnt[wave][d] + rank[i];
                if (lp < (uint32_t)kSortTile) {
                    s_keys[lp] = key[i];
                    s_vals[lp] = val[i];
                }
            }
        }
        __syncthreads();
        const uint32_t tile_n = min((uint32_t)kSortTile, n - tile * kSortTile);
        for (uint32_t idx = tid; idx < tile_n; idx += kSortThreads) {
            const K k = s_keys[idx];
            const uint32_t d = (uint32_t)(k >> shift) & mask;
            const uint32_t o = s_scatter[d] + idx;
            if (o < n) {  // always true for a consistent histogram; guards the buffer otherwise
                kout[o] = k;
                vout[o] = s_vals[idx];
            }
        }
        __syncthreads();
    }
}

template <typename K>
int sort_impl(Context* c, K* k0, uint32_t* v0, K* k1, uint32_t* v1, const uint32_t* n_dev, uint64_t n_host,
              uint64_t capacity, uint32_t begin_bit, uint32_t end_bit, FrameCounters* counters,
              int* result_index) {
    const uint32_t passes = end_bit > begin_bit ? ceil_div(end_bit - begin_bit, kRadixBits) : 0;
    *result_index = (int)(passes & 1u);
    if (passes == 0) return GSR_OK;
    if (passes > (uint32_t)kMaxPasses) return fail(c, GSR_ERR_ARGUMENT, "too many radix passes (%u)", passes);
    const uint32_t max_tiles = std::max<uint32_t>(1, ceil_div(capacity, kSortTile));
    uint32_t* hist = (uint32_t*)scratch(c, "sort_hist", ((size_t)kMaxPasses * kRadix + kMaxPasses) * 4);
    uint32_t* status = (uint32_t*)scratch(c, "sort_status", (size_t)passes * max_tiles * kRadix * 4);
    if (!hist || !status) return GSR_ERR_ALLOC;
    uint32_t* tickets = hist + kMaxPasses * kRadix;
    const uint32_t fill_blocks = std::min<uint32_t>(std::max<uint32_t>(1, max_tiles), 1024);
    hipLaunchKernelGGL(k_sort_init, dim3(fill_blocks), dim3(256), 0, c->stream, hist, status, n_dev, n_host,
                       passes, max_tiles);
    const uint32_t hist_blocks = std::min<uint32_t>(std::max<uint32_t>(1, ceil_div(capacity, 4096)),
                                                    (uint32_t)c->cu_count * 4);
    hipLaunchKernelGGL(k_sort_hist<K>, dim3(hist_blocks), dim3(256), 0, c->stream, k0, n_dev, n_host,
                       begin_bit, end_bit, passes, hist);
    const uint32_t sweep_blocks =
        std::min<uint32_t>(max_tiles, (uint32_t)c->cu_count * std::max<uint32_t>(1, c->sort_blocks_per_cu));
    K* kb[2] = {k0, k1};
    uint32_t* vb[2] = {v0, v1};
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = begin_bit + p * kRadixBits;
        const uint32_t bits = std::min<uint32_t>(kRadixBits, end_bit - shift);
        hipLaunchKernelGGL(k_onesweep<K>, dim3(sweep_blocks), dim3(kSortThreads), 0, c->stream, kb[p & 1],
                           vb[p & 1], kb[(p + 1) & 1], vb[(p + 1) & 1], n_dev, n_host, shift, bits,
                           hist + p * kRadix, tickets + p, status + (size_t)p * max_tiles * kRadix,
                           &counters->error);
    }
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

}  // namespace

int launch_sort(Context* c, uint32_t key_bytes, void* k0, uint32_t* v0, void* k1, uint32_t* v1,
                const uint32_t* n_dev, uint64_t n_host, uint64_t capacity, uint32_t begin_bit,
                uint32_t end_bit, FrameCounters* counters, int* result_index) {
    if (capacity > kMaxSortKeys) return fail(c, GSR_ERR_ARGUMENT, "sort capacity %llu exceeds %llu",
                                             (unsigned long long)capacity, (unsigned long long)kMaxSortKeys);
    if (end_bit > key_bytes * 8 || begin_bit > end_bit)
        return fail(c, GSR_ERR_ARGUMENT, "bad bit range [%u, %u) for %u-byte keys", begin_bit, end_bit, key_bytes);
    if (key_bytes == 8)
        return sort_impl<uint64_t>(c, (uint64_t*)k0, v0, (uint64_t*)k1, v1, n_dev, n_host, capacity, begin_bit,
                                   end_bit, counters, result_index);
    return sort_impl<uint32_t>(c, (uint32_t*)k0, v0, (uint32_t*)k1, v1, n_dev, n_host, capacity, begin_bit,
                               end_bit, counters, result_index);
}

}  // namespace gsr
