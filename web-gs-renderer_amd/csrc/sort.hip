// sort.hip — stable LSD radix sort of (key, value) pairs, one sweep per digit.
//
// Replaces the reference's 5-kernel-per-pass sort (src/shaders/radix_parallel.wgsl:26-160,
// driven by src/renderer/renderer.js:869-917).  Same result: a stable sort on
// the key bits [begin_bit, end_bit), hence the (key, value)-lexicographic
// order of the emitted instances.  Structure per sort:
//   k_sort_init      zero digit histograms, tile tickets and look-back words
//                    (only for the tiles the device-side count needs);
//   k_sort_hist      one read of the keys builds every pass's 256-bin digit
//                    histogram (LDS bins; a run of equal digits in
//                    consecutive lanes adds once; one global atomic per bin
//                    and block, one block per CU);
//   k_onesweep x P   persistent blocks take tiles by ticket, rank keys per
//                    wave64 with ballot digit matching (stable), chain
//                    per-digit tile prefixes with decoupled look-back, stage
//                    the tile digit-sorted in LDS and write it out in runs.
// Onesweep geometry: THREADS = 1024 / blocks-per-CU threads (1, 2 or 4
// blocks per CU) with 16 items per thread for 32-bit keys and 8 for 64-bit
// keys.  Large tiles keep the look-back chains short: a tile's look-back
// walks back through the tiles that started with it, 16 per round trip, so
// the count of tiles in flight sets its latency.
// The key count may live on the device (the frame pipeline), so every grid is
// sized from the capacity and blocks exit on the device count: no host
// readback, capturable in a hipGraph.
#include "gsr_internal.h"
#include "wave.h"

namespace gsr {
namespace {

constexpr uint32_t kAgg = 1u << 30, kInc = 2u << 30;
constexpr int kLookBack = 16;   // predecessor words fetched per look-back step
constexpr int kHistItems = 16;  // keys per thread per histogram block tile

template <typename K, int THREADS>
struct SortGeom {
    static constexpr int kThreads = THREADS;
    static constexpr int kWaves = THREADS / 64;
    static constexpr int kItems = sizeof(K) == 4 ? 16 : 8;
    static constexpr int kTile = THREADS * kItems;
};

__device__ __forceinline__ uint32_t read_n(const uint32_t* n_dev, uint64_t n_host) {
    return n_dev ? *n_dev : (uint32_t)n_host;
}

// Exclusive prefix over threads 0..255 of a block (waves 0-3).  Every thread
// of the block calls it (threads >= 256 pass 0); s4 holds 4 words.
__device__ __forceinline__ uint32_t digit_exclusive(uint32_t v, uint32_t* s4) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane == 63 && wave < 4) s4[wave] = incl;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w)
        if (w < (int)wave) pre += s4[w];
    __syncthreads();
    return pre + incl - v;
}

__global__ __launch_bounds__(256) void k_sort_init(uint32_t* __restrict__ hist_tickets,
                                                   uint32_t* __restrict__ status, const uint32_t* n_dev,
                                                   uint64_t n_host, uint32_t passes, uint32_t max_tiles,
                                                   uint32_t tile_keys) {
    const uint32_t n = read_n(n_dev, n_host);
    const uint32_t tiles = ceil_div(n, tile_keys);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = t0; i < (size_t)kMaxPasses * kRadix + kMaxPasses; i += stride) hist_tickets[i] = 0;
    const size_t per_pass = (size_t)tiles * kRadix;
    for (uint32_t p = 0; p < passes; ++p) {
        uint32_t* s = status + (size_t)p * max_tiles * kRadix;
        for (size_t i = t0; i < per_pass; i += stride) s[i] = 0;
    }
}

// 1024-thread blocks (one per CU) take 1024 x kHistItems-key tiles in a
// grid-stride loop; each thread issues all kHistItems loads of a tile before
// counting, so the loop is not bound by one load's latency per key.  Waves
// w and w + 4k share LDS copy w % 4 of the bins.
template <typename K>
__global__ __launch_bounds__(1024) void k_sort_hist(const K* __restrict__ keys, const uint32_t* n_dev,
                                                   uint64_t n_host, uint32_t begin_bit, uint32_t end_bit,
                                                   uint32_t passes, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_hist[4][kMaxPasses * kRadix];
    const uint32_t n = read_n(n_dev, n_host);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, copy = wave & 3;
    for (uint32_t w = 0; w < 4; ++w)
        for (uint32_t i = threadIdx.x; i < passes * kRadix; i += blockDim.x) s_hist[w][i] = 0;
    __syncthreads();
    constexpr uint32_t kBlockKeys = 1024 * kHistItems;
    for (uint32_t t = blockIdx.x; (uint64_t)t * kBlockKeys < n; t += gridDim.x) {
        const uint32_t wbase = t * kBlockKeys + wave * (64 * kHistItems);
        K k[kHistItems];
#pragma unroll
        for (int i = 0; i < kHistItems; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            k[i] = p < n ? keys[p] : (K)0;
        }
#pragma unroll
        for (int i = 0; i < kHistItems; ++i) {
            const uint32_t p0 = wbase + i * 64;
            if (p0 >= n) break;  // wave-uniform
            const uint32_t valid = min(64u, n - p0);  // lanes [0, valid) hold keys
            const bool ok = lane < valid;
            for (uint32_t p = 0; p < passes; ++p) {
                const uint32_t shift = begin_bit + p * kRadixBits;
                const uint32_t bits = min((uint32_t)kRadixBits, end_bit - shift);
                const uint32_t d = (uint32_t)(k[i] >> shift) & ((1u << bits) - 1u);
                // previous lane's digit (wave_shr:1); lane 0 always starts a run
                const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)d, 0x138, 0xF, 0xF, false);
                const bool head = ok && (lane == 0 || prev != d);
                const uint64_t heads = ballot(head);
                if (head) {
                    const uint64_t after = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
                    const uint32_t next = after ? (uint32_t)__builtin_ctzll(after) : valid;
                    atomicAdd(&s_hist[copy][p * kRadix + d], next - lane);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < passes * kRadix; i += blockDim.x) {
        const uint32_t v = s_hist[0][i] + s_hist[1][i] + s_hist[2][i] + s_hist[3][i];
        if (v) atomicAdd(&hist[i], v);
    }
}

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_onesweep(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* n_dev, uint64_t n_host, uint32_t shift, uint32_t bits,
    const uint32_t* __restrict__ hist, uint32_t* ticket, uint32_t* status, uint32_t* error) {
    using G = SortGeom<K, THREADS>;
    constexpr int WAVES = G::kWaves, ITEMS = G::kItems, TILE = G::kTile;
    __shared__ uint32_t s_wave_cnt[WAVES][kRadix];
    __shared__ uint32_t s_global[kRadix];
    __shared__ uint32_t s_local[kRadix];
    __shared__ uint32_t s_scatter[kRadix];
    __shared__ K s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s4[4];
    __shared__ uint32_t s_tile;

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool digit_thread = tid < (uint32_t)kRadix;  // threads 0..255 own one digit each
    const uint32_t n = read_n(n_dev, n_host);
    const uint32_t num_tiles = ceil_div(n, (uint32_t)TILE);
    const uint32_t mask = (1u << bits) - 1u;
    if (blockIdx.x >= num_tiles) return;  // uniform per block
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    {
        const uint32_t g = digit_exclusive(digit_thread ? hist[tid] : 0u, s4);
        if (digit_thread) s_global[tid] = g;
    }

    for (;;) {
        for (int i = tid; i < WAVES * kRadix; i += THREADS) (&s_wave_cnt[0][0])[i] = 0;
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= num_tiles) break;
        const uint32_t wbase = tile * TILE + wave * (64 * ITEMS);
        K key[ITEMS];
        uint32_t val[ITEMS];
        uint32_t rank[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            const bool ok = p < n;
            key[i] = ok ? kin[p] : (K)0;
            val[i] = ok ? vin[p] : 0u;
        }
        // Stable per-wave ranks: items in position order, lanes in lane order.
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
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
        // Per digit: wave prefixes, tile count, publication, look-back.
        uint32_t cnt = 0;
        uint32_t* my_status = status + (size_t)tile * kRadix + tid;
        if (digit_thread) {
#pragma unroll
            for (int w = 0; w < WAVES; ++w) {
                const uint32_t c = s_wave_cnt[w][tid];
                s_wave_cnt[w][tid] = cnt;
                cnt += c;
            }
            st_release_u32(my_status, (tile == 0 ? kInc : kAgg) | cnt);
        }
        const uint32_t local = digit_exclusive(cnt, s4);
        uint32_t excl = 0;
        if (digit_thread && tile > 0) {
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
        if (digit_thread) {
            s_local[tid] = local;
            s_scatter[tid] = s_global[tid] + excl - local;
        }
        __syncthreads();
        // Digit-sorted staging of the tile in LDS.
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            if (p < n) {
                const uint32_t d = (uint32_t)(key[i] >> shift) & mask;
                const uint32_t lp = s_local[d] + s_wave_cnt[wave][d] + rank[i];
                if (lp < (uint32_t)TILE) {
                    s_keys[lp] = key[i];
                    s_vals[lp] = val[i];
                }
            }
        }
        __syncthreads();
        const uint32_t tile_n = min((uint32_t)TILE, n - tile * TILE);
        for (uint32_t idx = tid; idx < tile_n; idx += THREADS) {
            const K k = s_keys[idx];
            const uint32_t d = (uint32_t)(k >> shift) & mask;
            const uint32_t o = s_scatter[d] + idx;
            if (o < n) {  // always true for a consistent histogram; guards the buffer otherwise
                kout[o] = k;
                vout[o] = s_vals[idx];
            }
        }
        // The next ticket is taken only now: a ticket held while this tile is
        // still in flight would hold up that tile's successors.
        if (tid == 0) s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
    }
}

template <typename K, int THREADS>
int sort_impl(Context* c, K* k0, uint32_t* v0, K* k1, uint32_t* v1, const uint32_t* n_dev, uint64_t n_host,
              uint64_t capacity, uint32_t begin_bit, uint32_t end_bit, FrameCounters* counters,
              int* result_index) {
    using G = SortGeom<K, THREADS>;
    const uint32_t passes = end_bit > begin_bit ? ceil_div(end_bit - begin_bit, kRadixBits) : 0;
    *result_index = (int)(passes & 1u);
    if (passes == 0) return GSR_OK;
    if (passes > (uint32_t)kMaxPasses) return fail(c, GSR_ERR_ARGUMENT, "too many radix passes (%u)", passes);
    const uint32_t max_tiles = std::max<uint32_t>(1, ceil_div(capacity, (uint64_t)G::kTile));
    uint32_t* hist = (uint32_t*)scratch(c, "sort_hist", ((size_t)kMaxPasses * kRadix + kMaxPasses) * 4);
    uint32_t* status = (uint32_t*)scratch(c, "sort_status", (size_t)passes * max_tiles * kRadix * 4);
    if (!hist || !status) return GSR_ERR_ALLOC;
    uint32_t* tickets = hist + kMaxPasses * kRadix;
    const uint32_t fill_blocks = std::min<uint32_t>(std::max<uint32_t>(1, max_tiles), 1024);
    hipLaunchKernelGGL(k_sort_init, dim3(fill_blocks), dim3(256), 0, c->stream, hist, status, n_dev, n_host,
                       passes, max_tiles, (uint32_t)G::kTile);
    // One block per CU: every block adds its bins to the global histogram with
    // one atomic each, and those atomics serialize per bin across blocks, so
    // fewer, larger blocks.
    const uint32_t hist_blocks = std::min<uint32_t>(
        std::max<uint32_t>(1, ceil_div(capacity, (uint64_t)1024 * kHistItems)), (uint32_t)c->cu_count);
    hipLaunchKernelGGL(k_sort_hist<K>, dim3(hist_blocks), dim3(1024), 0, c->stream, k0, n_dev, n_host,
                       begin_bit, end_bit, passes, hist);
    const uint32_t sweep_blocks = std::min<uint32_t>(max_tiles, (uint32_t)c->cu_count * (1024 / THREADS));
    K* kb[2] = {k0, k1};
    uint32_t* vb[2] = {v0, v1};
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = begin_bit + p * kRadixBits;
        const uint32_t bits = std::min<uint32_t>(kRadixBits, end_bit - shift);
        hipLaunchKernelGGL((k_onesweep<K, THREADS>), dim3(sweep_blocks), dim3(THREADS), 0, c->stream, kb[p & 1],
                           vb[p & 1], kb[(p + 1) & 1], vb[(p + 1) & 1], n_dev, n_host, shift, bits,
                           hist + p * kRadix, tickets + p, status + (size_t)p * max_tiles * kRadix,
                           &counters->error);
    }
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

// Onesweep blocks per CU: 1, 2 or 4 (1024-, 512- or 256-thread blocks).
template <typename K>
int sort_geometry(Context* c, K* k0, uint32_t* v0, K* k1, uint32_t* v1, const uint32_t* n_dev, uint64_t n_host,
                  uint64_t capacity, uint32_t begin_bit, uint32_t end_bit, uint32_t blocks_per_cu,
                  FrameCounters* counters, int* result_index) {
    switch (blocks_per_cu) {
        case 1:
            return sort_impl<K, 1024>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, counters,
                                      result_index);
        case 2:
            return sort_impl<K, 512>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, counters,
                                     result_index);
        case 4:
            return sort_impl<K, 256>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, counters,
                                     result_index);
        default:
            return fail(c, GSR_ERR_ARGUMENT, "sort blocks per CU %u not in {1, 2, 4}", blocks_per_cu);
    }
}

}  // namespace

int launch_sort(Context* c, uint32_t key_bytes, void* k0, uint32_t* v0, void* k1, uint32_t* v1,
                const uint32_t* n_dev, uint64_t n_host, uint64_t capacity, uint32_t begin_bit,
                uint32_t end_bit, uint32_t blocks_per_cu, FrameCounters* counters, int* result_index) {
    if (capacity > kMaxSortKeys) return fail(c, GSR_ERR_ARGUMENT, "sort capacity %llu exceeds %llu",
                                             (unsigned long long)capacity, (unsigned long long)kMaxSortKeys);
    if (end_bit > key_bytes * 8 || begin_bit > end_bit)
        return fail(c, GSR_ERR_ARGUMENT, "bad bit range [%u, %u) for %u-byte keys", begin_bit, end_bit, key_bytes);
    if (key_bytes == 8)
        return sort_geometry<uint64_t>(c, (uint64_t*)k0, v0, (uint64_t*)k1, v1, n_dev, n_host, capacity, begin_bit,
                                       end_bit, blocks_per_cu, counters, result_index);
    return sort_geometry<uint32_t>(c, (uint32_t*)k0, v0, (uint32_t*)k1, v1, n_dev, n_host, capacity, begin_bit,
                                   end_bit, blocks_per_cu, counters, result_index);
}

}  // namespace gsr
