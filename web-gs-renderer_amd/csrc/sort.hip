// sort.hip — stable LSD radix sort of (key, value) pairs, one sweep per digit.
//
// Replaces the reference's 5-kernel-per-pass sort (src/shaders/radix_parallel.wgsl:26-160,
// driven by src/renderer/renderer.js:869-917).  Same result: a stable sort on
// the key bits [begin_bit, end_bit), hence the (key, value)-lexicographic
// order of the emitted instances.  Structure per sort:
//   k_sort_init      zero digit histograms and tile tickets (skipped when
//                    the frame's preprocess grid already cleared them);
//   k_sort_hist      zero the first pass's look-back words (only for the
//                    tiles the device-side count needs); one read of the
//                    keys builds every pass's digit histogram (LDS bins; a
//                    run of equal digits in consecutive lanes adds once; one
//                    global atomic per bin and block, into one of several
//                    copies of the histograms that onesweep sums);
//   k_onesweep x P   persistent blocks take tiles by ticket, rank keys per
//                    wave64 with ballot digit matching (stable), chain
//                    per-digit tile prefixes with decoupled look-back, stage
//                    the tile digit-sorted in LDS and write it out in runs.
// Geometry: THREADS = 1024 / blocks-per-CU threads (1, 2 or 4 blocks per
// CU), digits of 8 to 11 bits (256 to 2048 buckets: a 33- or 34-bit
// view|tile|depth key takes 4 passes at 9 bits instead of 5 at 8, a 29-bit
// frame key 3 at 10 or 11 bits), and up to as many items per thread as the
// CU's LDS holds (SortGeom).  Large tiles keep the look-back
// chains short: a tile's look-back walks back through the tiles that
// started with it, so the count of tiles in flight sets its latency.  Below
// one full tile per resident block the tiles shrink to about one per block
// (onesweep_items, from the device-side count; each length is its own
// fully unrolled instantiation): a 3.3M-key frame sort then runs 14 items
// per thread in one round of 233 tiles, a 1M-key one 4 items in 245.  Where
// the block has more threads than digits, the threads of one digit share its
// look-back (4 predecessors per thread and round trip, the parts combined in
// order with lane shuffles); otherwise each digit's thread loads 16 per round
// trip.
// The key count may live on the device (the frame pipeline), so every grid is
// sized from the capacity and blocks exit on the device count: no host
// readback, capturable in a hipGraph.
#include <type_traits>

#include "gsr_internal.h"
#include "wave.h"

namespace gsr {
namespace {

constexpr uint32_t kAgg = 1u << 30, kInc = 2u << 30;
constexpr int kLookBack = 16;   // predecessor words fetched per look-back step (all digits of a thread)
constexpr int kHistItems = 8;   // keys per thread per histogram block tile
// Copies of the global digit histograms that the histogram blocks add into
// (block b into copy b % copies; onesweep sums them), so that the flush of
// some 400K atomics (1080p frame) does not serialise on one set of cache
// lines.  Measured in one box session on the 1080p frame's 32-bit keys: 1,
// 2, 4, 8 and 16 copies gave a sort of 0.1143, 0.1129, 0.1143, 0.1191 and
// 0.1267 ms; on the 64-view batch's 64-bit keys 1 and 8 copies were equal
// and 2 or 4 1.5% slower.  kHistCopies is the most any sort uses; scratch
// the caller provides (the no-allocation sort) keeps one copy.
constexpr uint32_t kHistCopies = 8;
template <typename K>
constexpr uint32_t hist_copies() { return sizeof(K) == 4 ? 2u : kHistCopies; }
// onesweep keys per thread at most for 64-bit keys in 1024-thread blocks
// (8 in smaller blocks, whose several blocks per CU need fewer registers)
constexpr int kMaxItems64 = 12;
constexpr size_t kLdsBytes = 160 * 1024;

// Onesweep geometry for keys K, THREADS threads per block and RB-bit digits.
// Each thread ranks up to kItems keys; the full tile is THREADS * kItems
// keys, the most items (up to 16, fewer for 64-bit keys) whose
// LDS (keys and values staged, 16-bit per-wave digit counts, three digit
// tables) fits one CU.  Digits are owned
// kDPT per thread (thread t: digits t * kDPT ...), so 2048 digits need no
// more than 512 threads.
template <typename K, int THREADS, int RB>
struct SortGeom {
    static constexpr int kThreads = THREADS;
    static constexpr int kWaves = THREADS / 64;
    static constexpr int kRadixN = 1 << RB;
    static constexpr int kDPT = kRadixN > THREADS ? kRadixN / THREADS : 1;
    // per-wave digit counts (at most 64 * items, prefixes at most the tile):
    // 16 bits when the 2048-bucket tables would not fit LDS otherwise
    using CntT = typename std::conditional<(RB > 8), uint16_t, uint32_t>::type;
    static constexpr int kDigitThreads = kRadixN / kDPT;
    static constexpr size_t lds(int items) {
        return (size_t)THREADS * items * (sizeof(K) + 4) + (size_t)kWaves * kRadixN * sizeof(CntT) + (size_t)3 * kRadixN * 4 +
               64 * 4 + 64;
    }
    static constexpr int fit(int items) { return items <= 1 || lds(items) <= kLdsBytes ? items : fit(items - 1); }
    static constexpr int kItems = fit(sizeof(K) == 4 ? 16 : THREADS == 1024 ? kMaxItems64 : 8);
    static constexpr int kTile = THREADS * kItems;
    static_assert(lds(kItems) <= kLdsBytes, "onesweep geometry does not fit LDS");
    static constexpr int kLB = kLookBack / kDPT > 4 ? kLookBack / kDPT : 4;  // predecessors per step
    // threads sharing one digit's look-back when the block has more threads
    // than digits, each loading its own kLBH predecessors of every step: few
    // loads per lane return soonest (2 to 4 measured best, 8 or more slower)
    static constexpr int kH = THREADS >= kRadixN ? THREADS / kRadixN : 1;
    static constexpr int kLBH = 4;
    static constexpr int kDigitWaves = (kDigitThreads + 63) / 64;
};

// Passes of RB-bit digits over key bits [begin_bit, end_bit).
__host__ __device__ constexpr uint32_t sort_passes(uint32_t begin_bit, uint32_t end_bit, uint32_t rb) {
    return end_bit > begin_bit ? (end_bit - begin_bit + rb - 1) / rb : 0;
}

__device__ __forceinline__ uint32_t read_n(const uint32_t* n_dev, uint64_t n_host) {
    return n_dev ? *n_dev : (uint32_t)n_host;
}

// Keys per thread of a onesweep tile for n keys and `slots` resident blocks
// of `threads`: the full max_items once n fills every block's tile, else
// the shortest of the tile lengths below that still gives at most one tile
// per block.  While a pass is one round of tiles, its time is one tile's
// chain (load, rank, look-back, scatter), which shorter tiles shorten; a
// second round would cost a whole chain more.  Choice k of max_items m:
// m, 7m/8, 3m/4, m/2, m/4 (at least 1).
__host__ __device__ constexpr uint32_t onesweep_item_choice(uint32_t m, int k) {
    const uint32_t c = k == 0 ? m : k == 1 ? m * 7 / 8 : k == 2 ? m * 3 / 4 : k == 3 ? m / 2 : m / 4;
    return c < 1 ? 1u : c;
}
__host__ __device__ constexpr uint32_t onesweep_items(uint32_t n, uint32_t slots, uint32_t threads,
                                                      uint32_t max_items) {
    const uint64_t per_item = (uint64_t)slots * threads;
    const uint64_t need = per_item ? (n + per_item - 1) / per_item : max_items;
    uint32_t items = max_items;
    for (int k = 1; k <= 4; ++k)
        if (onesweep_item_choice(max_items, k) >= need) items = onesweep_item_choice(max_items, k);
    return items;
}

// Exclusive prefix of v over the first WAVES waves of the block, in thread
// order.  Every thread of the block calls it (the result is meaningful for
// those waves only); s_w holds one word per wave.
template <int WAVES>
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_w) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane == 63 && wave < (uint32_t)WAVES) s_w[wave] = incl;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w)
        if (w < (int)wave) pre += s_w[w];
    __syncthreads();
    return pre + incl - v;
}

// Exclusive prefix over the digits of a block: thread t of the first WAVES
// waves holds the DPT values of digits t * DPT + k.
template <int WAVES, int DPT>
__device__ __forceinline__ void digits_exclusive(const uint32_t (&v)[DPT], uint32_t (&out)[DPT], uint32_t* s_w) {
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        out[k] = run;
        run += v[k];
    }
    const uint32_t base = block_exclusive<WAVES>(run, s_w);
#pragma unroll
    for (int k = 0; k < DPT; ++k) out[k] += base;
}

// Zeroes the histograms and tickets of every pass (unless the caller had a
// kernel of its own do it; the histogram kernel clears the first pass's
// look-back words, and each pass its successor's).
__global__ __launch_bounds__(256) void k_sort_init(uint32_t* __restrict__ hist_tickets, size_t hist_words) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < hist_words; i += stride) hist_tickets[i] = 0;
}

// 1024-thread blocks (up to four per CU) take 1024 x kHistItems-key tiles in
// a grid-stride loop; each thread issues all kHistItems loads of a tile before
// counting, so the loop is not bound by one load's latency per key.  8-bit
// digits: waves w and w + 4k share LDS copy w % 4 of the bins; wider digits
// spread over more bins and use one copy.
template <typename K, int RB>
__global__ __launch_bounds__(1024) void k_sort_hist(const K* __restrict__ keys, const uint32_t* n_dev,
                                                   uint64_t n_host, uint32_t begin_bit, uint32_t end_bit,
                                                   uint32_t passes, uint32_t* __restrict__ hist,
                                                   uint32_t* __restrict__ status0, uint32_t slots,
                                                   uint32_t sweep_threads, uint32_t max_items, uint32_t copies) {
    constexpr int R = 1 << RB;
    constexpr int COPIES = RB <= 8 ? 4 : 1;
    constexpr int MAXP = (int)(sizeof(K) * 8 + RB - 1) / RB;
    __shared__ uint32_t s_hist[COPIES][MAXP * R];
    const uint32_t n = read_n(n_dev, n_host);
    {   // the first onesweep pass's look-back words, for the tiles of n
        const uint32_t tile_keys = onesweep_items(n, slots, sweep_threads, max_items) * sweep_threads;
        const size_t words = (size_t)ceil_div(n, tile_keys) * R, stride = (size_t)gridDim.x * blockDim.x;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) status0[i] = 0;
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, copy = wave % COPIES;
    for (uint32_t w = 0; w < COPIES; ++w)
        for (uint32_t i = threadIdx.x; i < passes * R; i += blockDim.x) s_hist[w][i] = 0;
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
                const uint32_t shift = begin_bit + p * RB;
                const uint32_t bits = min((uint32_t)RB, end_bit - shift);
                const uint32_t d = (uint32_t)(k[i] >> shift) & ((1u << bits) - 1u);
                // previous lane's digit (wave_shr:1); lane 0 always starts a run
                const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)d, 0x138, 0xF, 0xF, false);
                const bool head = ok && (lane == 0 || prev != d);
                const uint64_t heads = ballot(head);
                if (head) {
                    const uint64_t after = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
                    const uint32_t next = after ? (uint32_t)__builtin_ctzll(after) : valid;
                    atomicAdd(&s_hist[copy][p * R + d], next - lane);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < passes * R; i += blockDim.x) {
        uint32_t v = 0;
#pragma unroll
        for (int w = 0; w < COPIES; ++w) v += s_hist[w][i];
        if (v) atomicAdd(&hist[(size_t)(blockIdx.x % copies) * passes * R + i], v);
    }
}

// LDS of a onesweep block (one allocation whatever the tile length).
template <typename K, int THREADS, int RB>
struct OnesweepLds {
    using G = SortGeom<K, THREADS, RB>;
    __align__(16) typename G::CntT s_wave_cnt[G::kWaves][G::kRadixN];
    uint32_t s_global[G::kRadixN];
    uint32_t s_local[G::kRadixN];
    uint32_t s_scatter[G::kRadixN];
    K s_keys[G::kTile];
    uint32_t s_vals[G::kTile];
    uint32_t s_w[G::kWaves];
    uint32_t s_tile;
};

// The tiles of one pass, IT keys per thread (a compile-time count, so the
// item loops unroll fully; k_onesweep picks the instantiation).
template <typename K, int THREADS, int RB, int IT>
__device__ __forceinline__ void onesweep_tiles(OnesweepLds<K, THREADS, RB>& L, const K* __restrict__ kin,
                                               const uint32_t* __restrict__ vin, K* __restrict__ kout,
                                               uint32_t* __restrict__ vout, uint32_t n, uint32_t num_tiles,
                                               uint32_t shift, uint32_t mask, uint32_t* ticket,
                                               uint32_t* status, uint32_t* error) {
    using G = SortGeom<K, THREADS, RB>;
    constexpr int WAVES = G::kWaves, R = G::kRadixN, DPT = G::kDPT, LB = G::kLB, H = G::kH, LBH = G::kLBH;
    using CntT = typename G::CntT;
    constexpr uint32_t tile_keys = (uint32_t)IT * THREADS;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool digit_thread = tid < (uint32_t)G::kDigitThreads;
    const uint32_t d0 = tid * DPT;  // first digit of this thread
    for (;;) {
        uint32_t* zc = reinterpret_cast<uint32_t*>(&L.s_wave_cnt[0][0]);  // cleared as 32-bit words
        for (int i = tid; i < (int)(WAVES * R * sizeof(CntT) / 4); i += THREADS) zc[i] = 0;
        __syncthreads();
        const uint32_t tile = L.s_tile;
        if (tile >= num_tiles) break;
        const uint32_t wbase = tile * tile_keys + wave * (64 * IT);
        const uint32_t wend = min(n, wbase + 64 * IT);  // this wave's keys: [wbase, wend)
        K key[IT];
        uint32_t val[IT];
        uint32_t rank[IT];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            const bool ok = p < wend;
            key[i] = ok ? kin[p] : (K)0;
            val[i] = ok ? vin[p] : 0u;
        }
        // Stable per-wave ranks: items in position order, lanes in lane order.
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            const bool ok = p < n;
            const uint32_t d = (uint32_t)(key[i] >> shift) & mask;
            // peers: the lanes whose digit equals this lane's, one ballot per
            // digit bit, unrolled over the digit width (a narrower last pass
            // has zero high bits everywhere, which keep every lane): keep the
            // lanes that agree, m &= ~(bb ^ bitv) with bitv the lane's bit
            // sign-extended to a word (one bit-field extract; the ballot is a
            // compare of it; the update one three-input bit operation per
            // half of the mask)
            const uint64_t okm = ballot(ok);
            uint32_t m_lo = (uint32_t)okm, m_hi = (uint32_t)(okm >> 32);
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const uint32_t bitv = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
                const uint64_t bb = ballot(bitv != 0u);
                m_lo &= ~((uint32_t)bb ^ bitv);
                m_hi &= ~((uint32_t)(bb >> 32) ^ bitv);
            }
            const uint64_t m = (uint64_t)m_hi << 32 | m_lo;
            const uint32_t below = count_below(m);
            uint32_t prior = 0;
            if (ok) prior = L.s_wave_cnt[wave][d];
            rank[i] = prior + below;
            if (ok && below == 0) L.s_wave_cnt[wave][d] = (CntT)(prior + (uint32_t)__popcll(m));
        }
        __syncthreads();
        // Per digit: wave prefixes, tile count, publication, look-back.
        uint32_t cnt[DPT];
        uint32_t* my_status = status + (size_t)tile * R + d0;
#pragma unroll
        for (int k = 0; k < DPT; ++k) {
            cnt[k] = 0;
            if (digit_thread) {
#pragma unroll
                for (int w = 0; w < WAVES; ++w) {
                    const uint32_t c = L.s_wave_cnt[w][d0 + k];
                    L.s_wave_cnt[w][d0 + k] = (CntT)cnt[k];
                    cnt[k] += c;
                }
                st_release_u32(my_status + k, (tile == 0 ? kInc : kAgg) | cnt[k]);
            }
        }
        uint32_t local[DPT];
        digits_exclusive<G::kDigitWaves, DPT>(cnt, local, L.s_w);
        if constexpr (H > 1) {
            // H adjacent lanes per digit: lane q of digit d loads predecessors
            // q * LBH + 1 .. (q + 1) * LBH back of each step of H * LBH (one
            // round trip), sums them nearest first up to the first inclusive
            // word, and the H partial results combine in order.  L.s_scatter
            // holds the digits' tile counts until the offsets replace them.
            if (digit_thread) {
                L.s_local[d0] = local[0];
                L.s_scatter[d0] = cnt[0];
            }
            __syncthreads();
            const uint32_t d = tid / H, q = tid % H;
            uint32_t ex = 0;
            if (tile > 0) {
                int64_t j = (int64_t)tile - 1 - (int64_t)q * LBH;  // this lane's nearest predecessor
                uint32_t spins = 0;
                for (;;) {
                    uint32_t w[LBH];
#pragma unroll
                    for (int t = 0; t < LBH; ++t)
                        w[t] = (j - t >= 0) ? ld_acquire_u32(status + (size_t)(j - t) * R + d) : kInc;
                    uint32_t sum = 0, found = 0;
#pragma unroll
                    for (int t = 0; t < LBH; ++t) {
                        if (found) continue;
                        uint32_t s = w[t];
                        while ((s >> 30) == 0) {
                            if (++spins > kSpinLimit) {
                                atomicOr(error, 2u);
                                s = kInc;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                            s = ld_acquire_u32(status + (size_t)(j - t) * R + d);
                        }
                        sum += s & kStatusValueMask;
                        found = (s & kInc) ? 1u : 0u;
                    }
                    // ordered combine: the nearer part's sum ends at its
                    // inclusive word, else the farther part's sum adds on
#pragma unroll
                    for (int o = 1; o < H; o <<= 1) {
                        const uint32_t ps = __shfl_xor(sum, o, H), pf = __shfl_xor(found, o, H);
                        const bool near = (q & (uint32_t)o) == 0;
                        const uint32_t nf = near ? found : pf;
                        sum = nf ? (near ? sum : ps) : sum + ps;
                        found |= pf;
                    }
                    ex += sum;
                    if (found) break;
                    j -= (int64_t)H * LBH;
                }
                if (q == 0) st_release_u32(status + (size_t)tile * R + d, kInc | ((ex + L.s_scatter[d]) & kStatusValueMask));
            }
            if (q == 0) L.s_scatter[d] = L.s_global[d] + ex - L.s_local[d];
        }
        uint32_t excl[DPT];
#pragma unroll
        for (int k = 0; k < DPT; ++k) excl[k] = 0;
        if (H == 1 && digit_thread && tile > 0) {
            // Batched look-back: the words of up to LB predecessors (all of
            // this thread's digits) are loaded together (independent loads,
            // one round trip), then consumed nearest first; only a word still
            // not ready is re-polled.
            int64_t j = (int64_t)tile - 1;
            uint32_t spins = 0;
            uint32_t open = (1u << DPT) - 1u;  // digits still without an inclusive prefix
            while (open != 0 && j >= 0) {
                uint32_t w[LB][DPT];
#pragma unroll
                for (int t = 0; t < LB; ++t)
#pragma unroll
                    for (int k = 0; k < DPT; ++k)
                        w[t][k] = (j - t >= 0) ? ld_acquire_u32(status + (size_t)(j - t) * R + d0 + k) : kInc;
#pragma unroll
                for (int t = 0; t < LB; ++t)
#pragma unroll
                    for (int k = 0; k < DPT; ++k) {
                        if (!((open >> k) & 1u) || j - t < 0) continue;
                        uint32_t s = w[t][k];
                        while ((s >> 30) == 0) {
                            if (++spins > kSpinLimit) {
                                atomicOr(error, 2u);
                                s = kInc;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                            s = ld_acquire_u32(status + (size_t)(j - t) * R + d0 + k);
                        }
                        excl[k] += s & kStatusValueMask;
                        if (s & kInc) open &= ~(1u << k);
                    }
                j -= LB;
            }
#pragma unroll
            for (int k = 0; k < DPT; ++k)
                st_release_u32(my_status + k, kInc | ((excl[k] + cnt[k]) & kStatusValueMask));
        }
        if (H == 1 && digit_thread) {
#pragma unroll
            for (int k = 0; k < DPT; ++k) {
                L.s_local[d0 + k] = local[k];
                L.s_scatter[d0 + k] = L.s_global[d0 + k] + excl[k] - local[k];
            }
        }
        __syncthreads();
        // Digit-sorted staging of the tile in LDS.
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t p = wbase + i * 64 + lane;
            if (p < n) {
                const uint32_t d = (uint32_t)(key[i] >> shift) & mask;
                const uint32_t lp = L.s_local[d] + L.s_wave_cnt[wave][d] + rank[i];
                if (lp < tile_keys) {
                    L.s_keys[lp] = key[i];
                    L.s_vals[lp] = val[i];
                }
            }
        }
        __syncthreads();
        const uint32_t tile_n = min(tile_keys, n - tile * tile_keys);
        for (uint32_t idx = tid; idx < tile_n; idx += THREADS) {
            const K k = L.s_keys[idx];
            const uint32_t d = (uint32_t)(k >> shift) & mask;
            const uint32_t o = L.s_scatter[d] + idx;
            if (o < n) {  // always true for a consistent histogram; guards the buffer otherwise
                kout[o] = k;
                vout[o] = L.s_vals[idx];
            }
        }
        // The next ticket is taken only now: a ticket held while this tile is
        // still in flight would hold up that tile's successors.
        if (tid == 0) L.s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
    }
}

// One pass.  status: this pass's look-back words (tile * R + digit), zero
// for every tile of the key count; clear_next: the next pass's words, which
// this pass zeroes (null on the last pass).
template <typename K, int THREADS, int RB>
__global__ __launch_bounds__(THREADS) void k_onesweep(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* n_dev, uint64_t n_host, uint32_t shift, uint32_t bits,
    const uint32_t* __restrict__ hist, uint32_t hist_copies, uint32_t hist_stride, uint32_t* ticket,
    uint32_t* status, uint32_t* clear_next,
    uint32_t* error, uint32_t slots) {
    using G = SortGeom<K, THREADS, RB>;
    constexpr int ITEMS = G::kItems, R = G::kRadixN, DPT = G::kDPT;
    static_assert(G::kH == 1 || DPT == 1, "shared look-back needs one digit per owner thread");
    __shared__ OnesweepLds<K, THREADS, RB> L;

    const uint32_t tid = threadIdx.x;
    const bool digit_thread = tid < (uint32_t)G::kDigitThreads;
    const uint32_t d0 = tid * DPT;  // first digit of this thread
    const uint32_t n = read_n(n_dev, n_host);
    const uint32_t items = onesweep_items(n, slots, THREADS, ITEMS);  // uniform
    const uint32_t tile_keys = items * THREADS;
    const uint32_t num_tiles = ceil_div(n, tile_keys);
    const uint32_t mask = (1u << bits) - 1u;
    if (clear_next) {
        const size_t words = (size_t)num_tiles * R, stride = (size_t)gridDim.x * THREADS;
        for (size_t i = (size_t)blockIdx.x * THREADS + tid; i < words; i += stride) clear_next[i] = 0;
    }
    if (blockIdx.x >= num_tiles) return;  // uniform per block
    if (tid == 0) L.s_tile = atomicAdd(ticket, 1u);
    {
        uint32_t h[DPT], g[DPT];
#pragma unroll
        for (int k = 0; k < DPT; ++k) {
            uint32_t s = 0;
            if (digit_thread)
                for (uint32_t cp = 0; cp < hist_copies; ++cp) s += hist[(size_t)cp * hist_stride + d0 + k];
            h[k] = s;
        }
        digits_exclusive<G::kDigitWaves, DPT>(h, g, L.s_w);
        if (digit_thread)
#pragma unroll
            for (int k = 0; k < DPT; ++k) L.s_global[d0 + k] = g[k];
    }

    // one instantiation per tile length onesweep_items can pick
    constexpr int I1 = onesweep_item_choice(ITEMS, 1), I2 = onesweep_item_choice(ITEMS, 2),
                  I3 = onesweep_item_choice(ITEMS, 3), I4 = onesweep_item_choice(ITEMS, 4);
    if (items == (uint32_t)ITEMS)
        onesweep_tiles<K, THREADS, RB, ITEMS>(L, kin, vin, kout, vout, n, num_tiles, shift, mask, ticket, status, error);
    else if (items == (uint32_t)I1)
        onesweep_tiles<K, THREADS, RB, I1>(L, kin, vin, kout, vout, n, num_tiles, shift, mask, ticket, status, error);
    else if (items == (uint32_t)I2)
        onesweep_tiles<K, THREADS, RB, I2>(L, kin, vin, kout, vout, n, num_tiles, shift, mask, ticket, status, error);
    else if (items == (uint32_t)I3)
        onesweep_tiles<K, THREADS, RB, I3>(L, kin, vin, kout, vout, n, num_tiles, shift, mask, ticket, status, error);
    else
        onesweep_tiles<K, THREADS, RB, I4>(L, kin, vin, kout, vout, n, num_tiles, shift, mask, ticket, status, error);
}

// Words of sort scratch for a geometry: the histograms (copies of them) and
// tickets, and two sets of look-back words.  A pass has at most max(resident
// blocks, full tiles of the capacity) tiles: the tile is shortened to give
// each resident block one tile while the keys do not fill them all
// (onesweep_items).
template <typename K, int THREADS, int RB>
SortWords sort_words_of(const Context* c, uint64_t capacity, uint32_t begin_bit, uint32_t end_bit,
                        uint32_t copies) {
    using G = SortGeom<K, THREADS, RB>;
    const uint32_t slots = (uint32_t)c->cu_count * (1024 / THREADS);
    const uint64_t max_tiles = std::max<uint64_t>(slots, (capacity + G::kTile - 1) / G::kTile);
    return SortWords{(size_t)sort_passes(begin_bit, end_bit, RB) * G::kRadixN * copies + kMaxPasses,
                     (size_t)2 * max_tiles * G::kRadixN};
}

template <typename K, int THREADS, int RB>
int sort_impl(Context* c, K* k0, uint32_t* v0, K* k1, uint32_t* v1, const uint32_t* n_dev, uint64_t n_host,
              uint64_t capacity, uint32_t begin_bit, uint32_t end_bit, FrameCounters* counters,
              int* result_index, bool hist_cleared, const SortMem* mem) {
    using G = SortGeom<K, THREADS, RB>;
    constexpr uint32_t R = G::kRadixN;
    const uint32_t passes = sort_passes(begin_bit, end_bit, RB);
    *result_index = (int)(passes & 1u);
    if (passes == 0) return GSR_OK;
    if (passes > (uint32_t)kMaxPasses) return fail(c, GSR_ERR_ARGUMENT, "too many radix passes (%u)", passes);
    const uint32_t slots = (uint32_t)c->cu_count * (1024 / THREADS);
    const uint32_t copies = mem ? 1u : hist_copies<K>();
    const SortWords words = sort_words_of<K, THREADS, RB>(c, capacity, begin_bit, end_bit, copies);
    const size_t hist_words = words.hist;
    const uint32_t max_tiles = (uint32_t)(words.status / (2 * R));
    uint32_t* hist;
    uint32_t* status;   // two sets of look-back words, alternating between passes
    if (mem) {
        if (mem->hist_words < words.hist || mem->status_words < words.status)
            return fail(c, GSR_ERR_ARGUMENT, "sort scratch too small: %zu + %zu words needed, %zu + %zu given",
                        words.hist, words.status, mem->hist_words, mem->status_words);
        hist = mem->hist;
        status = mem->status;
    } else {
        hist = (uint32_t*)scratch(c, "sort_hist", hist_words * 4);
        status = (uint32_t*)scratch(c, "sort_status", words.status * 4);
    }
    if (!hist || !status) return GSR_ERR_ALLOC;
    uint32_t* tickets = hist + (size_t)passes * R * copies;
    if (!hist_cleared)
        hipLaunchKernelGGL(k_sort_init, dim3((uint32_t)std::min<size_t>((hist_words + 255) / 256, 64)), dim3(256), 0,
                           c->stream, hist, hist_words);
    // Up to 4 blocks per CU: the key loads, not the per-bin global atomics of
    // every block, set the kernel's time (1080p frame sort 0.1186 -> 0.1163 ms
    // against one 16-item block per CU; fewer blocks were slower still).
    const uint32_t hist_blocks = std::min<uint32_t>(
        std::max<uint32_t>(1, ceil_div(capacity, (uint64_t)1024 * kHistItems)), (uint32_t)c->cu_count * 4);
    hipLaunchKernelGGL((k_sort_hist<K, RB>), dim3(hist_blocks), dim3(1024), 0, c->stream, k0, n_dev, n_host,
                       begin_bit, end_bit, passes, hist, status, slots, (uint32_t)THREADS, (uint32_t)G::kItems,
                       copies);
    K* kb[2] = {k0, k1};
    uint32_t* vb[2] = {v0, v1};
    uint32_t* st[2] = {status, status + (size_t)max_tiles * R};
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = begin_bit + p * RB;
        const uint32_t bits = std::min<uint32_t>(RB, end_bit - shift);
        hipLaunchKernelGGL((k_onesweep<K, THREADS, RB>), dim3(slots), dim3(THREADS), 0, c->stream, kb[p & 1],
                           vb[p & 1], kb[(p + 1) & 1], vb[(p + 1) & 1], n_dev, n_host, shift, bits, hist + p * R,
                           copies, passes * R, tickets + p, st[p & 1], p + 1 < passes ? st[(p + 1) & 1] : nullptr,
                           &counters->error, slots);
    }
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

// Onesweep blocks per CU: 1, 2 or 4 (1024-, 512- or 256-thread blocks).
template <typename K, int RB>
int sort_geometry(Context* c, K* k0, uint32_t* v0, K* k1, uint32_t* v1, const uint32_t* n_dev, uint64_t n_host,
                  uint64_t capacity, uint32_t begin_bit, uint32_t end_bit, uint32_t blocks_per_cu,
                  FrameCounters* counters, int* result_index, bool hist_cleared, const SortMem* mem) {
    switch (blocks_per_cu) {
        case 1:
            return sort_impl<K, 1024, RB>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, counters,
                                          result_index, hist_cleared, mem);
        case 2:
            return sort_impl<K, 512, RB>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, counters,
                                          result_index, hist_cleared, mem);
        case 4:
            return sort_impl<K, 256, RB>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, counters,
                                          result_index, hist_cleared, mem);
        default:
            return fail(c, GSR_ERR_ARGUMENT, "sort blocks per CU %u not in {1, 2, 4}", blocks_per_cu);
    }
}

template <typename K>
int sort_digits(Context* c, K* k0, uint32_t* v0, K* k1, uint32_t* v1, const uint32_t* n_dev, uint64_t n_host,
                uint64_t capacity, uint32_t begin_bit, uint32_t end_bit, uint32_t blocks_per_cu, uint32_t digit_bits,
                FrameCounters* counters, int* result_index, bool hist_cleared, const SortMem* mem) {
    switch (digit_bits) {
        case 8:
            return sort_geometry<K, 8>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, blocks_per_cu,
                                       counters, result_index, hist_cleared, mem);
        case 9:
            return sort_geometry<K, 9>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit, blocks_per_cu,
                                       counters, result_index, hist_cleared, mem);
        case 10:
            return sort_geometry<K, 10>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit,
                                        blocks_per_cu, counters, result_index, hist_cleared, mem);
        case 11:
            return sort_geometry<K, 11>(c, k0, v0, k1, v1, n_dev, n_host, capacity, begin_bit, end_bit,
                                        blocks_per_cu, counters, result_index, hist_cleared, mem);
        default:
            return fail(c, GSR_ERR_ARGUMENT, "sort digit width %u not in 8..11", digit_bits);
    }
}

}  // namespace

uint32_t default_digit_bits(uint32_t key_bytes, uint32_t end_bit) {
    if (key_bytes != 8) return 8;  // wider digits would halve the 32-bit key tile (LDS)
    // 64-bit keys keep 9216-key or longer tiles up to 10-bit digits: the
    // narrowest width with the fewest passes
    uint32_t best = 8;
    for (uint32_t rb = 9; rb <= 10; ++rb)
        if (sort_passes(0, end_bit, rb) < sort_passes(0, end_bit, best)) best = rb;
    return best;
}

uint32_t sort_result_index(uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits) {
    return sort_passes(begin_bit, end_bit, digit_bits) & 1u;
}

FillRegion sort_hist_region(Context* c, uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits) {
    const size_t words =
        (size_t)sort_passes(begin_bit, end_bit, digit_bits) * (1u << digit_bits) * kHistCopies + kMaxPasses;
    uint32_t* p = (uint32_t*)scratch(c, "sort_hist", words * 4);
    return FillRegion{p, p ? words : 0};
}

uint32_t sort_default_tile_keys(uint32_t key_bytes) {
    return key_bytes == 8 ? (uint32_t)SortGeom<uint64_t, 1024, 8>::kTile : (uint32_t)SortGeom<uint32_t, 1024, 8>::kTile;
}

SortWords sort_scratch_words(const Context* c, uint32_t key_bytes, uint64_t capacity, uint32_t begin_bit,
                             uint32_t end_bit, uint32_t blocks_per_cu, uint32_t digit_bits) {
    if (digit_bits == 8 && blocks_per_cu == 1)
        return key_bytes == 8 ? sort_words_of<uint64_t, 1024, 8>(c, capacity, begin_bit, end_bit, 1)
                              : sort_words_of<uint32_t, 1024, 8>(c, capacity, begin_bit, end_bit, 1);
    return SortWords{~(size_t)0, ~(size_t)0};  // only the default geometry is sized here
}

int launch_sort(Context* c, uint32_t key_bytes, void* k0, uint32_t* v0, void* k1, uint32_t* v1,
                const uint32_t* n_dev, uint64_t n_host, uint64_t capacity, uint32_t begin_bit,
                uint32_t end_bit, uint32_t blocks_per_cu, FrameCounters* counters, int* result_index,
                uint32_t digit_bits, bool hist_cleared, const SortMem* mem) {
    if (capacity > kMaxSortKeys) return fail(c, GSR_ERR_ARGUMENT, "sort capacity %llu exceeds %llu",
                                             (unsigned long long)capacity, (unsigned long long)kMaxSortKeys);
    if (end_bit > key_bytes * 8 || begin_bit > end_bit)
        return fail(c, GSR_ERR_ARGUMENT, "bad bit range [%u, %u) for %u-byte keys", begin_bit, end_bit, key_bytes);
    if (key_bytes == 8)
        return sort_digits<uint64_t>(c, (uint64_t*)k0, v0, (uint64_t*)k1, v1, n_dev, n_host, capacity, begin_bit,
                                     end_bit, blocks_per_cu, digit_bits, counters, result_index, hist_cleared, mem);
    return sort_digits<uint32_t>(c, (uint32_t*)k0, v0, (uint32_t*)k1, v1, n_dev, n_host, capacity, begin_bit,
                                 end_bit, blocks_per_cu, digit_bits, counters, result_index, hist_cleared, mem);
}

}  // namespace gsr
