// scan.hip — exclusive scan of the per-splat tile counts into instance
// offsets (src/shaders/offset_scan.wgsl:22-68), plus the device-side sizing
// of src/shaders/indirect_arg.wgsl:14-28 (n = min(I, capacity)).
//
// Single pass with decoupled look-back: 1024-thread blocks take 8192-entry
// tiles in ticket order; each thread sums 8 consecutive entries (two 16-byte
// loads when the arrays are 16-byte aligned), waves scan with DPP, and the
// first wave of each block chains the tile prefix through 64-bit status
// words (flag in bits 62-63, running total in bits 0-61), so totals beyond
// 2^32 are detected rather than wrapped.  (1080p, 1M entries: 16384-entry
// tiles 10.7 us, 8192 9.7, 4096 11.7.)
#include "gsr_internal.h"
#include "wave.h"

namespace gsr {
namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint64_t kFlagAgg = 1ull << 62, kFlagInc = 2ull << 62, kValMask = (1ull << 62) - 1;
constexpr int kLookBack = 16;

// counts[0..n) -> out[0..n], out[n] = total.  status[0] is the tile ticket,
// status[1 + t] the look-back word of tile t; all zeroed before the launch.
// VEC: in and out are 16-byte aligned (vector loads and stores away from the
// tail).
template <bool VEC>
__global__ __launch_bounds__(kScanThreads) void k_scan(const uint32_t* in, uint32_t n,
                                                       uint32_t* out,  // may alias in
                                                       uint64_t* __restrict__ status_base,
                                                       FrameCounters* counters, uint64_t capacity) {
    constexpr int WAVES = kScanThreads / 64;
    uint64_t* status = status_base + 1;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[WAVES];
    __shared__ uint64_t s_prefix;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd((uint32_t*)status_base, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t total_len = (uint64_t)n + 1;  // entry n is the virtual 0 giving the total
    const uint32_t num_tiles = ceil_div(total_len, kScanTile);
    const uint64_t base = (uint64_t)tile * kScanTile + (uint64_t)tid * kScanItems;
    uint32_t v[kScanItems];
    if (VEC && base + kScanItems <= n) {
        const uint4* in4 = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const uint4 t = in4[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) {
            const uint64_t p = base + i;
            v[i] = p < n ? in[p] : 0u;
        }
    }
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) sum += v[i];
    const uint32_t incl = wave_inclusive_sum(sum);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_prefix = 0, block_total = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const uint32_t t = s_wave[w];
        if (w < (int)wave) wave_prefix += t;
        block_total += t;
    }
    if (tile == 0) {
        if (tid == 0) {
            st_release_u64(&status[0], kFlagInc | block_total);
            s_prefix = 0;
        }
    } else if (wave == 0) {
        if (lane == 0) st_release_u64(&status[tile], kFlagAgg | block_total);
        // Look-back by the whole first wave: lane l polls predecessor
        // j - l, 64 per round trip.  The exclusive prefix sums the words up
        // to the nearest inclusive one; a round is used only once every word
        // it needs is published (the others are polled again).
        uint64_t excl = 0;
        int64_t j = (int64_t)tile - 1;
        uint32_t spins = 0;
        bool done = false;
        while (!done) {
            const int64_t p = j - (int64_t)lane;
            uint64_t s = p >= 0 ? ld_acquire_u64(&status[p]) : kFlagInc;  // before tile 0: an inclusive 0
            for (;;) {
                const uint64_t inc = ballot((s >> 62) == 2u);
                const uint32_t first = inc ? (uint32_t)__builtin_ctzll(inc) : 63u;
                const uint64_t need = (2ull << first) - 1ull;  // lanes 0 .. first (all 64 for 63)
                if ((ballot((s >> 62) != 0u) & need) == need) {
                    uint64_t v = lane <= first ? (s & kValMask) : 0ull;
#pragma unroll
                    for (int m = 32; m >= 1; m >>= 1) v += (uint64_t)__shfl_xor((long long)v, m, 64);
                    excl += v;
                    done = inc != 0;
                    break;
                }
                if (++spins > kSpinLimit) {
                    if (lane == 0) atomicOr(&counters->error, 1u);
                    done = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                if ((s >> 62) == 0u) s = ld_acquire_u64(&status[p]);  // p >= 0: the virtual word is published
            }
            j -= 64;
        }
        if (lane == 0) {
            st_release_u64(&status[tile], kFlagInc | ((excl + block_total) & kValMask));
            s_prefix = excl;
        }
    }
    if (tid == 0 && tile == num_tiles - 1) {
        // (the last tile's s_prefix is written above by this same thread)
        const uint64_t total = s_prefix + block_total;
        counters->instances = total;
        counters->n = (uint32_t)(total < capacity ? total : capacity);
    }
    __syncthreads();
    uint64_t run = s_prefix + wave_prefix + (incl - sum);
    uint32_t o[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        o[i] = (uint32_t)run;
        run += v[i];
    }
    if (VEC && base + kScanItems <= (uint64_t)n + 1) {
        uint4* out4 = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) out4[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i)
            if (base + i <= n) out[base + i] = o[i];
    }
}

}  // namespace

size_t scan_status_words(uint32_t n) { return ((size_t)ceil_div((uint64_t)n + 1, kScanTile) + 1) * 2; }

uint32_t* scan_status_buffer(Context* c, uint32_t n) {
    return (uint32_t*)scratch(c, "scan_status", scan_status_words(n) * 4);
}

int launch_scan(Context* c, const uint32_t* in, uint32_t n, uint32_t* out, FrameCounters* counters,
                uint64_t capacity, bool status_cleared) {
    const uint32_t tiles = ceil_div((uint64_t)n + 1, kScanTile);
    uint64_t* status = (uint64_t*)scan_status_buffer(c, n);
    if (!status) return GSR_ERR_ALLOC;
    if (!status_cleared) {
        const int rc = launch_fill_u32(c, (uint32_t*)status, 0u, scan_status_words(n));
        if (rc) return rc;
    }
    const bool vec = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 16 == 0);
    if (vec)
        hipLaunchKernelGGL(k_scan<true>, dim3(tiles), dim3(kScanThreads), 0, c->stream, in, n, out, status, counters,
                           capacity);
    else
        hipLaunchKernelGGL(k_scan<false>, dim3(tiles), dim3(kScanThreads), 0, c->stream, in, n, out, status,
                           counters, capacity);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

}  // namespace gsr
