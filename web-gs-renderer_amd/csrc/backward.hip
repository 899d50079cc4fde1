// backward.hip — adjoint of the forward frame (the reference is forward-only;
// this differentiates src/shaders/raseter.wgsl:35-117 and
// src/shaders/preprocess.wgsl:58-284).
//
// k_raster_backward: one wave64 per 8x8 pixel block of a tile, one pixel per
//   lane, as in the forward.  Each pixel replays its splats back to front,
//   starting at its stored last contributor with its stored final
//   transmittance; alphas are recomputed with the forward's exact FP32 steps
//   (so the blend decisions are the forward's), T is recovered as
//   T * rcp(1 - alpha), and the light arriving from behind each splat is
//   accumulated on the way (as one dL/dRGB-weighted sum per pixel).  Each
//   lane stores its pixel's dL/dpower and blend weight per blended splat in
//   the wave's LDS; every few splats the wave sums them per splat over pixel
//   rows (moments of dL/dpower, opacity and colour terms) and adds the nine
//   sums to the splat's 2-D gradient with float atomics.
// k_preprocess_backward: one thread per splat, turns the raster's sums into
//   the 2-D gradient (pixel position, conic, opacity, colour) and chains it
//   back to the 59 parameters.
//
// 2-D gradient layout per splat as returned in grad2d_out (12 floats): d/dx,
// d/dy of the pixel position; d/dA, d/dB, d/dC of the conic; d/d opacity;
// d/d(r, g, b); pad.  The raster's sums that it is made from take 9 floats
// per entry (no pad).
#include <cstddef>
#include <cstring>

#include "cull.h"
#include "gsr_internal.h"
#include "wave.h"

namespace gsr {
namespace {

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, m, 64));
    return v;
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Splats whose per-pixel terms a wave collects before it sums them per
// splat (see backward_block).
constexpr int kBwdChunk = 4;

// LDS of one wave of backward_block (the workgroup holds 4; 5.5 KiB each, so
// that seven workgroups, 28 waves, fit a CU: with the staged records padded
// to 48 bytes (6 KiB per wave) six fitted, and with 8 KiB per wave and 8
// collected splats five).  The staged batch is three arrays (conic and
// opacity; colour with the splat index in w; pixel position), each read
// back with one same-address load.  pair[c] holds collected splat c's terms
// with pixel p at p ^ 8c, so that the flush's reads of different splats fall
// in different banks.
struct BwdWaveLds {
    float4 co[64];                  // the staged batch: -A/2, -B, -C/2 and opacity
    float4 cl[64];                  //   colour, splat index bits
    float2 p[64];                   //   pixel position
    float2 pair[kBwdChunk][64];     // per collected splat and pixel: (gpow, w)
    float4 g[64];                   // per pixel: dL/d(r, g, b)
};
static_assert(sizeof(BwdWaveLds) == 5632, "backward LDS per wave");

// Which of the nine per-splat totals lane l16 = lane % 16 of a splat's
// 16-lane group holds after backward_flush's reduce-scatter (4 bits per
// lane, 0xF: a duplicate, not added).
constexpr uint64_t kFlushTotalOfLane = 0xFFF8F765FF43F210ull;

// Floats per 2-D gradient-sum entry (one per splat, or per (view, splat) in
// a batch): the five position/conic moments, the opacity term and the
// three colour terms, packed without padding.
constexpr uint32_t kSumFloats = 9;

// Sums the collected splats' terms per splat and adds them to the splats'
// 2-D gradient sums.  Lane l works on collected splat c = l / 16, block
// column s = l % 8 and half h = (l / 8) % 2 of the rows (pixels s + 8 i for
// i = 4h .. 4h + 3: dx is constant per lane); the 16 lanes of a splat then
// combine their nine partial sums by a reduce-scatter of four DPP steps
// (partners l ^ 8, l ^ 7, l ^ 2, l ^ 1): at each step a lane keeps about
// half of its values, adding the partner's copy of them, so the step
// costs about as many adds as the lane keeps values (13, 7, 4 and 3
// instructions with the selects, against 9 per step for a full reduction of
// every value).  Nine lanes end with one total each and add it with one
// atomic instruction, whose lanes address one splat's entry each (the
// atomics' cost is per addressed cache line).
// slots: 6 bits per collected splat, its slot in the staged batch.
__device__ __forceinline__ void backward_flush(BwdWaveLds& L, uint32_t n, uint64_t slots, uint32_t x0,
                                               uint32_t y0, uint32_t splats, float* __restrict__ grad2d) {
    static_assert(kBwdChunk == 4, "one splat per 16 lanes");
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t c = lane >> 4, half = (lane >> 3) & 1u, col = lane & 7;
    // the collected terms were written by every lane of the wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float sg = 0.f, sy = 0.f, syy = 0.f, sr = 0.f, sgg = 0.f, sb = 0.f, dx = 0.f;
    float op = 1.f;
    uint32_t gid = 0xFFFFFFFFu;
    if (c < n) {
        const uint32_t sl = (uint32_t)(slots >> (6 * c)) & 63u;
        const float2 sp = L.p[sl];
        const float ix = sp.x, iy = sp.y;
        op = L.co[sl].w;
        gid = __float_as_uint(L.cl[sl].w);
        dx = ix - (float)(x0 + col);
        const float fy0 = (float)y0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = half * 4 + k;
            const float2 pr = L.pair[c][col + 8 * (i ^ c)];  // pixel col + 8 i
            const float4 gv = L.g[col + 8 * i];
            const float dy = iy - (fy0 + (float)i);
            const float gy = pr.x * dy;
            sg += pr.x;
            sy += gy;
            syy = __builtin_fmaf(gy, dy, syy);
            sr = __builtin_fmaf(pr.y, gv.x, sr);
            sgg = __builtin_fmaf(pr.y, gv.y, sgg);
            sb = __builtin_fmaf(pr.y, gv.z, sb);
        }
    }
    // moments with the lane's constant dx: sum g dx = dx sum g, and so on;
    // opacity: sum of e gal over the pixels off the 0.99 clamp = sum gpow / op
    // (v_rcp_f32, 1 ulp, instead of an IEEE division sequence)
    const float v[9] = {dx * sg, sy, dx * dx * sg, dx * sy, syy, sg * __builtin_amdgcn_rcpf(op), sr, sgg, sb};
    const bool b0 = col & 1, b1 = col & 2, b2 = col & 4;
    // partner l ^ 8 (row_ror:8 within the 16-lane row): the first half keeps
    // totals 0-4, the second 5-8
    float w[5];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float keep = half ? v[k + 5] : v[k], send = half ? v[k] : v[k + 5];
        w[k] = keep + dpp_f32<0x128>(send);
    }
    w[4] = v[4] + dpp_f32<0x128>(v[4]);
    // partner l ^ 7 (row_half_mirror): columns 0-3 keep w0-w2, 4-7 w3, w4
    float x[3];
    x[0] = (b2 ? w[3] : w[0]) + dpp_f32<0x141>(b2 ? w[0] : w[3]);
    x[1] = (b2 ? w[4] : w[1]) + dpp_f32<0x141>(b2 ? w[1] : w[4]);
    x[2] = w[2] + dpp_f32<0x141>(w[2]);
    // partner l ^ 2 (quad_perm [2,3,0,1])
    const float y0v = (b1 ? x[2] : x[0]) + dpp_f32<0x4E>(b1 ? x[0] : x[2]);
    const float y1v = x[1] + dpp_f32<0x4E>(x[1]);
    // partner l ^ 1 (quad_perm [1,0,3,2])
    const float mine = (b0 ? y1v : y0v) + dpp_f32<0xB1>(b0 ? y0v : y1v);
    // the totals are formed by every lane before the branch below: sunk into
    // it, the last DPP step could not fold into its add
    asm volatile("" ::"v"(mine));
    // the collected slots are reused by the next chunk
    __builtin_amdgcn_wave_barrier();
    const uint32_t slot = (uint32_t)(kFlushTotalOfLane >> (4 * (lane & 15))) & 0xFu;
    if (c < n && gid < splats && slot != 0xFu) {
        float* dst = grad2d + (size_t)gid * kSumFloats;
        unsafeAtomicAdd(dst + slot, mine);
    }
}

// One 8x8 block (pixels x0.. x0 + 7, y0.. y0 + 7 of an image of W x H whose
// planes start at tfinal / last / dLdC) replaying the tile's instances
// [start, min(range_end, start + max last contributor)) back to front, one
// pixel per lane, the batch of 64 splats staged in the wave's LDS and read
// back with same-address loads.  For every splat that some pixel blended,
// each lane stores its pixel's gpow = dL/dpower and w = alpha T into the
// wave's chunk buffer; every kBwdChunk such splats (and at the end of each
// batch, whose staged records the flush reads) backward_flush turns
// the buffer into per-splat sums: the moments of
// gpow (g dx, g dy, g dx^2, g dx dy, g dy^2), the opacity term and the colour
// terms, added to grad2d per entry of vals.  Summing per splat over block
// columns replaces a 64-lane reduction per splat and block by a few adds per
// pixel and a four-step DPP reduce-scatter of the nine sums over 16 lanes.
__device__ __forceinline__ void backward_block(uint32_t W, uint32_t H, uint32_t x0, uint32_t y0, uint32_t start,
                                               uint32_t range_end, uint32_t splats, const uint32_t* __restrict__ vals,
                                               const SplatRec* __restrict__ rec, const float* __restrict__ tfinal,
                                               const uint32_t* __restrict__ last, const float* __restrict__ dLdC,
                                               float* __restrict__ grad2d) {
    __shared__ BwdWaveLds s_bwd[4];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t x = x0 + lane % 8, y = y0 + lane / 8;
    const float fx = (float)x, fy = (float)y;
    const size_t plane = (size_t)W * H;
    // GS: sum over channels of dL/dRGB times the colour the splats behind the
    // current one contribute (unnormalised: each c_j alpha_j T_j), so that
    // dL/dalpha-type terms T_i (c_i - behind_i / T_{i+1}) . g collapse to
    // T_b (c . g) - GS / (1 - alpha)
    float T = 1.0f, gr = 0.0f, gg = 0.0f, gb = 0.0f, GS = 0.0f;
    uint32_t lst = 0;
    if (x < W && y < H) {
        const size_t o = (size_t)y * W + x;
        T = tfinal[o];
        lst = last[o];
        gr = dLdC[o];
        gg = dLdC[plane + o];
        gb = dLdC[2 * plane + o];
    }
    const uint32_t maxlast = wave_max_u32(lst);
    if (maxlast == 0 || splats == 0) return;  // nothing blended (no record to read)
    const uint32_t stop = min(range_end, start + maxlast);  // exclusive end of the replay
    if (stop <= start) return;
    const float rxa = (float)x0, rxb = (float)min(x0 + 7, W - 1);
    const float rya = (float)y0, ryb = (float)min(y0 + 7, H - 1);
    // the wave's LDS as a wave-uniform (scalar) index, so that a record's
    // LDS address is scalar arithmetic
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    BwdWaveLds& L = s_bwd[wv];
    L.g[lane] = make_float4(gr, gg, gb, 0.0f);
    uint32_t n = 0;       // splats collected in the chunk buffer (wave-uniform)
    uint64_t slots = 0;   // their slots in the staged batch, 6 bits each
    for (int64_t b = (int64_t)start + (int64_t)((stop - start - 1) / 64) * 64; b >= (int64_t)start; b -= 64) {
        // as in the forward: lanes without a valid instance read a valid
        // record (the replay's last value, splat 0) and are not live
        const uint32_t idx = (uint32_t)b + lane;
        const uint32_t g = vals[min(idx, stop - 1)];
        const bool valid = idx < stop && g < splats;
        const uint32_t gid = valid ? g : 0xFFFFFFFFu;
        const SplatRec sr = rec[valid ? g : 0u];
        const float4 co = sr.co, cl = sr.cl;
        const float2 p = make_float2(sr.pq.x, sr.pq.y);
        const bool live = valid && reaches_rect(p.x, p.y, co.x, co.y, co.z, sr.pq.z, rxa, rxb, rya, ryb);
        // Stage the batch in this wave's LDS slots (as in the forward).
        __builtin_amdgcn_wave_barrier();
        L.co[lane] = make_float4(-0.5f * co.x, -co.y, -0.5f * co.z, co.w);  // as in the forward
        L.cl[lane] = make_float4(cl.x, cl.y, cl.z, __uint_as_float(gid));
        L.p[lane] = p;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // live splats back to front (highest lane first); base: 1 + the
        // batch's position in the range, as a scalar, so that a splat's
        // position is a scalar add
        const uint32_t base = (uint32_t)__builtin_amdgcn_readfirstlane((int)((uint32_t)b - start + 1));
        for (uint64_t lm = ballot(live); lm != 0;) {
            const int j = 63 - (int)__builtin_clzll(lm);
            lm &= ~(1ull << j);
            const uint32_t pos = base + (uint32_t)j;
            const float4 sco = L.co[j], scl = L.cl[j];
            const float2 sp = L.p[j];
            const float A = sco.x, B = sco.y, C = sco.z, op = sco.w;  // -A/2, -B, -C/2: power only
            const float cr = scl.x, cg = scl.y, cb = scl.z;
            const float dx = sp.x - fx, dy = sp.y - fy;
            const float power = __builtin_fmaf(dx, __builtin_fmaf(A, dx, B * dy), (C * dy) * dy);
            const float e = exp_blend(power);
            const float oe = op * e;
            const float alpha = fminf(kAlphaClamp, oe);
            // blended in the forward: the forward's skip tests, as predicates
            const bool hit = (pos <= lst) & !(power > 0.0f) & !(power < -5.55f) & !(alpha < kAlphaMin);
            // Undo the blend (v_rcp_f32: 1 ulp, no IEEE division sequence);
            // alpha <= 0.99 keeps the recovered T finite in every lane.
            const float r = __builtin_amdgcn_rcpf(1.0f - alpha);
            const float Tb = T * r;
            const float gc = __builtin_fmaf(cr, gr, __builtin_fmaf(cg, gg, cb * gb));
            const float gal = __builtin_fmaf(Tb, gc, -(r * GS));
            asm volatile("" ::"v"(gal));  // keeps gal (and the colour read) ahead of the branch
            if (ballot(hit) == 0) continue;
            // dL/dpower (zero on the 0.99 clamp, where alpha does not depend
            // on the power) and the blend weight; both zero where not blended
            const bool grad_ok = hit & (oe <= kAlphaClamp);
            const float w = hit ? alpha * Tb : 0.0f;
            const float gpow = grad_ok ? alpha * gal : 0.0f;
            // state before this splat where it was blended (w = 0 elsewhere)
            T = hit ? Tb : T;
            GS = __builtin_fmaf(w, gc, GS);
            L.pair[n][lane ^ (n << 3)] = make_float2(gpow, w);
            slots |= (uint64_t)j << (6 * n);
            if (++n == (uint32_t)kBwdChunk) {
                backward_flush(L, n, slots, x0, y0, splats, grad2d);
                n = 0;
                slots = 0;
            }
        }
        // the next batch overwrites the staged records the slots refer to
        if (n) {
            backward_flush(L, n, slots, x0, y0, splats, grad2d);
            n = 0;
            slots = 0;
        }
    }
}

// True when the forward whose state the backward replays overflowed its
// instance capacity or had a bounded wait give up: its sorted instances,
// ranges and per-pixel state are then incomplete.
__device__ __forceinline__ bool frame_failed(const FrameCounters* counters, uint64_t capacity) {
    return counters->instances > capacity || counters->error != 0u;
}

template <int TS>
__global__ __launch_bounds__(256) void k_raster_backward(
    uint32_t W, uint32_t H, uint32_t tiles_x, uint32_t num_tiles, uint32_t splats, const FrameCounters* counters,
    uint64_t capacity, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vals,
    const SplatRec* __restrict__ rec, const float* __restrict__ tfinal, const uint32_t* __restrict__ last,
    const float* __restrict__ dLdC, float* __restrict__ grad2d, const uint32_t* __restrict__ order) {
    constexpr uint32_t SIDE = TS / 8, BLOCKS = SIDE * SIDE;
    const uint32_t wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t slot = wv / BLOCKS, sub = wv % BLOCKS;
    if (slot >= num_tiles || frame_failed(counters, capacity)) return;
    const uint32_t tile = order[slot];
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const uint2 r = ranges[tile];
    backward_block(W, H, tx * TS + (sub % SIDE) * 8, ty * TS + (sub / SIDE) * 8, r.x, min(r.y, counters->n), splats,
                   vals, rec, tfinal, last, dLdC, grad2d);
}

// View batches: the tiles of every view, view by view and heaviest first
// within a view (order[] over the global tile ids); views[v] = {W, H,
// tiles_x, tile_base}, pix[v] the float offset of view v's [3][H][W] image
// (and pix[v] / 3 its pixel offset).  Entries (vals) are v * n + g.
template <int TS>
__global__ __launch_bounds__(256) void k_raster_backward_batch(
    const uint4* __restrict__ views, const uint64_t* __restrict__ pix, uint32_t num_views, uint32_t total_tiles,
    uint32_t entries, const uint32_t* n_dev, const uint2* __restrict__ ranges, const uint32_t* __restrict__ vals,
    const SplatRec* __restrict__ rec, const float* __restrict__ tfinal, const uint32_t* __restrict__ last,
    const float* __restrict__ dLdC, float* __restrict__ grad2d, const uint32_t* __restrict__ order) {
    constexpr uint32_t SIDE = TS / 8, BLOCKS = SIDE * SIDE;
    const uint32_t wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t slot = wv / BLOCKS, sub = wv % BLOCKS;
    if (slot >= total_tiles) return;
    const uint32_t t = order[slot];
    uint32_t lo = 0, hi = num_views;  // view holding global tile t: last v with tile_base <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].w <= t) lo = mid;
        else hi = mid;
    }
    const uint4 vi = views[lo];
    const uint32_t local = t - vi.w;
    const uint32_t tx = local % vi.z, ty = local / vi.z;
    const uint2 r = ranges[t];
    const uint64_t po = pix[lo] / 3;
    backward_block(vi.x, vi.y, tx * TS + (sub % SIDE) * 8, ty * TS + (sub / SIDE) * 8, r.x, min(r.y, *n_dev), entries,
                   vals, rec, tfinal + po, last + po, dLdC + pix[lo], grad2d);
}

constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f, SH_C2_1 = -1.0925484305920792f,
                SH_C2_2 = 0.31539156525252005f, SH_C2_3 = -1.0925484305920792f,
                SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f, SH_C3_1 = 2.890611442640554f,
                SH_C3_2 = -0.4570457994644658f, SH_C3_3 = 0.3731763325901154f,
                SH_C3_4 = -0.4570457994644658f, SH_C3_5 = 1.445305721320277f,
                SH_C3_6 = -0.5900435899266435f;

// The raster's per-splat sums (moments of dL/dpower in slots 0-4, then
// opacity and colour) in the 2-D gradient layout, except the position:
// d/dA = -Mxx / 2, d/dB = -Mxy, d/dC = -Myy / 2; Gv[0], Gv[1] keep the
// moments Mx, My, which view_adjoint turns into d/dx = -(A Mx + B My),
// d/dy = -(C My + B Mx) with the conic (A, B, C) it recomputes.
__device__ __forceinline__ void sums_to_2d(const float (&S)[kSumFloats], float (&Gv)[9]) {
    Gv[0] = S[0];
    Gv[1] = S[1];
    Gv[2] = -0.5f * S[2];
    Gv[3] = -S[3];
    Gv[4] = -0.5f * S[4];
    Gv[5] = S[5];
    Gv[6] = S[6];
    Gv[7] = S[7];
    Gv[8] = S[8];
}

// One entry's sums: 36 bytes at a 4-byte alignment, which the loads and
// stores of consecutive threads cover without gaps.  Returns whether any
// of them is nonzero (sums start at +0 and atomic adds never make -0, so
// a zero entry is all +0 and needs no clearing).
__device__ __forceinline__ bool load_sums(const float* __restrict__ e, float (&S)[kSumFloats]) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < (int)kSumFloats; ++k) {
        S[k] = e[k];
        any |= S[k] != 0.0f;
    }
    return any;
}

__device__ __forceinline__ void clear_sums(float* __restrict__ e) {
#pragma unroll
    for (int k = 0; k < (int)kSumFloats; ++k) e[k] = 0.0f;
}

// View-independent part of one splat for the preprocess adjoint (the
// forward's covariance construction, preprocess.wgsl order).
struct SplatShape {
    float px, py, pz;
    float sv[3];                    // exp(log-scale)
    float qw, qx, qy, qz, ql;       // normalised rotation, length of the raw one
    float R[3][3], M[3][3], S[3][3];  // rotation, R diag(s), 3-D covariance
};

// One splat's parameter channels: read from the planar scene, or from
// registers the caller filled.
struct ParGlobal {
    const float* __restrict__ P;
    size_t N;
    uint32_t g;
    __device__ float operator()(int ch) const { return P[(size_t)ch * N + g]; }
};
struct ParRegs {
    const float* v;  // the 59 channels
    __device__ float operator()(int ch) const { return v[ch]; }
};
// The 48 SH channels from a thread's column of an LDS table (row k: SH value
// k, f_dc then f_rest), the others from the planar scene.
constexpr int kShChannels = 48;
struct ParShLds {
    const float* __restrict__ P;
    size_t N;
    uint32_t g;
    const float (*sh)[256];
    uint32_t col;
    __device__ float operator()(int ch) const {
        if (ch >= CH_DC && ch < CH_DC + 3) return sh[ch - CH_DC][col];
        if (ch >= CH_REST) return sh[3 + ch - CH_REST][col];
        return P[(size_t)ch * N + g];
    }
};

template <typename Par>
__device__ __forceinline__ void splat_shape(const Par& par, SplatShape& s) {
    s.px = par(CH_POS);
    s.py = par(CH_POS + 1);
    s.pz = par(CH_POS + 2);
#pragma unroll
    for (int c = 0; c < 3; ++c) s.sv[c] = gsr_exp(par(CH_SCALE + c));
    const float rw = par(CH_ROT + 0), rx = par(CH_ROT + 1);
    const float ry = par(CH_ROT + 2), rz = par(CH_ROT + 3);
    s.ql = sqrtf(((rx * rx + ry * ry) + rz * rz) + rw * rw);
    const float qw = rw / s.ql, qx = rx / s.ql, qy = ry / s.ql, qz = rz / s.ql;
    s.qw = qw;
    s.qx = qx;
    s.qy = qy;
    s.qz = qz;
    s.R[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz); s.R[0][1] = 2.0f * (qx * qy - qw * qz); s.R[0][2] = 2.0f * (qx * qz + qw * qy);
    s.R[1][0] = 2.0f * (qx * qy + qw * qz); s.R[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz); s.R[1][2] = 2.0f * (qy * qz - qw * qx);
    s.R[2][0] = 2.0f * (qx * qz - qw * qy); s.R[2][1] = 2.0f * (qy * qz + qw * qx); s.R[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) s.M[r][c] = s.R[r][c] * s.sv[c];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            s.S[i][j] = s.M[i][0] * s.M[j][0] + s.M[i][1] * s.M[j][1] + s.M[i][2] * s.M[j][2];
}

// Adjoint of one view's projection of the splat, for the 2-D gradient Gv
// (pixel x, y; conic A, B, C; opacity; r, g, b).  Returns false when the
// splat is behind the near plane of this view (no gradient).  Otherwise:
//   gp    dL/dposition through the pixel position and the covariance;
//   gsym  dL/dS + (dL/dS)^T of the 3-D covariance;
//   gch   dL/d(SH colour) per channel (zero where the forward clamped it);
//   basis the 15 higher SH basis values of the view direction;
// and the view direction's part of dL/dposition is added to gp.
// Gv[0], Gv[1] arrive as the raster's moments sum g dx, sum g dy and leave
// as d/dx, d/dy, formed with the conic of this function's own covariance
// chain.
template <typename Par, typename Shape>
__device__ __forceinline__ bool view_adjoint(const Par& par, const Shape& s,
                                             const CamK& cam, float (&Gv)[9], float (&gp)[3],
                                             float (&gsym)[3][3], float (&gch)[3], float (&basis)[15]) {
    const float px = s.px, py = s.py, pz = s.pz;
    const float* V = cam.view;
    const float* PV = cam.view_proj;
    const float tx = ((V[0] * px + V[4] * py) + V[8] * pz) + V[12];
    const float ty = ((V[1] * px + V[5] * py) + V[9] * pz) + V[13];
    const float tz = ((V[2] * px + V[6] * py) + V[10] * pz) + V[14];
    if (!(tz < -kNearView)) return false;
    // covariance chain
    const float limx = kFrustumGuard * cam.tan_fov[0], limy = kFrustumGuard * cam.tan_fov[1];
    const float vx = tx / tz, vy = ty / tz;
    const bool clx = vx < -limx || vx > limx, cly = vy < -limy || vy > limy;
    const float ux = fminf(limx, fmaxf(-limx, vx)) * tz;
    const float uy = fminf(limy, fmaxf(-limy, vy)) * tz;
    const float uz = tz;
    const float fx = cam.focal[0], fy = cam.focal[1];
    const float iuz = 1.0f / uz, iuz2 = iuz * iuz;
    const float J00 = fx * iuz, J02 = -fx * ux * iuz2, J11 = fy * iuz, J12 = -fy * uy * iuz2;
    float Tm[2][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Tm[0][c] = J00 * V[c * 4 + 0] + J02 * V[c * 4 + 2];
        Tm[1][c] = J11 * V[c * 4 + 1] + J12 * V[c * 4 + 2];
    }
    float St0[3], St1[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        St0[i] = s.S[i][0] * Tm[0][0] + s.S[i][1] * Tm[0][1] + s.S[i][2] * Tm[0][2];
        St1[i] = s.S[i][0] * Tm[1][0] + s.S[i][1] * Tm[1][1] + s.S[i][2] * Tm[1][2];
    }
    const float a = Tm[0][0] * St0[0] + Tm[0][1] * St0[1] + Tm[0][2] * St0[2] + 0.3f;
    const float b = Tm[0][0] * St1[0] + Tm[0][1] * St1[1] + Tm[0][2] * St1[2];
    const float c = Tm[1][0] * St1[0] + Tm[1][1] * St1[1] + Tm[1][2] * St1[2] + 0.3f;
    const float det = a * c - b * b;
    {   // the pixel position's gradient from its moments, with the conic
        // (c, -b, a) / det of this chain (not read from the forward's record)
        const float A = c / det, B = -b / det, C = a / det;
        const float mx = Gv[0], my = Gv[1];
        Gv[0] = -__builtin_fmaf(A, mx, B * my);
        Gv[1] = -__builtin_fmaf(C, my, B * mx);
    }
    {   // pixel position <- ndc <- clip <- p
        const float cx = ((PV[0] * px + PV[4] * py) + PV[8] * pz) + PV[12];
        const float cy = ((PV[1] * px + PV[5] * py) + PV[9] * pz) + PV[13];
        const float cw = ((PV[3] * px + PV[7] * py) + PV[11] * pz) + PV[15];
        const float w = cw + kSmall;
        const float gnx = Gv[0] * cam.texture_size[0] * 0.5f, gny = Gv[1] * cam.texture_size[1] * 0.5f;
        const float iw = 1.0f / w;
        const float gcx = gnx * iw, gcy = gny * iw, gcw = -(gnx * cx + gny * cy) * iw * iw;
        gp[0] = PV[0] * gcx + PV[1] * gcy + PV[3] * gcw;
        gp[1] = PV[4] * gcx + PV[5] * gcy + PV[7] * gcw;
        gp[2] = PV[8] * gcx + PV[9] * gcy + PV[11] * gcw;
    }
    const float id2 = 1.0f / (det * det);
    const float gA = Gv[2], gB = Gv[3], gC = Gv[4];
    const float ga = (gA * (-c * c) + gB * (b * c) + gC * (-b * b)) * id2;
    const float gbb = (gA * (2.0f * b * c) - gB * (det + 2.0f * b * b) + gC * (2.0f * a * b)) * id2;
    const float gc = (gA * (-b * b) + gB * (a * b) + gC * (-a * a)) * id2;
    float gt0[3], gt1[3], gS[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        gt0[i] = 2.0f * ga * St0[i] + gbb * St1[i];
        gt1[i] = gbb * St0[i] + 2.0f * gc * St1[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            gS[i][j] = ga * Tm[0][i] * Tm[0][j] + gbb * Tm[0][i] * Tm[1][j] + gc * Tm[1][i] * Tm[1][j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) gsym[i][j] = gS[i][j] + gS[j][i];
    float gJ00 = 0.f, gJ02 = 0.f, gJ11 = 0.f, gJ12 = 0.f;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
        gJ00 += gt0[cc] * V[cc * 4 + 0];
        gJ02 += gt0[cc] * V[cc * 4 + 2];
        gJ11 += gt1[cc] * V[cc * 4 + 1];
        gJ12 += gt1[cc] * V[cc * 4 + 2];
    }
    const float gux = gJ02 * (-fx * iuz2), guy = gJ12 * (-fy * iuz2);
    const float guz = gJ00 * (-fx * iuz2) + gJ02 * (2.0f * fx * ux * iuz2 * iuz) + gJ11 * (-fy * iuz2) +
                      gJ12 * (2.0f * fy * uy * iuz2 * iuz);
    float gtx = clx ? 0.0f : gux, gty = cly ? 0.0f : guy, gtz = guz;
    if (clx) gtz += gux * (vx > 0.0f ? limx : -limx);
    if (cly) gtz += guy * (vy > 0.0f ? limy : -limy);
    gp[0] += V[0] * gtx + V[1] * gty + V[2] * gtz;
    gp[1] += V[4] * gtx + V[5] * gty + V[6] * gtz;
    gp[2] += V[8] * gtx + V[9] * gty + V[10] * gtz;
    // colour
    const float vx3 = px - cam.eye[0], vy3 = py - cam.eye[1], vz3 = pz - cam.eye[2];
    const float vl = sqrtf(vx3 * vx3 + vy3 * vy3 + vz3 * vz3);
    const float X = vx3 / vl, Y = vy3 / vl, Z = vz3 / vl;
    const float XX = X * X, YY = Y * Y, ZZ = Z * Z;
    const float bs[15] = {-SH_C1 * Y, SH_C1 * Z, -SH_C1 * X, SH_C2_0 * X * Y, SH_C2_1 * Y * Z,
                          SH_C2_2 * (2.0f * ZZ - XX - YY), SH_C2_3 * X * Z, SH_C2_4 * (XX - YY),
                          SH_C3_0 * Y * (3.0f * XX - YY), SH_C3_1 * X * Y * Z, SH_C3_2 * Y * (4.0f * ZZ - XX - YY),
                          SH_C3_3 * Z * (2.0f * ZZ - 3.0f * XX - 3.0f * YY), SH_C3_4 * X * (4.0f * ZZ - XX - YY),
                          SH_C3_5 * Z * (XX - YY), SH_C3_6 * X * (XX - 3.0f * YY)};
    const float dbx[15] = {0.f, 0.f, -SH_C1, SH_C2_0 * Y, 0.f, SH_C2_2 * (-2.0f * X), SH_C2_3 * Z, SH_C2_4 * 2.0f * X,
                           SH_C3_0 * Y * 6.0f * X, SH_C3_1 * Y * Z, SH_C3_2 * Y * (-2.0f * X), SH_C3_3 * Z * (-6.0f * X),
                           SH_C3_4 * (4.0f * ZZ - 3.0f * XX - YY), SH_C3_5 * Z * 2.0f * X, SH_C3_6 * (3.0f * XX - 3.0f * YY)};
    const float dby[15] = {-SH_C1, 0.f, 0.f, SH_C2_0 * X, SH_C2_1 * Z, SH_C2_2 * (-2.0f * Y), 0.f, SH_C2_4 * (-2.0f * Y),
                           SH_C3_0 * (3.0f * XX - 3.0f * YY), SH_C3_1 * X * Z, SH_C3_2 * (4.0f * ZZ - XX - 3.0f * YY),
                           SH_C3_3 * Z * (-6.0f * Y), SH_C3_4 * X * (-2.0f * Y), SH_C3_5 * Z * (-2.0f * Y),
                           SH_C3_6 * X * (-6.0f * Y)};
    const float dbz[15] = {0.f, SH_C1, 0.f, 0.f, SH_C2_1 * Y, SH_C2_2 * 4.0f * Z, SH_C2_3 * X, 0.f, 0.f,
                           SH_C3_1 * X * Y, SH_C3_2 * Y * 8.0f * Z, SH_C3_3 * (6.0f * ZZ - 3.0f * XX - 3.0f * YY),
                           SH_C3_4 * X * 8.0f * Z, SH_C3_5 * (XX - YY), 0.f};
#pragma unroll
    for (int k = 0; k < 15; ++k) basis[k] = bs[k];
    float gdx = 0.f, gdy = 0.f, gdz = 0.f;
    const float gcol[3] = {Gv[6], Gv[7], Gv[8]};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float acc = SH_C0 * par(CH_DC + ch);
        float f[15];
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            f[k] = par(CH_REST + 15 * ch + k);
            acc += bs[k] * f[k];
        }
        const float gv = (acc + 0.5f < 0.0f) ? 0.0f : gcol[ch];
        gch[ch] = gv;
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            gdx += gv * dbx[k] * f[k];
            gdy += gv * dby[k] * f[k];
            gdz += gv * dbz[k] * f[k];
        }
    }
    const float dd = X * gdx + Y * gdy + Z * gdz;
    const float ivl = 1.0f / vl;
    gp[0] += (gdx - X * dd) * ivl;
    gp[1] += (gdy - Y * dd) * ivl;
    gp[2] += (gdz - Z * dd) * ivl;
    return true;
}

// dL/dS + (dL/dS)^T -> gs: dL/d log-scale (3), dL/d raw quaternion (4).
__device__ __forceinline__ void shape_adjoint(const SplatShape& s, const float (&gsym)[3][3], float (&gs)[7]) {
    float gM[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            gM[r][cc] = gsym[r][0] * s.M[0][cc] + gsym[r][1] * s.M[1][cc] + gsym[r][2] * s.M[2][cc];
    float gR[3][3];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
        float gsc = 0.f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            gR[r][cc] = gM[r][cc] * s.sv[cc];
            gsc += gM[r][cc] * s.R[r][cc];
        }
        gs[cc] = gsc * s.sv[cc];
    }
    const float qw = s.qw, qx = s.qx, qy = s.qy, qz = s.qz;
    const float gqw = 2.0f * (-qz * gR[0][1] + qy * gR[0][2] + qz * gR[1][0] - qx * gR[1][2] - qy * gR[2][0] + qx * gR[2][1]);
    const float gqx = 2.0f * (qy * gR[0][1] + qz * gR[0][2] + qy * gR[1][0] - 2.0f * qx * gR[1][1] - qw * gR[1][2] +
                              qz * gR[2][0] + qw * gR[2][1] - 2.0f * qx * gR[2][2]);
    const float gqy = 2.0f * (-2.0f * qy * gR[0][0] + qx * gR[0][1] + qw * gR[0][2] + qx * gR[1][0] + qz * gR[1][2] -
                              qw * gR[2][0] + qz * gR[2][1] - 2.0f * qy * gR[2][2]);
    const float gqz = 2.0f * (-2.0f * qz * gR[0][0] - qw * gR[0][1] + qx * gR[0][2] + qw * gR[1][0] - 2.0f * qz * gR[1][1] +
                              qy * gR[1][2] + qx * gR[2][0] + qy * gR[2][1]);
    const float dotq = qw * gqw + qx * gqx + qy * gqy + qz * gqz;
    const float iql = 1.0f / s.ql;
    gs[3] = (gqw - qw * dotq) * iql;
    gs[4] = (gqx - qx * dotq) * iql;
    gs[5] = (gqy - qy * dotq) * iql;
    gs[6] = (gqz - qz * dotq) * iql;
}

template <typename Par>
__device__ __forceinline__ float opacity_of(const Par& par) {
    return 1.0f / (1.0f + gsr_exp(-par(CH_OPACITY)));
}

// grad2d arrives with the raster's per-splat sums and leaves cleared for the
// next backward; out2d (nullable) receives the 2-D gradient (the position's
// two terms formed with the conic that view_adjoint recomputes), grad the 59
// parameter gradients.  When the forward failed (frame_failed) every output
// is a quiet NaN, so a caller that skips the error gsr_synchronize reports
// cannot take the result for a gradient.
__global__ __launch_bounds__(256) void k_preprocess_backward(const float* __restrict__ P, uint32_t n, CamK cam,
                                                             const FrameCounters* __restrict__ counters,
                                                             uint64_t capacity, float* __restrict__ grad2d,
                                                             float4* __restrict__ out2d,
                                                             float* __restrict__ grad) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = n;
    if (frame_failed(counters, capacity)) {
        const float q = __builtin_nanf("");
        if (out2d)
            for (int k = 0; k < 3; ++k) out2d[(size_t)g * 3 + k] = make_float4(q, q, q, q);
        for (int ch = 0; ch < (int)kParams; ++ch) grad[(size_t)ch * N + g] = q;
        return;
    }
    // every load up front (sums, the 59 parameters): one memory round trip
    // per thread instead of one per dependent stage
    float S[kSumFloats];
    const bool any = load_sums(grad2d + (size_t)g * kSumFloats, S);
    float pv[kParams];
#pragma unroll
    // streaming loads, as in preprocess (backward preprocess 113 -> 106 us)
    for (int ch = 0; ch < (int)kParams; ++ch) pv[ch] = __builtin_nontemporal_load(P + (size_t)ch * N + g);
    const ParRegs par{pv};
    // the 2-D gradient with the position's two moments left as they are
    // (view_adjoint turns them into d/dx, d/dy: no read of the forward's
    // record)
    float Gv[9];
    sums_to_2d(S, Gv);
    SplatShape s;
    float gp[3], gsym[3][3], gch[3], basis[15], gs[7], gop = 0.0f;
    bool ok = any;
    if (ok) {
        splat_shape(par, s);
        ok = view_adjoint(par, s, cam, Gv, gp, gsym, gch, basis);
    }
    if (any) clear_sums(grad2d + (size_t)g * kSumFloats);  // read: cleared for the next backward's sums
    if (out2d) {  // zeros (not -0) for a splat without sums
        out2d[(size_t)g * 3] = any ? make_float4(ok ? Gv[0] : 0.0f, ok ? Gv[1] : 0.0f, Gv[2], Gv[3])
                                   : make_float4(0.f, 0.f, 0.f, 0.f);
        out2d[(size_t)g * 3 + 1] = any ? make_float4(Gv[4], Gv[5], Gv[6], Gv[7]) : make_float4(0.f, 0.f, 0.f, 0.f);
        out2d[(size_t)g * 3 + 2] = make_float4(any ? Gv[8] : 0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (ok) {
        shape_adjoint(s, gsym, gs);
        const float op = opacity_of(par);
        gop = Gv[5] * op * (1.0f - op);
    }
    // Every lane stores through the same instructions, zeros where the splat
    // has no gradient: stores of one line from two branches far apart in
    // time let a partly written line leave the cache and be written twice.
    // (The selects keep the arrays unset on that path: no registers for zeros.)
#pragma unroll
    for (int k = 0; k < 7; ++k) grad[(size_t)(CH_SCALE + k) * N + g] = ok ? gs[k] : 0.0f;  // scale, rotation
    grad[(size_t)CH_OPACITY * N + g] = gop;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        grad[(size_t)(CH_DC + ch) * N + g] = ok ? SH_C0 * gch[ch] : 0.0f;
#pragma unroll
        for (int k = 0; k < 15; ++k)
            grad[(size_t)(CH_REST + 15 * ch + k) * N + g] = ok ? basis[k] * gch[ch] : 0.0f;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) grad[(size_t)(CH_POS + c) * N + g] = ok ? gp[c] : 0.0f;
}

// The part of SplatShape the per-view adjoint reads.
struct SplatCov {
    float px, py, pz;
    float S[3][3];
};

// An opaque copy of a pointer: loads through it are not hoisted out of the
// loop it is made in or merged with loads through the original, so values
// the compiler would otherwise keep in registers across a long loop are
// re-read (from cache) where they are used.
template <typename T>
__device__ __forceinline__ const T* opaque(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// View batches: one thread per splat walks the views, converting each
// entry's sums (entry v * n + g of grad2d, conic from rec at the same index)
// and summing the parameter gradient over views.  Every gradient is linear
// in the view's 2-D gradient given the view's forward decisions, so the sums
// are taken before the view-independent factors: dL/dS + (dL/dS)^T (6
// values) before the scale / rotation chain, dL/d opacity before the sigmoid
// slope, dL/d colour before the SH-DC constant.  grad2d_out (nullable)
// receives every entry in the 2-D gradient layout.
// At most 168 VGPRs: 3 waves per SIMD (the 48 KB SH table allows 3 workgroups per CU)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_preprocess_backward_batch(const float* __restrict__ P, uint32_t n,
                                                                   const CamK* __restrict__ cams, uint32_t views,
                                                                   float* __restrict__ grad2d,
                                                                   float4* __restrict__ grad2d_out,
                                                                   float* __restrict__ grad) {
    // The SH coefficients are read per view; from HBM that was 192 B per
    // touched (view, splat), which missed L2 (the per-view sums stream past
    // it).  Each thread keeps its own in a column of this table (48 KB per
    // workgroup, which the kernel's 3 waves per SIMD leave room for); only
    // the thread itself reads its column, so no barrier is needed.
    __shared__ float s_sh[kShChannels][256];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = n;
    const uint32_t col = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kShChannels; ++k)
        s_sh[k][col] = __builtin_nontemporal_load(P + (size_t)(k < 3 ? CH_DC + k : CH_REST + k - 3) * N + g);
    SplatCov cv;
    bool shaped = false, touched = false;
    float sp[3] = {0.f, 0.f, 0.f}, ss[6] = {}, sop = 0.0f, sch[3] = {0.f, 0.f, 0.f}, srest[45] = {};
    for (uint32_t v = 0; v < views; ++v) {
        const size_t e = (size_t)v * N + g;
        float S[kSumFloats];
        const bool any = load_sums(grad2d + e * kSumFloats, S);
        if (any) clear_sums(grad2d + e * kSumFloats);  // read: cleared for the next backward's sums
        // the 2-D gradient with the position's two moments left as they are:
        // view_adjoint turns them into d/dx, d/dy with the conic it
        // recomputes (no read of the forward's record)
        float Gv[9];
        sums_to_2d(S, Gv);
        if (!any) {
            if (grad2d_out) {
                grad2d_out[e * 3] = make_float4(0.f, 0.f, 0.f, 0.f);
                grad2d_out[e * 3 + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
                grad2d_out[e * 3 + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            continue;
        }
        const float* Pv = opaque(P);  // the shape read where it is needed, not held through the loop
        if (!shaped) {
            SplatShape full;
            splat_shape(ParGlobal{Pv, N, g}, full);
            cv.px = full.px;
            cv.py = full.py;
            cv.pz = full.pz;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) cv.S[i][j] = full.S[i][j];
            shaped = true;
        }
        float gp[3], gsym[3][3], gch[3], basis[15];
        const bool ok = view_adjoint(ParShLds{Pv, N, g, s_sh, col}, cv, cams[v], Gv, gp, gsym, gch, basis);
        if (grad2d_out) {
            grad2d_out[e * 3] = make_float4(ok ? Gv[0] : 0.f, ok ? Gv[1] : 0.f, Gv[2], Gv[3]);
            grad2d_out[e * 3 + 1] = make_float4(Gv[4], Gv[5], Gv[6], Gv[7]);
            grad2d_out[e * 3 + 2] = make_float4(Gv[8], 0.0f, 0.0f, 0.0f);
        }
        if (!ok) continue;
        touched = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) sp[c] += gp[c];
        ss[0] += gsym[0][0];
        ss[1] += gsym[0][1];
        ss[2] += gsym[0][2];
        ss[3] += gsym[1][1];
        ss[4] += gsym[1][2];
        ss[5] += gsym[2][2];
        sop += Gv[5];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            sch[ch] += gch[ch];
#pragma unroll
            for (int k = 0; k < 15; ++k) srest[15 * ch + k] = __builtin_fmaf(basis[k], gch[ch], srest[15 * ch + k]);
        }
    }
    // an untouched splat keeps its zero sums and gets zeros through the same
    // stores as the others (one store instruction per channel and wave)
    float gs[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gop = 0.0f;
    if (touched) {
        const float* Pe = opaque(P);  // the shape again, not kept live through the loop
        SplatShape s;
        splat_shape(ParGlobal{Pe, N, g}, s);
        const float sym[3][3] = {{ss[0], ss[1], ss[2]}, {ss[1], ss[3], ss[4]}, {ss[2], ss[4], ss[5]}};
        shape_adjoint(s, sym, gs);
        const float op = opacity_of(ParGlobal{Pe, N, g});
        gop = sop * op * (1.0f - op);
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) grad[(size_t)(CH_SCALE + k) * N + g] = gs[k];  // scale, then rotation
    grad[(size_t)CH_OPACITY * N + g] = gop;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        grad[(size_t)(CH_DC + ch) * N + g] = SH_C0 * sch[ch];
#pragma unroll
        for (int k = 0; k < 15; ++k) grad[(size_t)(CH_REST + 15 * ch + k) * N + g] = srest[15 * ch + k];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) grad[(size_t)(CH_POS + c) * N + g] = sp[c];
}

}  // namespace

int launch_backward(Context* c, const float* params, uint32_t count, const CamK& cam, const float* dLdC,
                    float* grad, float* grad2d_out) {
    if (count == 0) return GSR_OK;  // no splats, no gradients
    const uint32_t W = c->state_w, H = c->state_h, tile = c->state_tile;
    const uint32_t tx = ceil_div(W, tile), ty = ceil_div(H, tile);
    const size_t N = count;
    float* g2 = (float*)buffer(c, "grad2d", N * kSumFloats * 4);
    if (!g2) return GSR_ERR_ALLOC;
    // the previous backward cleared the sums it read, unless the buffer is
    // new or of another size
    int rc = c->grad2d_clean == N ? GSR_OK : launch_fill_u32(c, (uint32_t*)g2, 0u, N * kSumFloats);
    c->grad2d_clean = 0;
    if (rc) return rc;
    const uint32_t* vals = (const uint32_t*)buffer(c, c->state_sorted_index ? "vals1" : "vals0", 0);
    const uint2* ranges = (const uint2*)buffer(c, "ranges", 0);
    const SplatRec* rec = (const SplatRec*)buffer(c, "splat_rec", 0);
    const float* tfinal = (const float*)buffer(c, "tfinal", 0);
    const uint32_t* last = (const uint32_t*)buffer(c, "last", 0);
    const FrameCounters* counters = (const FrameCounters*)buffer(c, "counters", 0);
    mark(c, "backward raster");
    const uint32_t tiles = tx * ty;
    const uint32_t blocks = ceil_div(tiles * (tile * tile / 64), 4u);
    // the forward's heaviest-first order of the same tile ranges
    const uint32_t* order = (const uint32_t*)buffer(c, "frame_tile_order", 0);
    if (!order) return GSR_ERR_ALLOC;
    switch (tile) {
        case 8:
            hipLaunchKernelGGL(k_raster_backward<8>, dim3(blocks), dim3(256), 0, c->stream, W, H, tx, tiles, count,
                               counters, c->state_capacity, ranges, vals, rec, tfinal, last, dLdC, g2, order);
            break;
        case 16:
            hipLaunchKernelGGL(k_raster_backward<16>, dim3(blocks), dim3(256), 0, c->stream, W, H, tx, tiles, count,
                               counters, c->state_capacity, ranges, vals, rec, tfinal, last, dLdC, g2, order);
            break;
        case 32:
            hipLaunchKernelGGL(k_raster_backward<32>, dim3(blocks), dim3(256), 0, c->stream, W, H, tx, tiles, count,
                               counters, c->state_capacity, ranges, vals, rec, tfinal, last, dLdC, g2, order);
            break;
        default:
            return fail(c, GSR_ERR_STATE, "bad saved tile size %u", tile);
    }
    GSR_HIP(c, hipGetLastError());
    mark(c, "backward preprocess");
    hipLaunchKernelGGL(k_preprocess_backward, dim3(ceil_div(count, 256)), dim3(256), 0, c->stream, params, count, cam,
                       counters, c->state_capacity, g2, (float4*)grad2d_out, grad);
    GSR_HIP(c, hipGetLastError());
    c->grad2d_clean = N;
    mark_end(c, false);
    return GSR_OK;
}

int launch_backward_batch(Context* c, const float* params, uint32_t count, uint32_t views, uint32_t tile,
                          uint32_t total_tiles, const CamK* cams, const uint4* info, const uint64_t* pix,
                          int sorted_index, const float* dLdC, float* grad, float* grad2d_out) {
    if (count == 0) return GSR_OK;  // no splats, no gradients
    const size_t NV = (size_t)count * views;
    float* g2 = (float*)buffer(c, "b_grad2d", NV * kSumFloats * 4);
    uint32_t* order = (uint32_t*)buffer(c, "b_tile_order", (size_t)total_tiles * 4);
    if (!g2 || !order) return GSR_ERR_ALLOC;
    // the previous batched backward cleared the sums it read, unless the
    // buffer is new or of another shape
    int rc = c->b_grad2d_clean == NV ? GSR_OK : launch_fill_u32(c, (uint32_t*)g2, 0u, NV * kSumFloats);
    c->b_grad2d_clean = 0;
    if (rc) return rc;
    const uint32_t* vals = (const uint32_t*)buffer(c, sorted_index ? "b_vals1" : "b_vals0", 0);
    const uint2* ranges = (const uint2*)buffer(c, "b_ranges", 0);
    const SplatRec* rec = (const SplatRec*)buffer(c, "b_splat_rec", 0);
    const float* tfinal = (const float*)buffer(c, "b_tfinal", 0);
    const uint32_t* last = (const uint32_t*)buffer(c, "b_last", 0);
    const FrameCounters* counters = (const FrameCounters*)buffer(c, "b_counters", 0);
    mark(c, "backward raster");
    // view by view (heaviest first within a view): the atomics of a stretch of
    // the pass then land in one view's 2-D gradient sums, not all views'
    if ((rc = launch_tile_order(c, ranges, total_tiles, order, info, views))) return rc;
    const uint32_t blocks = ceil_div(total_tiles * (tile * tile / 64), 4u);
    switch (tile) {
        case 8:
            hipLaunchKernelGGL(k_raster_backward_batch<8>, dim3(blocks), dim3(256), 0, c->stream, info, pix, views,
                               total_tiles, (uint32_t)NV, &counters->n, ranges, vals, rec, tfinal, last, dLdC, g2,
                               order);
            break;
        case 16:
            hipLaunchKernelGGL(k_raster_backward_batch<16>, dim3(blocks), dim3(256), 0, c->stream, info, pix, views,
                               total_tiles, (uint32_t)NV, &counters->n, ranges, vals, rec, tfinal, last, dLdC, g2,
                               order);
            break;
        case 32:
            hipLaunchKernelGGL(k_raster_backward_batch<32>, dim3(blocks), dim3(256), 0, c->stream, info, pix, views,
                               total_tiles, (uint32_t)NV, &counters->n, ranges, vals, rec, tfinal, last, dLdC, g2,
                               order);
            break;
        default:
            return fail(c, GSR_ERR_STATE, "bad saved tile size %u", tile);
    }
    GSR_HIP(c, hipGetLastError());
    mark(c, "backward preprocess");
    hipLaunchKernelGGL(k_preprocess_backward_batch, dim3(ceil_div(count, 256)), dim3(256), 0, c->stream, params, count,
                       cams, views, g2, (float4*)grad2d_out, grad);
    GSR_HIP(c, hipGetLastError());
    c->b_grad2d_clean = NV;
    mark_end(c, false);
    return GSR_OK;
}

}  // namespace gsr

using namespace gsr;

extern "C" {

int gsr_backward(gsr_context* ctx, const float* params_dev, uint32_t count, const gsr_camera* cam,
                 const float* dL_drgb_dev, float* grad_params_dev, float* grad2d_dev) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!cam || ((!params_dev || !dL_drgb_dev || !grad_params_dev) && count > 0))
        return fail(c, GSR_ERR_ARGUMENT, "gsr_backward: null argument");
    if (!c->have_state || c->state_count != count || c->state_w != (uint32_t)cam->texture_size[0] ||
        c->state_h != (uint32_t)cam->texture_size[1])
        return fail(c, GSR_ERR_STATE, "gsr_backward needs the matching gsr_forward with save_state = 1 first");
    CamK ck;
    std::memcpy(&ck, cam, sizeof(ck));
    ck.count = count;
    // the raster adjoint replays the saved forward and the preprocess adjoint
    // recomputes the projection from these arguments: both must be the forward's
    if (std::memcmp(&ck, &c->state_cam, offsetof(CamK, pad)) != 0)
        return fail(c, GSR_ERR_STATE, "gsr_backward: the camera differs from the saved forward's");
    if (count > 0 && params_dev != c->state_params)
        return fail(c, GSR_ERR_STATE, "gsr_backward: the parameters are not the saved forward's");
    hipSetDevice(c->device);
    return launch_backward(c, params_dev, count, ck, dL_drgb_dev, grad_params_dev, grad2d_dev);
}

}  // extern "C"
