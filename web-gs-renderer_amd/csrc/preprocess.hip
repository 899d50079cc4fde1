// preprocess.hip — per-splat projection, 2-D covariance, conic, radius, SH
// colour, tile rectangle and depth key (src/shaders/preprocess.wgsl:58-284).
//
// One thread per splat, 256-thread blocks (4 wave64).  The scene is planar
// ([59][N] float32), so each of the 59 per-splat loads is one fully coalesced
// 256-byte wave access, all 59 issued before any arithmetic so that a
// thread waits for memory once.  FP32 operation order is the
// oracle's (no contraction, fmaf where written, IEEE sqrt/div, gsr_exp).
// The view-independent part (covariance, opacity) is split from the
// per-camera part so that view batches compute it once per splat; the
// arithmetic and its order are the same either way.
#include "cull.h"
#include "gsr_internal.h"

namespace gsr {
namespace {

constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f, SH_C2_1 = -1.0925484305920792f,
                SH_C2_2 = 0.31539156525252005f, SH_C2_3 = -1.0925484305920792f,
                SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f, SH_C3_1 = 2.890611442640554f,
                SH_C3_2 = -0.4570457994644658f, SH_C3_3 = 0.3731763325901154f,
                SH_C3_4 = -0.4570457994644658f, SH_C3_5 = 1.445305721320277f,
                SH_C3_6 = -0.5900435899266435f;

// The view-independent part of a splat: position, 3-D covariance
// Sigma = (R diag(s)) (R diag(s))^T (quantToMat3 / scaleToMat3,
// preprocess.wgsl:173-201; symmetric, so six entries) and opacity.
struct SplatFixed {
    float px, py, pz;
    float S00, S01, S02, S11, S12, S22;
    float opacity;
};

// Scene parameters are read once per frame (per view batch): loaded as
// non-temporal (streaming) so they do not displace what the later stages
// reuse from L2 and the MALL.  Measured on the 1M-splat 1080p frame: forward
// preprocess 62 -> 60 us, and 84 -> 57 us when a backward runs between frames.
__device__ __forceinline__ float ld_param(const float* p) { return __builtin_nontemporal_load(p); }

__device__ __forceinline__ SplatFixed splat_fixed(const float* __restrict__ P, size_t N, uint32_t g) {
    SplatFixed f;
    f.px = ld_param(P + (CH_POS * N + g));
    f.py = ld_param(P + ((CH_POS + 1) * N + g));
    f.pz = ld_param(P + ((CH_POS + 2) * N + g));
    const float s0 = gsr_exp(ld_param(P + ((CH_SCALE + 0) * N + g)));
    const float s1 = gsr_exp(ld_param(P + ((CH_SCALE + 1) * N + g)));
    const float s2 = gsr_exp(ld_param(P + ((CH_SCALE + 2) * N + g)));
    float qw = ld_param(P + ((CH_ROT + 0) * N + g)), qx = ld_param(P + ((CH_ROT + 1) * N + g));
    float qy = ld_param(P + ((CH_ROT + 2) * N + g)), qz = ld_param(P + ((CH_ROT + 3) * N + g));
    const float ql = sqrtf(((qx * qx + qy * qy) + qz * qz) + qw * qw);
    qx = qx / ql; qy = qy / ql; qz = qz / ql; qw = qw / ql;
    const float x2 = qx + qx, y2 = qy + qy, z2 = qz + qz;
    const float xx = qx * x2, xy_ = qx * y2, xz = qx * z2;
    const float yy = qy * y2, yz = qy * z2, zz = qz * z2;
    const float wx = qw * x2, wy = qw * y2, wz = qw * z2;
    // R (row, col) as quantToMat3 (preprocess.wgsl:182-201), times diag(s)
    const float m00 = (1.0f - (yy + zz)) * s0, m01 = (xy_ - wz) * s1, m02 = (xz + wy) * s2;
    const float m10 = (xy_ + wz) * s0, m11 = (1.0f - (xx + zz)) * s1, m12 = (yz - wx) * s2;
    const float m20 = (xz - wy) * s0, m21 = (yz + wx) * s1, m22 = (1.0f - (xx + yy)) * s2;
    // Sigma = M M^T; the products commute, so Sij == Sji exactly
    f.S00 = (m00 * m00 + m01 * m01) + m02 * m02;
    f.S01 = (m00 * m10 + m01 * m11) + m02 * m12;
    f.S02 = (m00 * m20 + m01 * m21) + m02 * m22;
    f.S11 = (m10 * m10 + m11 * m11) + m12 * m12;
    f.S12 = (m10 * m20 + m11 * m21) + m12 * m22;
    f.S22 = (m20 * m20 + m21 * m21) + m22 * m22;
    f.opacity = 1.0f / (1.0f + gsr_exp(-ld_param(P + (CH_OPACITY * N + g))));
    return f;
}

// The splat seen by one camera.  sh(i) returns SH coefficient i of the
// splat: i = ch for the DC terms, 3 + 15 ch + k for the rest.  Outputs as
// gsr_preprocess documents them.
template <typename SH>
__device__ __forceinline__ void splat_view(const SplatFixed& f, SH sh, const CamK& cam, uint32_t tile,
                                           uint32_t& ntiles, uint4& meta, float4& conic, float2& xy,
                                           float4& rgb) {
    const float px = f.px, py = f.py, pz = f.pz;
    const float* V = cam.view;
    const float* PV = cam.view_proj;
    const float tx = ((V[0] * px + V[4] * py) + V[8] * pz) + V[12];
    const float ty = ((V[1] * px + V[5] * py) + V[9] * pz) + V[13];
    const float tz = ((V[2] * px + V[6] * py) + V[10] * pz) + V[14];
    ntiles = 0;
    meta = make_uint4(0, 0, 0, 0);
    conic = make_float4(0.f, 0.f, 0.f, 0.f);
    xy = make_float2(0.f, 0.f);
    rgb = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tz < -kNearView) {
        const float cx = ((PV[0] * px + PV[4] * py) + PV[8] * pz) + PV[12];
        const float cy = ((PV[1] * px + PV[5] * py) + PV[9] * pz) + PV[13];
        const float cz = ((PV[2] * px + PV[6] * py) + PV[10] * pz) + PV[14];
        const float cw = ((PV[3] * px + PV[7] * py) + PV[11] * pz) + PV[15];
        const float w = cw + kSmall;
        const float ndcx = cx / w, ndcy = cy / w, ndcz = cz / w;
        const float W = cam.texture_size[0], H = cam.texture_size[1];
        const float pixx = ((ndcx + 1.0f) * W - 1.0f) * 0.5f;
        const float pixy = ((ndcy + 1.0f) * H - 1.0f) * 0.5f;
        xy = make_float2(pixx, pixy);
        const float S00 = f.S00, S01 = f.S01, S02 = f.S02, S11 = f.S11, S12 = f.S12, S22 = f.S22;
        const float S10 = S01, S20 = S02, S21 = S12;
        // calculateCovariance2D (preprocess.wgsl:251-269)
        const float limx = kFrustumGuard * cam.tan_fov[0];
        const float limy = kFrustumGuard * cam.tan_fov[1];
        const float ux = fminf(limx, fmaxf(-limx, tx / tz)) * tz;
        const float uy = fminf(limy, fmaxf(-limy, ty / tz)) * tz;
        const float uz = tz;
        const float fx = cam.focal[0], fy = cam.focal[1];
        const float uz2 = uz * uz;
        const float J00 = fx / uz, J02 = (-fx * ux) / uz2;
        const float J11 = fy / uz, J12 = (-fy * uy) / uz2;
        const float T00 = J00 * V[0] + J02 * V[2], T01 = J00 * V[4] + J02 * V[6], T02 = J00 * V[8] + J02 * V[10];
        const float T10 = J11 * V[1] + J12 * V[2], T11 = J11 * V[5] + J12 * V[6], T12 = J11 * V[9] + J12 * V[10];
        const float U00 = (T00 * S00 + T01 * S10) + T02 * S20;
        const float U01 = (T00 * S01 + T01 * S11) + T02 * S21;
        const float U02 = (T00 * S02 + T01 * S12) + T02 * S22;
        const float U10 = (T10 * S00 + T11 * S10) + T12 * S20;
        const float U11 = (T10 * S01 + T11 * S11) + T12 * S21;
        const float U12 = (T10 * S02 + T11 * S12) + T12 * S22;
        const float a = ((U00 * T00 + U01 * T01) + U02 * T02) + 0.3f;
        const float b = (U00 * T10 + U01 * T11) + U02 * T12;
        const float c = ((U10 * T10 + U11 * T11) + U12 * T12) + 0.3f;
        const float det = a * c - b * b;
        const float mid = 0.5f * (a + c);
        const float lambda = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
        const float radius = ceilf(3.0f * sqrtf(lambda));
        conic = make_float4(c / det, -b / det, a / det, f.opacity);
        if (!(pixx + radius < 0.0f || pixx - radius > W || pixy + radius < 0.0f || pixy - radius > H)) {
            // calculateColor (preprocess.wgsl:219-249)
            float dx = px - cam.eye[0], dy = py - cam.eye[1], dz = pz - cam.eye[2];
            const float dl = sqrtf((dx * dx + dy * dy) + dz * dz);
            dx = dx / dl; dy = dy / dl; dz = dz / dl;
            const float X = dx, Y = dy, Z = dz;
            const float XX = X * X, YY = Y * Y, ZZ = Z * Z, XY = X * Y, YZ = Y * Z, XZ = X * Z;
            float basis[15];
            basis[0] = -(SH_C1 * Y);
            basis[1] = SH_C1 * Z;
            basis[2] = -(SH_C1 * X);
            basis[3] = SH_C2_0 * XY;
            basis[4] = SH_C2_1 * YZ;
            basis[5] = SH_C2_2 * ((2.0f * ZZ - XX) - YY);
            basis[6] = SH_C2_3 * XZ;
            basis[7] = SH_C2_4 * (XX - YY);
            basis[8] = (SH_C3_0 * Y) * (3.0f * XX - YY);
            basis[9] = (SH_C3_1 * XY) * Z;
            basis[10] = (SH_C3_2 * Y) * ((4.0f * ZZ - XX) - YY);
            basis[11] = (SH_C3_3 * Z) * ((2.0f * ZZ - 3.0f * XX) - 3.0f * YY);
            basis[12] = (SH_C3_4 * X) * ((4.0f * ZZ - XX) - YY);
            basis[13] = (SH_C3_5 * Z) * (XX - YY);
            basis[14] = (SH_C3_6 * X) * (XX - 3.0f * YY);
            float col[3];
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                float acc = SH_C0 * sh(ch);
#pragma unroll
                for (int k = 0; k < 15; ++k) acc = __builtin_fmaf(basis[k], sh(3 + 15 * ch + k), acc);
                acc = acc + 0.5f;
                col[ch] = fmaxf(acc, 0.0f);
            }
            rgb = make_float4(col[0], col[1], col[2], 0.0f);
            // tile rectangle, depth16, tiles touched (preprocess.wgsl:122-143)
            const uint32_t Wi = (uint32_t)W, Hi = (uint32_t)H;
            const uint32_t TXn = (Wi + tile - 1) / tile, TYn = (Hi + tile - 1) / tile;
            const float ts = (float)tile;
            const float hx = (float)(TXn - 1), hy = (float)(TYn - 1);
            const uint32_t x0 = (uint32_t)fminf(fmaxf(floorf((pixx - radius) / ts), 0.0f), hx);
            const uint32_t y0 = (uint32_t)fminf(fmaxf(floorf((pixy - radius) / ts), 0.0f), hy);
            const uint32_t x1 = (uint32_t)fminf(fmaxf(floorf((pixx + radius) / ts), 0.0f), hx);
            const uint32_t y1 = (uint32_t)fminf(fmaxf(floorf((pixy + radius) / ts), 0.0f), hy);
            const float zsat = fminf(fmaxf(ndcz, 0.0f), 1.0f);
            const uint32_t depth16 = (uint32_t)(zsat * 65535.0f);
            ntiles = (x1 - x0 + 1) * (y1 - y0 + 1);
            meta = make_uint4((y0 << 16) | x0, (y1 << 16) | x1, depth16, 0u);
        }
    }
}

// SH coefficients held in registers.
struct ShRegs {
    const float* v;
    __device__ float operator()(int i) const { return v[i]; }
};

__device__ __forceinline__ SplatRec make_rec(const float4& conic, const float2& xy, const float4& rgb) {
    SplatRec r;
    r.co = conic;
    r.pq = make_float4(xy.x, xy.y, reach_qmax(conic.w), 0.0f);
    r.cl = rgb;
    r.pad = make_float4(0.f, 0.f, 0.f, 0.f);
    return r;
}

// The batch path's record store: the 16 bytes of padding are left unwritten
// (no kernel reads them).  6M splats x 8 views: 768 MB fewer bytes written,
// preprocess 1.83 -> 1.77 ms; the plain frame keeps the full store, which
// was 2 us faster there.
__device__ __forceinline__ void store_rec_unpadded(SplatRec* __restrict__ dst, const float4& conic, const float2& xy,
                                                   const float4& rgb) {
    dst->co = conic;
    dst->pq = make_float4(xy.x, xy.y, reach_qmax(conic.w), 0.0f);
    dst->cl = rgb;
}

// PACKED: write records (rec_out) instead of the three arrays.
template <bool PACKED>
__global__ __launch_bounds__(256) void k_preprocess(const float* __restrict__ P, uint32_t n, CamK cam,
                                                    uint32_t tile, uint32_t* __restrict__ tiles_out,
                                                    uint4* __restrict__ meta_out,
                                                    float4* __restrict__ conic_out,
                                                    float2* __restrict__ xy_out,
                                                    float4* __restrict__ rgb_out, SplatRec* __restrict__ rec_out,
                                                    ZeroRegions clear) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    zero_regions_part(clear, g, (size_t)gridDim.x * blockDim.x);
    if (g >= n) return;
    uint32_t ntiles;
    uint4 meta;
    float4 conic;
    float2 xy;
    float4 rgb;
    // All 48 SH coefficients are loaded up front with the other 11 values:
    // one memory round trip per thread instead of a second one after the
    // culls (a culled splat costs its 192 SH bytes; about 1% of a frame's)
    float shv[48];
#pragma unroll
    for (int i = 0; i < 48; ++i) shv[i] = ld_param(P + (size_t)(i < 3 ? CH_DC + i : CH_REST + i - 3) * n + g);
    splat_view(splat_fixed(P, n, g), ShRegs{shv}, cam, tile, ntiles, meta, conic, xy, rgb);
    tiles_out[g] = ntiles;
    meta_out[g] = meta;
    if (PACKED) {
        rec_out[g] = make_rec(conic, xy, rgb);
    } else {
        conic_out[g] = conic;
        xy_out[g] = xy;
        rgb_out[g] = rgb;
    }
}

__global__ __launch_bounds__(256) void k_pack(const float4* __restrict__ conic, const float2* __restrict__ xy,
                                              const float4* __restrict__ rgb, uint32_t n,
                                              SplatRec* __restrict__ rec) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    SplatRec r;
    r.co = conic[g];
    const float2 p = xy[g];
    r.pq = make_float4(p.x, p.y, reach_qmax(r.co.w), 0.0f);
    r.cl = rgb[g];
    r.pad = make_float4(0.f, 0.f, 0.f, 0.f);
    rec[g] = r;
}

// View batches: a workgroup takes 64 splats.  Their view-independent parts
// (position, covariance, opacity) and SH coefficients are computed and staged
// in LDS once; then thread (s, q) = (tid % 64, tid / 64) projects splat s for
// views q, q + 4, q + 8, ...  Outputs of view v sit at v * n + g, so each
// wave writes one view's 64 consecutive records.  (6M splats x 8 views:
// 2.00 ms with 32 splats x 8 phases, 1.83 with 64 x 4, 1.93 with 128 x 2,
// 2.77 with 256 x 1.)
constexpr int kBatchSplats = 64;
constexpr int kBatchLanes = 4;  // view phases per splat

struct ShShared {
    const float* sh;  // [48][kBatchSplats]
    uint32_t s;
    __device__ float operator()(int i) const { return sh[i * kBatchSplats + s]; }
};

__global__ __launch_bounds__(256) void k_preprocess_batch(const float* __restrict__ P, uint32_t n,
                                                          const CamK* __restrict__ cams, uint32_t views,
                                                          uint32_t tile, uint32_t* __restrict__ tiles_out,
                                                          uint4* __restrict__ meta_out,
                                                          SplatRec* __restrict__ rec_out) {
    __shared__ float s_sh[48 * kBatchSplats];
    __shared__ SplatFixed s_fixed[kBatchSplats];
    const uint32_t tid = threadIdx.x;
    const uint32_t g0 = blockIdx.x * kBatchSplats;
    const uint32_t count = min((uint32_t)kBatchSplats, n - g0);
    for (uint32_t i = tid; i < 48 * kBatchSplats; i += blockDim.x) {
        const uint32_t k = i / kBatchSplats, s = i % kBatchSplats;
        const uint32_t ch = k < 3 ? CH_DC + k : CH_REST + k - 3;
        s_sh[i] = s < count ? ld_param(P + (size_t)ch * n + g0 + s) : 0.0f;
    }
    if (tid < count) s_fixed[tid] = splat_fixed(P, n, g0 + tid);
    __syncthreads();
    const uint32_t s = tid % kBatchSplats, q = tid / kBatchSplats;
    if (s >= count) return;
    const uint32_t g = g0 + s;
    const SplatFixed f = s_fixed[s];
    for (uint32_t v = q; v < views; v += kBatchLanes) {
        uint32_t ntiles;
        uint4 meta;
        float4 conic;
        float2 xy;
        float4 rgb;
        splat_view(f, ShShared{s_sh, s}, cams[v], tile, ntiles, meta, conic, xy, rgb);
        const size_t o = (size_t)v * n + g;
        tiles_out[o] = ntiles;
        meta_out[o] = meta;
        store_rec_unpadded(rec_out + o, conic, xy, rgb);
    }
}

}  // namespace

int launch_preprocess_batch(Context* c, const float* params, uint32_t n, const CamK* cams, uint32_t views,
                            uint32_t tile, uint32_t* tiles, uint4* meta, SplatRec* rec) {
    if (n == 0 || views == 0) return GSR_OK;
    hipLaunchKernelGGL(k_preprocess_batch, dim3(ceil_div(n, kBatchSplats)), dim3(kBatchSplats * kBatchLanes), 0,
                       c->stream, params, n, cams, views, tile, tiles, meta, rec);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

int launch_preprocess(Context* c, const float* params, uint32_t n, const CamK& cam, uint32_t tile,
                      uint32_t* tiles, uint4* meta, float4* conic, float2* xy, float4* rgb, SplatRec* rec,
                      const FillRegion* clear, int nclear) {
    if (nclear < 0 || nclear > 4) return fail(c, GSR_ERR_ARGUMENT, "preprocess: %d regions to clear", nclear);
    if (n == 0) return clear && nclear ? launch_zero_regions(c, clear, nclear) : GSR_OK;
    const ZeroRegions z = make_zero_regions(clear, clear ? nclear : 0);
    if (rec)
        hipLaunchKernelGGL(k_preprocess<true>, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, params, n, cam, tile,
                           tiles, meta, nullptr, nullptr, nullptr, rec, z);
    else
        hipLaunchKernelGGL(k_preprocess<false>, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, params, n, cam, tile,
                           tiles, meta, conic, xy, rgb, nullptr, z);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

int launch_pack_splats(Context* c, const float4* conic, const float2* xy, const float4* rgb, uint32_t n,
                       SplatRec* rec) {
    if (n == 0) return GSR_OK;
    hipLaunchKernelGGL(k_pack, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, conic, xy, rgb, n, rec);
    GSR_HIP(c, hipGetLastError());
    return GSR_OK;
}

}  // namespace gsr
