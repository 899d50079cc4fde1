// gsr_oracle.cpp — CPU restatement of the reference splat pass graph.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
// renderer: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it.  The product never calls it.
//
// It restates, in FP32 and FP64, the per-frame pipeline of
// KoS-Y1/Web-GS-Renderer:
//   preprocess   src/shaders/preprocess.wgsl:58-284
//   offset scan  src/shaders/offset_scan.wgsl:22-68   (exclusive, N+1 entries)
//   emit         src/shaders/emit.wgsl:28-67          (row-major slots, cap)
//   sizing       src/shaders/indirect_arg.wgsl:14-28  (n = min(I, cap))
//   sort         src/shaders/radix_parallel.wgsl:26-160 (stable LSD on the key)
//   tile ranges  src/shaders/tile_ranges.wgsl:7-35    (untouched tiles (0,0))
//   raster       src/shaders/raseter.wgsl:35-117      (front-to-back blend)
//   pass order   src/renderer/renderer.js:822-932
//
// FP32 arithmetic order.  The WGSL leaves the rounding of exp and the
// contraction of a*b+c to the WebGPU implementation.  This restatement fixes
// one order: no implicit contraction (built with -ffp-contract=off), fmaf only
// where written, IEEE sqrt and division, and exp from exp_f32() below
// (Cody-Waite reduction, degree-7 Taylor polynomial).  The HIP kernels follow
// the same order, so their FP32 outputs are comparable bit for bit.  FP64
// versions (libm exp, same formulas) measure the FP32 error.
//
// Pins (tests/test_oracle.py): exp_f32 against libm within 2 ulp; the integer
// chain against closed forms / brute force; preprocess special cases
// (isotropic splat, on-axis Jacobian, SH degree 0, clamp, sigmoid(0));
// FP32 vs FP64 agreement; raster closed forms on 1-3 splats.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// The 176-byte GlobalUniforms block (preprocess.wgsl:2-12).
struct Cam {
    float view[16];
    float view_proj[16];
    float eye[3];
    uint32_t count;
    float focal[2];
    float tan_fov[2];
    float texture_size[2];
    uint32_t pad[2];
};
static_assert(sizeof(Cam) == 176, "GlobalUniforms is 176 bytes");

// exp for FP32: x = n*ln2 + r, |r| <= ln2/2 (Cody & Waite, "Software Manual
// for the Elementary Functions", 1980), e^r by its degree-7 Taylor polynomial
// in Horner form with fused multiply-adds, scaled by 2^n.  Results below the
// normal range are returned as 0.
inline float exp_f32(float x) {
    if (x != x) return x;
    if (x > 88.8f) return INFINITY;
    if (x < -87.0f) return 0.0f;
    const float n = rintf(x * 0x1.715476p+0f);          // log2(e)
    float r = fmaf(n, -0x1.62e430p-1f, x);               // ln2, high part
    r = fmaf(n, 0x1.05c610p-29f, r);                     // ln2 - high part (negated)
    float p = 0x1.a01a02p-13f;                           // 1/5040
    p = fmaf(p, r, 0x1.6c16c2p-10f);                     // 1/720
    p = fmaf(p, r, 0x1.111112p-7f);                      // 1/120
    p = fmaf(p, r, 0x1.555556p-5f);                      // 1/24
    p = fmaf(p, r, 0x1.555556p-3f);                      // 1/6
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)n);
}

template <typename T> struct M;
template <> struct M<float> {
    static float ex(float x) { return exp_f32(x); }
    static float fma(float a, float b, float c) { return fmaf(a, b, c); }
};
template <> struct M<double> {
    static double ex(double x) { return std::exp(x); }
    static double fma(double a, double b, double c) { return std::fma(a, b, c); }
};

// SH constants (preprocess.wgsl:41-54).
constexpr double SH_C0 = 0.28209479177387814;
constexpr double SH_C1 = 0.4886025119029199;
constexpr double SH_C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                             -1.0925484305920792, 0.5462742152960396};
constexpr double SH_C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                             0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                             -0.5900435899266435};

template <typename T> struct Pre {
    uint32_t tiles;
    uint32_t meta[4];
    T conic[4];
    T xy[2];
    T rgb[4];
};

// Packed block layout of src/loader/loader.js:75-85.
struct Packed {
    const float* base;
    uint32_t n;
    float pos(uint32_t g, int k) const { return base[(size_t)g * 3 + k]; }
    float lscale(uint32_t g, int k) const { return base[(size_t)n * 3 + (size_t)g * 3 + k]; }
    float rot(uint32_t g, int k) const { return base[(size_t)n * 6 + (size_t)g * 4 + k]; }
    float dc(uint32_t g, int k) const { return base[(size_t)n * 10 + (size_t)g * 3 + k]; }
    float opacity(uint32_t g) const { return base[(size_t)n * 13 + g]; }
    // getShRest (preprocess.wgsl:159-162): coefficient c of channel ch.
    float rest(uint32_t g, int c, int ch) const {
        return base[(size_t)n * 14 + (size_t)g * 45 + c + 15 * ch];
    }
};

// preprocess.wgsl:58-144 for one splat.
template <typename T>
void preprocess_one(const Packed& P, uint32_t g, const Cam& cam, uint32_t tile, Pre<T>& o) {
    std::memset(&o, 0, sizeof(o));
    const T px = P.pos(g, 0), py = P.pos(g, 1), pz = P.pos(g, 2);
    const float* V = cam.view;
    const float* PV = cam.view_proj;
    // positionView = view * (p, 1)   (:80)
    const T tx = ((T(V[0]) * px + T(V[4]) * py) + T(V[8]) * pz) + T(V[12]);
    const T ty = ((T(V[1]) * px + T(V[5]) * py) + T(V[9]) * pz) + T(V[13]);
    const T tz = ((T(V[2]) * px + T(V[6]) * py) + T(V[10]) * pz) + T(V[14]);
    // near cull (:83-85)
    if (tz >= T(-0.2f)) return;
    // clip -> ndc -> pixel (:87-93)
    const T cx = ((T(PV[0]) * px + T(PV[4]) * py) + T(PV[8]) * pz) + T(PV[12]);
    const T cy = ((T(PV[1]) * px + T(PV[5]) * py) + T(PV[9]) * pz) + T(PV[13]);
    const T cz = ((T(PV[2]) * px + T(PV[6]) * py) + T(PV[10]) * pz) + T(PV[14]);
    const T cw = ((T(PV[3]) * px + T(PV[7]) * py) + T(PV[11]) * pz) + T(PV[15]);
    const T w = cw + T(1e-7f);
    const T ndcx = cx / w, ndcy = cy / w, ndcz = cz / w;
    const T W = T(cam.texture_size[0]), H = T(cam.texture_size[1]);
    const T pixx = ((ndcx + T(1)) * W - T(1)) * T(0.5);
    const T pixy = ((ndcy + T(1)) * H - T(1)) * T(0.5);
    o.xy[0] = pixx;
    o.xy[1] = pixy;
    // scale and rotation (:95-100, quantToMat3 :182-201)
    const T s[3] = {M<T>::ex(T(P.lscale(g, 0))), M<T>::ex(T(P.lscale(g, 1))),
                    M<T>::ex(T(P.lscale(g, 2)))};
    T qx = P.rot(g, 1), qy = P.rot(g, 2), qz = P.rot(g, 3), qw = P.rot(g, 0);
    const T ql = std::sqrt(((qx * qx + qy * qy) + qz * qz) + qw * qw);
    qx = qx / ql; qy = qy / ql; qz = qz / ql; qw = qw / ql;
    const T x2 = qx + qx, y2 = qy + qy, z2 = qz + qz;
    const T xx = qx * x2, xy = qx * y2, xz = qx * z2;
    const T yy = qy * y2, yz = qy * z2, zz = qz * z2;
    const T wx = qw * x2, wy = qw * y2, wz = qw * z2;
    T R[3][3];  // R[row][col]
    R[0][0] = T(1) - (yy + zz); R[1][0] = xy + wz;           R[2][0] = xz - wy;
    R[0][1] = xy - wz;          R[1][1] = T(1) - (xx + zz);  R[2][1] = yz + wx;
    R[0][2] = xz + wy;          R[1][2] = yz - wx;           R[2][2] = T(1) - (xx + yy);
    // Sigma = (R S)(R S)^T
    T Mm[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Mm[r][c] = R[r][c] * s[c];
    T S[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            S[i][j] = (Mm[i][0] * Mm[j][0] + Mm[i][1] * Mm[j][1]) + Mm[i][2] * Mm[j][2];
    // calculateCovariance2D (:251-269)
    const T limx = T(1.3f) * T(cam.tan_fov[0]);
    const T limy = T(1.3f) * T(cam.tan_fov[1]);
    const T ux = std::min(limx, std::max(-limx, tx / tz)) * tz;
    const T uy = std::min(limy, std::max(-limy, ty / tz)) * tz;
    const T uz = tz;
    const T fx = cam.focal[0], fy = cam.focal[1];
    const T uz2 = uz * uz;
    const T J00 = fx / uz, J02 = (-fx * ux) / uz2;
    const T J11 = fy / uz, J12 = (-fy * uy) / uz2;
    // T = J * Wv with Wv[r][c] = view[c*4+r]
    T Tm[2][3];
    for (int c = 0; c < 3; ++c) {
        Tm[0][c] = J00 * T(V[c * 4 + 0]) + J02 * T(V[c * 4 + 2]);
        Tm[1][c] = J11 * T(V[c * 4 + 1]) + J12 * T(V[c * 4 + 2]);
    }
    T U[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c)
            U[r][c] = (Tm[r][0] * S[0][c] + Tm[r][1] * S[1][c]) + Tm[r][2] * S[2][c];
    const T a = ((U[0][0] * Tm[0][0] + U[0][1] * Tm[0][1]) + U[0][2] * Tm[0][2]) + T(0.3f);
    const T b = (U[0][0] * Tm[1][0] + U[0][1] * Tm[1][1]) + U[0][2] * Tm[1][2];
    const T c = ((U[1][0] * Tm[1][0] + U[1][1] * Tm[1][1]) + U[1][2] * Tm[1][2]) + T(0.3f);
    const T det = a * c - b * b;
    // calculateRadius (:271-275)
    const T mid = T(0.5) * (a + c);
    const T lambda = mid + std::sqrt(std::max(T(0.1f), mid * mid - det));
    const T radius = std::ceil(T(3) * std::sqrt(lambda));
    o.conic[0] = c / det;
    o.conic[1] = -b / det;
    o.conic[2] = a / det;
    o.conic[3] = T(1) / (T(1) + M<T>::ex(-T(P.opacity(g))));
    // isInFrustum (:277-284)
    if (pixx + radius < T(0) || pixx - radius > W || pixy + radius < T(0) || pixy - radius > H)
        return;
    // calculateColor (:219-249), d = normalize(p - eye)
    T dx = px - T(cam.eye[0]), dy = py - T(cam.eye[1]), dz = pz - T(cam.eye[2]);
    const T dl = std::sqrt((dx * dx + dy * dy) + dz * dz);
    dx = dx / dl; dy = dy / dl; dz = dz / dl;
    const T X = dx, Y = dy, Z = dz;
    const T XX = X * X, YY = Y * Y, ZZ = Z * Z, XY = X * Y, YZ = Y * Z, XZ = X * Z;
    T basis[15];
    basis[0] = -(T(float(SH_C1)) * Y);
    basis[1] = T(float(SH_C1)) * Z;
    basis[2] = -(T(float(SH_C1)) * X);
    basis[3] = T(float(SH_C2[0])) * XY;
    basis[4] = T(float(SH_C2[1])) * YZ;
    basis[5] = T(float(SH_C2[2])) * ((T(2) * ZZ - XX) - YY);
    basis[6] = T(float(SH_C2[3])) * XZ;
    basis[7] = T(float(SH_C2[4])) * (XX - YY);
    basis[8] = (T(float(SH_C3[0])) * Y) * (T(3) * XX - YY);
    basis[9] = (T(float(SH_C3[1])) * XY) * Z;
    basis[10] = (T(float(SH_C3[2])) * Y) * ((T(4) * ZZ - XX) - YY);
    basis[11] = (T(float(SH_C3[3])) * Z) * ((T(2) * ZZ - T(3) * XX) - T(3) * YY);
    basis[12] = (T(float(SH_C3[4])) * X) * ((T(4) * ZZ - XX) - YY);
    basis[13] = (T(float(SH_C3[5])) * Z) * (XX - YY);
    basis[14] = (T(float(SH_C3[6])) * X) * (XX - T(3) * YY);
    for (int ch = 0; ch < 3; ++ch) {
        T acc = T(float(SH_C0)) * T(P.dc(g, ch));
        for (int k = 0; k < 15; ++k) acc = M<T>::fma(basis[k], T(P.rest(g, k, ch)), acc);
        acc = acc + T(0.5);
        o.rgb[ch] = std::max(acc, T(0));
    }
    // tile rect, depth16, tiles touched (:122-143)
    const uint32_t Wi = (uint32_t)cam.texture_size[0], Hi = (uint32_t)cam.texture_size[1];
    const uint32_t TXn = (Wi + tile - 1) / tile, TYn = (Hi + tile - 1) / tile;
    const T ts = T(tile);
    auto tclamp = [](T v, T hi) { return std::min(std::max(v, T(0)), hi); };
    const uint32_t x0 = (uint32_t)tclamp(std::floor((pixx - radius) / ts), T(TXn - 1));
    const uint32_t y0 = (uint32_t)tclamp(std::floor((pixy - radius) / ts), T(TYn - 1));
    const uint32_t x1 = (uint32_t)tclamp(std::floor((pixx + radius) / ts), T(TXn - 1));
    const uint32_t y1 = (uint32_t)tclamp(std::floor((pixy + radius) / ts), T(TYn - 1));
    const T zsat = std::min(std::max(ndcz, T(0)), T(1));
    const uint32_t depth16 = (uint32_t)(zsat * T(65535));
    o.tiles = (x1 - x0 + 1) * (y1 - y0 + 1);
    o.meta[0] = (y0 << 16) | x0;
    o.meta[1] = (y1 << 16) | x1;
    o.meta[2] = depth16;
    o.meta[3] = 0;
}

template <typename T>
void preprocess_all(const float* packed, uint32_t n, const Cam& cam, uint32_t tile,
                    uint32_t* tiles, uint32_t* meta, T* conic, T* xy, T* rgb) {
    Packed P{packed, n};
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < (int64_t)n; ++g) {
        Pre<T> o;
        preprocess_one<T>(P, (uint32_t)g, cam, tile, o);
        tiles[g] = o.tiles;
        for (int k = 0; k < 4; ++k) meta[g * 4 + k] = o.meta[k];
        for (int k = 0; k < 4; ++k) conic[g * 4 + k] = o.conic[k];
        xy[g * 2] = o.xy[0];
        xy[g * 2 + 1] = o.xy[1];
        for (int k = 0; k < 4; ++k) rgb[g * 4 + k] = o.rgb[k];
    }
}

// offset_scan.wgsl:22-68 — exclusive prefix sum; total returned.
uint64_t scan_exclusive(const uint32_t* in, uint32_t n, uint32_t* out) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = in[i];
        out[i] = (uint32_t)acc;
        acc += v;
    }
    return acc;
}

// emit.wgsl:28-67 — slots offsets[g]+k, row-major over the tile rect.
void emit_all(const uint32_t* offsets, const uint32_t* meta, uint32_t n, uint32_t width,
              uint32_t tile, uint64_t cap, uint64_t* keys, uint32_t* vals) {
    const uint32_t TXn = (width + tile - 1) / tile;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t g = 0; g < (int64_t)n; ++g) {
        const uint64_t base = offsets[g];
        const uint32_t cnt = offsets[g + 1] - offsets[g];
        if (cnt == 0) continue;
        const uint32_t m0 = meta[g * 4], m1 = meta[g * 4 + 1];
        const uint32_t x0 = m0 & 0xFFFF, y0 = m0 >> 16, x1 = m1 & 0xFFFF, y1 = m1 >> 16;
        const uint64_t d16 = meta[g * 4 + 2] & 0xFFFF;
        uint64_t slot = base;
        for (uint32_t ty = y0; ty <= y1; ++ty)
            for (uint32_t tx = x0; tx <= x1; ++tx) {
                if (slot >= cap) goto next;
                keys[slot] = ((uint64_t)(ty * TXn + tx) << 16) | d16;
                vals[slot] = (uint32_t)g;
                ++slot;
            }
    next:;
    }
}

// Stable sort by key (radix_parallel.wgsl is a stable LSD sort, so the result
// equals the (key, value)-lexicographic order of the emitted array).
void sort_stable(uint64_t* keys, uint32_t* vals, uint64_t n) {
    std::vector<uint64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return keys[a] < keys[b]; });
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> v2(n);
    for (uint64_t i = 0; i < n; ++i) { k2[i] = keys[idx[i]]; v2[i] = vals[idx[i]]; }
    std::memcpy(keys, k2.data(), n * 8);
    std::memcpy(vals, v2.data(), n * 4);
}

// Parallel stable LSD radix sort (8-bit digits over [0, end_bit)), used by
// the timed CPU pipeline; tests check it equals sort_stable.
void sort_radix_parallel(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t end_bit) {
    std::vector<uint64_t> kb(n);
    std::vector<uint32_t> vb(n);
    uint64_t* ks = keys; uint32_t* vs = vals;
    uint64_t* kd = kb.data(); uint32_t* vd = vb.data();
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<uint64_t> hist((size_t)nt * 256);
    for (uint32_t shift = 0; shift < end_bit; shift += 8) {
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
            uint64_t* h = &hist[(size_t)t * 256];
            for (uint64_t i = b; i < e; ++i) ++h[(ks[i] >> shift) & 0xFF];
#pragma omp barrier
#pragma omp single
            {
                uint64_t acc = 0;
                for (int d = 0; d < 256; ++d)
                    for (int tt = 0; tt < nt; ++tt) {
                        const uint64_t c = hist[(size_t)tt * 256 + d];
                        hist[(size_t)tt * 256 + d] = acc;
                        acc += c;
                    }
            }
            for (uint64_t i = b; i < e; ++i) {
                const uint64_t p = h[(ks[i] >> shift) & 0xFF]++;
                kd[p] = ks[i];
                vd[p] = vs[i];
            }
        }
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    if (ks != keys) {
        std::memcpy(keys, ks, n * 8);
        std::memcpy(vals, vs, n * 4);
    }
}

// tile_ranges.wgsl:7-35 on sorted keys, tile = key >> 16.
void tile_ranges(const uint64_t* keys, uint64_t n, uint32_t num_tiles, uint32_t* ranges) {
    std::memset(ranges, 0, (size_t)num_tiles * 8);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t t = keys[i] >> 16;
        if (i == 0) {
            ranges[t * 2] = 0;
        } else {
            const uint64_t p = keys[i - 1] >> 16;
            if (t > p) {
                ranges[p * 2 + 1] = (uint32_t)i;
                ranges[t * 2] = (uint32_t)i;
            }
        }
        if (i == n - 1) ranges[t * 2 + 1] = (uint32_t)n;
    }
}

// raseter.wgsl:35-117 for every pixel: integer pixel coordinates, blend in
// range order, alpha clamp 0.99, cutoff 1/255, stop before T < 1e-4.
// last[] is the 1-based range position of the last blended splat.
template <typename T>
void raster_all(uint32_t W, uint32_t H, uint32_t tile, const uint32_t* ranges, const uint32_t* vals,
                const T* conic, const T* xy, const T* rgb, T* out, T* tfinal, uint32_t* last) {
    const uint32_t TXn = (W + tile - 1) / tile, TYn = (H + tile - 1) / tile;
    const int64_t ntile = (int64_t)TXn * TYn;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < ntile; ++t) {
        const uint32_t tx = (uint32_t)(t % TXn), ty = (uint32_t)(t / TXn);
        const uint32_t start = ranges[t * 2], end = ranges[t * 2 + 1];
        for (uint32_t ly = 0; ly < tile; ++ly)
            for (uint32_t lx = 0; lx < tile; ++lx) {
                const uint32_t x = tx * tile + lx, y = ty * tile + ly;
                if (x >= W || y >= H) continue;
                const T fxp = T(x), fyp = T(y);
                T Tr = 1, c0 = 0, c1 = 0, c2 = 0;
                uint32_t lastc = 0;
                for (uint32_t i = start; i < end; ++i) {
                    const uint32_t g = vals[i];
                    const T A = conic[g * 4], B = conic[g * 4 + 1], C = conic[g * 4 + 2];
                    const T op = conic[g * 4 + 3];
                    const T dx = xy[g * 2] - fxp, dy = xy[g * 2 + 1] - fyp;
                    // power = -0.5 (A dx^2 + C dy^2) - B dx dy  (calculatePower :113-117)
                    const T q = M<T>::fma(A * dx, dx, (C * dy) * dy);
                    const T power = M<T>::fma(T(-0.5), q, -((B * dx) * dy));
                    if (power > T(0)) continue;
                    const T alpha = std::min(T(0.99f), op * M<T>::ex(power));
                    if (alpha < T(1.0f / 255.0f)) continue;
                    const T testT = Tr * (T(1) - alpha);
                    if (testT < T(0.0001f)) break;
                    const T wgt = alpha * Tr;
                    c0 = M<T>::fma(rgb[g * 4], wgt, c0);
                    c1 = M<T>::fma(rgb[g * 4 + 1], wgt, c1);
                    c2 = M<T>::fma(rgb[g * 4 + 2], wgt, c2);
                    Tr = testT;
                    lastc = i - start + 1;
                }
                const size_t p = (size_t)y * W + x;
                out[p] = c0;
                out[(size_t)W * H + p] = c1;
                out[(size_t)2 * W * H + p] = c2;
                if (tfinal) tfinal[p] = Tr;
                if (last) last[p] = lastc;
            }
    }
}

template <typename T>
int render(const float* packed, uint32_t n, const Cam& cam, uint32_t tile, uint64_t cap, T* out,
           T* tfinal, uint32_t* last, uint64_t* instances, double* stage_ms, int use_radix) {
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const uint32_t W = (uint32_t)cam.texture_size[0], H = (uint32_t)cam.texture_size[1];
    const uint32_t TXn = (W + tile - 1) / tile, TYn = (H + tile - 1) / tile;
    std::vector<uint32_t> tiles(n), meta((size_t)n * 4), offsets((size_t)n + 1);
    std::vector<T> conic((size_t)n * 4), xy((size_t)n * 2), rgb((size_t)n * 4);
    auto t0 = clk::now();
    preprocess_all<T>(packed, n, cam, tile, tiles.data(), meta.data(), conic.data(), xy.data(), rgb.data());
    auto t1 = clk::now();
    std::vector<uint32_t> counts(tiles);
    counts.push_back(0);
    const uint64_t total = scan_exclusive(counts.data(), n + 1, offsets.data());
    auto t2 = clk::now();
    if (cap == 0) cap = (uint64_t)n * 64;
    const uint64_t m = std::min(total, cap);
    std::vector<uint64_t> keys(m);
    std::vector<uint32_t> vals(m);
    emit_all(offsets.data(), meta.data(), n, W, tile, cap, keys.data(), vals.data());
    auto t3 = clk::now();
    uint32_t tbits = 0;
    while ((1ull << tbits) < (uint64_t)TXn * TYn) ++tbits;
    if (use_radix)
        sort_radix_parallel(keys.data(), vals.data(), m, 16 + tbits);
    else
        sort_stable(keys.data(), vals.data(), m);
    auto t4 = clk::now();
    std::vector<uint32_t> ranges((size_t)TXn * TYn * 2);
    tile_ranges(keys.data(), m, TXn * TYn, ranges.data());
    auto t5 = clk::now();
    raster_all<T>(W, H, tile, ranges.data(), vals.data(), conic.data(), xy.data(), rgb.data(), out,
                  tfinal, last);
    auto t6 = clk::now();
    if (instances) *instances = total;
    if (stage_ms) {
        stage_ms[0] = ms(t0, t1);
        stage_ms[1] = ms(t1, t2);
        stage_ms[2] = ms(t2, t3);
        stage_ms[3] = ms(t3, t4);
        stage_ms[4] = ms(t4, t5);
        stage_ms[5] = ms(t5, t6);
    }
    return total > cap ? 1 : 0;
}

}  // namespace

extern "C" {

float orc_exp_f32(float x) { return exp_f32(x); }

void orc_exp_f32_array(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = exp_f32(x[i]);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void orc_preprocess_f32(const float* packed, uint32_t n, const void* cam, uint32_t tile,
                        uint32_t* tiles, uint32_t* meta, float* conic, float* xy, float* rgb) {
    preprocess_all<float>(packed, n, *(const Cam*)cam, tile, tiles, meta, conic, xy, rgb);
}

void orc_preprocess_f64(const float* packed, uint32_t n, const void* cam, uint32_t tile,
                        uint32_t* tiles, uint32_t* meta, double* conic, double* xy, double* rgb) {
    preprocess_all<double>(packed, n, *(const Cam*)cam, tile, tiles, meta, conic, xy, rgb);
}

uint64_t orc_scan(const uint32_t* in, uint32_t n, uint32_t* out) { return scan_exclusive(in, n, out); }

void orc_emit(const uint32_t* offsets, const uint32_t* meta, uint32_t n, uint32_t width, uint32_t tile,
              uint64_t cap, uint64_t* keys, uint32_t* vals) {
    emit_all(offsets, meta, n, width, tile, cap, keys, vals);
}

void orc_sort_stable(uint64_t* keys, uint32_t* vals, uint64_t n) { sort_stable(keys, vals, n); }

void orc_sort_radix(uint64_t* keys, uint32_t* vals, uint64_t n, uint32_t end_bit) {
    sort_radix_parallel(keys, vals, n, end_bit);
}

void orc_tile_ranges(const uint64_t* keys, uint64_t n, uint32_t num_tiles, uint32_t* ranges) {
    tile_ranges(keys, n, num_tiles, ranges);
}

void orc_raster_f32(uint32_t W, uint32_t H, uint32_t tile, const uint32_t* ranges, const uint32_t* vals,
                    const float* conic, const float* xy, const float* rgb, float* out, float* tfinal,
                    uint32_t* last) {
    raster_all<float>(W, H, tile, ranges, vals, conic, xy, rgb, out, tfinal, last);
}

void orc_raster_f64(uint32_t W, uint32_t H, uint32_t tile, const uint32_t* ranges, const uint32_t* vals,
                    const double* conic, const double* xy, const double* rgb, double* out,
                    double* tfinal, uint32_t* last) {
    raster_all<double>(W, H, tile, ranges, vals, conic, xy, rgb, out, tfinal, last);
}

int orc_render_f32(const float* packed, uint32_t n, const void* cam, uint32_t tile, uint64_t cap,
                   float* out, float* tfinal, uint32_t* last, uint64_t* instances, double* stage_ms,
                   int use_radix) {
    return render<float>(packed, n, *(const Cam*)cam, tile, cap, out, tfinal, last, instances, stage_ms,
                         use_radix);
}

int orc_render_f64(const float* packed, uint32_t n, const void* cam, uint32_t tile, uint64_t cap,
                   double* out, double* tfinal, uint32_t* last, uint64_t* instances) {
    return render<double>(packed, n, *(const Cam*)cam, tile, cap, out, tfinal, last, instances, nullptr,
                          0);
}

}  // extern "C"
