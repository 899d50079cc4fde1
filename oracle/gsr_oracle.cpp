// gsr_oracle.cpp — CPU restatement of the reference splat pass graph.
//
// TEST INFRASTRUCTURE ONLY.  This library is the CPU reference the HIP
// renderer is checked against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
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
// where written, IEEE sqrt and division, exp from exp_f32() below
// (Cody-Waite reduction, degree-7 Taylor polynomial) in preprocess and from
// exp_blend_f32() (exp2 reduction, degree-5 polynomial, for powers in
// [-5.55, 0]) in the blend, whose power and transmittance follow
// power_f32() and transmit_f32().  The HIP kernels follow the same order, so their
// FP32 outputs are comparable bit for bit.  FP64 versions (libm exp, same
// formulas) measure the FP32 error.
//
// Pins (tests/test_oracle.py): exp_f32 against libm within 2 ulp, exp_blend_f32
// within 2.7 ulp on [-5.55, 0] and its ldexp equal, on every float there, to
// the kernels' exponent insertion by shift-and-add; the FP32 blend decision (blend_range_f32 and
// the alpha tests) against FP64 libm over every power from 0 down to -inf,
// and a raster of far thin splats against the FP64 raster; power_f32 exact on
// small integers and within 6 FP32 roundings of calculatePower in FP64,
// transmit_f32 bit-equal to the rounded exact T (1 - alpha); the integer
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

// e^p for the blend's alpha, p in [-5.55, 0] (raster only; the GPU kernels
// use the same steps): 2^(p log2 e), k = p log2 e rounded to an integer by
// adding 1.5 * 2^23, f = fma(p, log2 e, -k), 2^f by a degree-5 polynomial
// (coefficients fitted by iteratively reweighted least squares for relative
// error on [-1/2, 1/2], oracle/exp_blend_fit.py), times 2^k.  Pinned within
// 2.7 ulp of libm's exp on [-5.55, 0] (tests/test_oracle.py).
// The two parts of exp_blend_f32: q = 2^f and the bits of t = 1.5 * 2^23 + k.
inline void exp_blend_parts(float p, float& q, uint32_t& tu) {
    const float kLog2e = 0x1.715476p+0f, kRound = 0x1.8p+23f;
    const float t = fmaf(p, kLog2e, kRound);
    const float f = fmaf(p, kLog2e, -(t - kRound));
    q = 0x1.5bba1p-10f;
    q = fmaf(q, f, 0x1.3cea88p-7f);
    q = fmaf(q, f, 0x1.c6b752p-5f);
    q = fmaf(q, f, 0x1.ebf9bcp-3f);
    q = fmaf(q, f, 0x1.62e42ap-1f);
    q = fmaf(q, f, 1.0f);
    std::memcpy(&tu, &t, 4);
}

inline float exp_blend_f32(float p) {
    float q;
    uint32_t tu;
    exp_blend_parts(p, q, tu);
    return ldexpf(q, (int32_t)(tu - 0x4B400000u));  // 0x4B400000: bits of 1.5 * 2^23
}

// Which powers the FP32 blend evaluates.  raseter.wgsl:84-92 skips a splat
// at a pixel when power > 0, and when alpha = min(0.99, op e^power) < 1/255.
// The opacity is a sigmoid (op <= 1) and e^-5.55 = 0.003887 < 1/255 =
// 0.003922, so every power below -5.55 is skipped by the alpha test as well.
// The FP32 paths (raster, the FP32 decision replay of the adjoint, the FP32
// backward twin) drop those powers before the exponential, exactly as the
// GPU kernels do: exp_blend_f32 is accurate on [-5.55, 0] only, and below
// about -8.7e6 its 2^k scale wraps (inf, NaN at -inf).  A NaN power (from a
// non-finite conic) passes both tests, as in the kernels.  The FP64 paths
// keep the reference's tests as written (libm exp is exact enough there),
// which is what pins this cut (tests/test_oracle.py).
inline bool blend_range_f32(float power) { return !(power > 0.0f) && !(power < -5.55f); }

// The FP32 power of a splat at pixel offset (dx, dy) = (splat - pixel):
// calculatePower (raseter.wgsl:113-117), -0.5 (A dx^2 + C dy^2) - B dx dy,
// evaluated as dx (a dx + b dy) + c dy^2 with a = -A/2, b = -B, c = -C/2
// (scalings by -1/2 and -1, exact), two FMAs and three products, the order
// the kernels use.
inline float power_f32(float A, float B, float C, float dx, float dy) {
    return fmaf(dx, fmaf(-0.5f * A, dx, (-B) * dy), ((-0.5f * C) * dy) * dy);
}
// The transmittance after blending alpha: T (1 - alpha) (raseter.wgsl:93) as
// one FMA, T - alpha T, the kernels' order.
inline float transmit_f32(float T, float alpha) { return fmaf(-alpha, T, T); }

template <typename T> struct M;
template <> struct M<float> {
    static float ex(float x) { return exp_f32(x); }
    static float ex_blend(float x) { return exp_blend_f32(x); }
    static bool in_blend_range(float p) { return blend_range_f32(p); }
    static float fma(float a, float b, float c) { return fmaf(a, b, c); }
    static float power(float A, float B, float C, float dx, float dy) { return power_f32(A, B, C, dx, dy); }
    static float transmit(float T, float alpha) { return transmit_f32(T, alpha); }
};
template <> struct M<double> {
    static double ex(double x) { return std::exp(x); }
    static double ex_blend(double x) { return std::exp(x); }
    static bool in_blend_range(double p) { return !(p > 0.0); }
    static double fma(double a, double b, double c) { return std::fma(a, b, c); }
    // the reference's formulas as written
    static double power(double A, double B, double C, double dx, double dy) {
        return -0.5 * (A * dx * dx + C * dy * dy) - B * dx * dy;
    }
    static double transmit(double T, double alpha) { return T * (1.0 - alpha); }
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
    uint8_t flags;   // bit 0-2: colour channel clamped at 0; bit 3/4: x/y frustum-guard clamp active
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
    if (tx / tz < -limx || tx / tz > limx) o.flags |= 8;
    if (ty / tz < -limy || ty / tz > limy) o.flags |= 16;
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
        if (acc < T(0)) o.flags |= (uint8_t)(1u << ch);
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
                    uint32_t* tiles, uint32_t* meta, T* conic, T* xy, T* rgb, uint8_t* flags = nullptr) {
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
        if (flags) flags[g] = o.flags;
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
inline uint64_t fnv(uint64_t h, uint64_t v) {
    for (int k = 0; k < 8; ++k) {
        h ^= (v >> (8 * k)) & 0xFF;
        h *= 1099511628211ull;
    }
    return h;
}

// sig (optional, per pixel): hash of the pixel's discrete decisions (which
// splats it blends and where it stops), to tell when two frames lie on the
// same smooth branch of the renderer.
// evals (optional): pixel-splat evaluations performed under the
// early-termination rules (the raster's algorithmic VALU work).
// One pixel (integer coordinates fxp, fyp) blended over `len` splats in
// order; splat(k) gives the k-th splat's conic (A, B, C), opacity, pixel
// position and colour, id(k) its index.  Results in (c0, c1, c2), Tr,
// lastc (1-based position of the last blended splat); h hashes the pixel's
// discrete decisions, ev counts evaluations.
template <typename T> struct SplatV {
    T A, B, C, op, x, y, r, g, b;
};
template <typename T, typename Splat, typename Id>
inline void blend_pixel(T fxp, T fyp, uint32_t len, Splat splat, Id id, T& c0, T& c1, T& c2, T& Tr,
                        uint32_t& lastc, uint64_t& h, uint64_t& ev) {
    Tr = 1;
    c0 = c1 = c2 = 0;
    lastc = 0;
    for (uint32_t i = 0; i < len; ++i) {
        ++ev;
        const SplatV<T> s = splat(i);
        const T dx = s.x - fxp, dy = s.y - fyp;
        // power = -0.5 (A dx^2 + C dy^2) - B dx dy  (calculatePower :113-117)
        const T power = M<T>::power(s.A, s.B, s.C, dx, dy);
        if (!M<T>::in_blend_range(power)) continue;
        const T alpha = std::min(T(0.99f), s.op * M<T>::ex_blend(power));
        if (alpha < T(1.0f / 255.0f)) continue;
        const T testT = M<T>::transmit(Tr, alpha);
        if (testT < T(0.0001f)) {
            h = fnv(h, 0xFFFFFFFFull + i);
            break;
        }
        h = fnv(h, ((uint64_t)id(i) << 1) | (s.op * M<T>::ex_blend(power) > T(0.99f) ? 1u : 0u));
        const T wgt = alpha * Tr;
        c0 = M<T>::fma(s.r, wgt, c0);
        c1 = M<T>::fma(s.g, wgt, c1);
        c2 = M<T>::fma(s.b, wgt, c2);
        Tr = testT;
        lastc = i + 1;
    }
}

// sig (optional, per pixel): hash of the pixel's discrete decisions (which
// splats it blends and where it stops), to tell when two frames lie on the
// same smooth branch of the renderer.
// evals (optional): pixel-splat evaluations performed under the
// early-termination rules (the raster's algorithmic VALU work).
template <typename T>
void raster_all(uint32_t W, uint32_t H, uint32_t tile, const uint32_t* ranges, const uint32_t* vals,
                const T* conic, const T* xy, const T* rgb, T* out, T* tfinal, uint32_t* last,
                uint64_t* sig = nullptr, uint64_t* evals = nullptr) {
    const uint32_t TXn = (W + tile - 1) / tile, TYn = (H + tile - 1) / tile;
    const int64_t ntile = (int64_t)TXn * TYn;
    uint64_t ev = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : ev)
    for (int64_t t = 0; t < ntile; ++t) {
        const uint32_t tx = (uint32_t)(t % TXn), ty = (uint32_t)(t / TXn);
        const uint32_t start = ranges[t * 2], end = ranges[t * 2 + 1];
        auto splat = [&](uint32_t k) {
            const uint32_t g = vals[start + k];
            return SplatV<T>{conic[g * 4], conic[g * 4 + 1], conic[g * 4 + 2], conic[g * 4 + 3], xy[g * 2],
                             xy[g * 2 + 1], rgb[g * 4], rgb[g * 4 + 1], rgb[g * 4 + 2]};
        };
        auto id = [&](uint32_t k) { return vals[start + k]; };
        for (uint32_t ly = 0; ly < tile; ++ly)
            for (uint32_t lx = 0; lx < tile; ++lx) {
                const uint32_t x = tx * tile + lx, y = ty * tile + ly;
                if (x >= W || y >= H) continue;
                T Tr, c0, c1, c2;
                uint32_t lastc;
                uint64_t h = 1469598103934665603ull;
                blend_pixel<T>(T(x), T(y), end > start ? end - start : 0, splat, id, c0, c1, c2, Tr, lastc, h, ev);
                const size_t p = (size_t)y * W + x;
                out[p] = c0;
                out[(size_t)W * H + p] = c1;
                out[(size_t)2 * W * H + p] = c2;
                if (tfinal) tfinal[p] = Tr;
                if (last) last[p] = lastc;
                if (sig) sig[p] = h;
            }
    }
    if (evals) *evals = ev;
}

template <typename T>
int render(const float* packed, uint32_t n, const Cam& cam, uint32_t tile, uint64_t cap, T* out,
           T* tfinal, uint32_t* last, uint64_t* instances, double* stage_ms, int use_radix,
           uint64_t* signature = nullptr, uint64_t* evals = nullptr) {
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
    std::vector<uint64_t> psig(signature ? (size_t)W * H : 0);
    raster_all<T>(W, H, tile, ranges.data(), vals.data(), conic.data(), xy.data(), rgb.data(), out,
                  tfinal, last, signature ? psig.data() : nullptr, evals);
    auto t6 = clk::now();
    if (signature) {
        uint64_t h = fnv(1469598103934665603ull, total);
        for (uint64_t i = 0; i < m; ++i) h = fnv(h, keys[i] ^ ((uint64_t)vals[i] << 40));
        for (uint64_t v : psig) h = fnv(h, v);
        *signature = h;
    }
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


// ---------------------------------------------------------------------------
// FP64 adjoint of the frame (the reference has no backward; this is the
// analytic derivative of the forward above).  Gradient layout per splat
// (grad2d, 12 doubles): d/d(x, y) of the pixel position, d/d(A, B, C) of the
// conic, d/d opacity, d/d(r, g, b), 3 unused.  Parameter gradients are
// planar [59][N] like the device layout.

constexpr int G2 = 12;

// Per-splat preprocess adjoint (chain of preprocess.wgsl:80-135 in FP64).
// Writes d/d(parameter ch) to out[ch * stride + index].
// flags (may be null): the FP32 forward's clamp decisions for this splat
// (see Pre::flags); without them the FP64 values decide.
void preprocess_backward_one(const Packed& P, uint32_t g, const Cam& cam, const double* G, double* out,
                             size_t stride, size_t index, const uint8_t* flags = nullptr) {
    auto put = [&](int ch, double v) { out[(size_t)ch * stride + index] = v; };
    const double px = P.pos(g, 0), py = P.pos(g, 1), pz = P.pos(g, 2);
    const float* V = cam.view;
    const float* PV = cam.view_proj;
    const double tx = V[0] * px + V[4] * py + V[8] * pz + V[12];
    const double ty = V[1] * px + V[5] * py + V[9] * pz + V[13];
    const double tz = V[2] * px + V[6] * py + V[10] * pz + V[14];
    double gp[3] = {0, 0, 0};
    bool any = false;
    for (int k = 0; k < 9; ++k) any |= G[k] != 0.0;
    if (!any || tz >= -0.2) {
        for (int ch = 0; ch < 59; ++ch) put(ch, 0.0);
        return;
    }
    // pixel position <- clip <- p
    const double cx = PV[0] * px + PV[4] * py + PV[8] * pz + PV[12];
    const double cy = PV[1] * px + PV[5] * py + PV[9] * pz + PV[13];
    const double cw = PV[3] * px + PV[7] * py + PV[11] * pz + PV[15];
    const double w = cw + double(1e-7f);
    const double W = cam.texture_size[0], H = cam.texture_size[1];
    const double gnx = G[0] * W * 0.5, gny = G[1] * H * 0.5;
    const double gcx = gnx / w, gcy = gny / w, gcw = -(gnx * cx + gny * cy) / (w * w);
    for (int c = 0; c < 3; ++c) gp[c] += PV[c * 4 + 0] * gcx + PV[c * 4 + 1] * gcy + PV[c * 4 + 3] * gcw;
    // forward intermediates of the covariance
    const double s[3] = {std::exp((double)P.lscale(g, 0)), std::exp((double)P.lscale(g, 1)),
                         std::exp((double)P.lscale(g, 2))};
    const double rq[4] = {P.rot(g, 0), P.rot(g, 1), P.rot(g, 2), P.rot(g, 3)};  // w x y z
    const double ql = std::sqrt(rq[0] * rq[0] + rq[1] * rq[1] + rq[2] * rq[2] + rq[3] * rq[3]);
    const double qw = rq[0] / ql, qx = rq[1] / ql, qy = rq[2] / ql, qz = rq[3] / ql;
    double R[3][3];
    R[0][0] = 1 - 2 * (qy * qy + qz * qz); R[0][1] = 2 * (qx * qy - qw * qz); R[0][2] = 2 * (qx * qz + qw * qy);
    R[1][0] = 2 * (qx * qy + qw * qz); R[1][1] = 1 - 2 * (qx * qx + qz * qz); R[1][2] = 2 * (qy * qz - qw * qx);
    R[2][0] = 2 * (qx * qz - qw * qy); R[2][1] = 2 * (qy * qz + qw * qx); R[2][2] = 1 - 2 * (qx * qx + qy * qy);
    double Mm[3][3], S[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Mm[r][c] = R[r][c] * s[c];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[i][j] = Mm[i][0] * Mm[j][0] + Mm[i][1] * Mm[j][1] + Mm[i][2] * Mm[j][2];
    const double limx = double(1.3f) * cam.tan_fov[0], limy = double(1.3f) * cam.tan_fov[1];
    const double vx = tx / tz, vy = ty / tz;
    const bool clx = flags ? (*flags & 8) != 0 : (vx < -limx || vx > limx);
    const bool cly = flags ? (*flags & 16) != 0 : (vy < -limy || vy > limy);
    const double ux = (clx ? (vx > 0 ? limx : -limx) : vx) * tz;
    const double uy = (cly ? (vy > 0 ? limy : -limy) : vy) * tz;
    const double uz = tz;
    const double fx = cam.focal[0], fy = cam.focal[1];
    const double J00 = fx / uz, J02 = -fx * ux / (uz * uz), J11 = fy / uz, J12 = -fy * uy / (uz * uz);
    double Tm[2][3];
    for (int c = 0; c < 3; ++c) {
        Tm[0][c] = J00 * V[c * 4 + 0] + J02 * V[c * 4 + 2];
        Tm[1][c] = J11 * V[c * 4 + 1] + J12 * V[c * 4 + 2];
    }
    double St0[3], St1[3];
    for (int i = 0; i < 3; ++i) {
        St0[i] = S[i][0] * Tm[0][0] + S[i][1] * Tm[0][1] + S[i][2] * Tm[0][2];
        St1[i] = S[i][0] * Tm[1][0] + S[i][1] * Tm[1][1] + S[i][2] * Tm[1][2];
    }
    const double a = Tm[0][0] * St0[0] + Tm[0][1] * St0[1] + Tm[0][2] * St0[2] + double(0.3f);
    const double b = Tm[0][0] * St1[0] + Tm[0][1] * St1[1] + Tm[0][2] * St1[2];
    const double c = Tm[1][0] * St1[0] + Tm[1][1] * St1[1] + Tm[1][2] * St1[2] + double(0.3f);
    const double det = a * c - b * b, det2 = det * det;
    // conic (A, B, C) = (c, -b, a) / det
    const double gA = G[2], gB = G[3], gC = G[4];
    const double ga = gA * (-c * c / det2) + gB * (b * c / det2) + gC * (-b * b / det2);
    const double gb = gA * (2 * b * c / det2) + gB * (-(det + 2 * b * b) / det2) + gC * (2 * a * b / det2);
    const double gc = gA * (-b * b / det2) + gB * (a * b / det2) + gC * (-a * a / det2);
    // (a, b, c) = (t0' S t0, t0' S t1, t1' S t1) with t0, t1 the rows of T = J W
    double gt0[3], gt1[3], gS[3][3];
    for (int i = 0; i < 3; ++i) {
        gt0[i] = 2 * ga * St0[i] + gb * St1[i];
        gt1[i] = gb * St0[i] + 2 * gc * St1[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            gS[i][j] = ga * Tm[0][i] * Tm[0][j] + gb * Tm[0][i] * Tm[1][j] + gc * Tm[1][i] * Tm[1][j];
    double gJ00 = 0, gJ02 = 0, gJ11 = 0, gJ12 = 0;
    for (int cc = 0; cc < 3; ++cc) {
        gJ00 += gt0[cc] * V[cc * 4 + 0];
        gJ02 += gt0[cc] * V[cc * 4 + 2];
        gJ11 += gt1[cc] * V[cc * 4 + 1];
        gJ12 += gt1[cc] * V[cc * 4 + 2];
    }
    const double uz2 = uz * uz, uz3 = uz2 * uz;
    const double gux = gJ02 * (-fx / uz2), guy = gJ12 * (-fy / uz2);
    const double guz = gJ00 * (-fx / uz2) + gJ02 * (2 * fx * ux / uz3) + gJ11 * (-fy / uz2) + gJ12 * (2 * fy * uy / uz3);
    double gt[3] = {clx ? 0.0 : gux, cly ? 0.0 : guy, guz};
    if (clx) gt[2] += gux * (vx > 0 ? limx : -limx);
    if (cly) gt[2] += guy * (vy > 0 ? limy : -limy);
    for (int cc = 0; cc < 3; ++cc) gp[cc] += V[cc * 4 + 0] * gt[0] + V[cc * 4 + 1] * gt[1] + V[cc * 4 + 2] * gt[2];
    // Sigma = M M'  ->  dM = (gS + gS') M
    double gM[3][3];
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += (gS[r][k] + gS[k][r]) * Mm[k][cc];
            gM[r][cc] = acc;
        }
    double gR[3][3], gls[3];
    for (int cc = 0; cc < 3; ++cc) {
        double gsc = 0;
        for (int r = 0; r < 3; ++r) {
            gR[r][cc] = gM[r][cc] * s[cc];
            gsc += gM[r][cc] * R[r][cc];
        }
        gls[cc] = gsc * s[cc];
    }
    // R(q) for the normalised quaternion, then through the normalisation
    const double gqw = 2 * (-qz * gR[0][1] + qy * gR[0][2] + qz * gR[1][0] - qx * gR[1][2] - qy * gR[2][0] + qx * gR[2][1]);
    const double gqx = 2 * (qy * gR[0][1] + qz * gR[0][2] + qy * gR[1][0] - 2 * qx * gR[1][1] - qw * gR[1][2] +
                            qz * gR[2][0] + qw * gR[2][1] - 2 * qx * gR[2][2]);
    const double gqy = 2 * (-2 * qy * gR[0][0] + qx * gR[0][1] + qw * gR[0][2] + qx * gR[1][0] + qz * gR[1][2] -
                            qw * gR[2][0] + qz * gR[2][1] - 2 * qy * gR[2][2]);
    const double gqz = 2 * (-2 * qz * gR[0][0] - qw * gR[0][1] + qx * gR[0][2] + qw * gR[1][0] - 2 * qz * gR[1][1] +
                            qy * gR[1][2] + qx * gR[2][0] + qy * gR[2][1]);
    const double qh[4] = {qw, qx, qy, qz}, gqh[4] = {gqw, gqx, gqy, gqz};
    const double dotq = qh[0] * gqh[0] + qh[1] * gqh[1] + qh[2] * gqh[2] + qh[3] * gqh[3];
    for (int k = 0; k < 4; ++k) put(6 + k, (gqh[k] - qh[k] * dotq) / ql);
    for (int k = 0; k < 3; ++k) put(3 + k, gls[k]);
    // opacity = sigmoid(logit)
    const double op = 1.0 / (1.0 + std::exp(-(double)P.opacity(g)));
    put(13, G[5] * op * (1 - op));
    // colour: SH degree 3 at d = normalize(p - eye), clamped at 0
    double v3[3] = {px - cam.eye[0], py - cam.eye[1], pz - cam.eye[2]};
    const double vl = std::sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
    const double X = v3[0] / vl, Y = v3[1] / vl, Z = v3[2] / vl;
    const double XX = X * X, YY = Y * Y, ZZ = Z * Z;
    const double C1 = float(SH_C1);
    const double C2[5] = {float(SH_C2[0]), float(SH_C2[1]), float(SH_C2[2]), float(SH_C2[3]), float(SH_C2[4])};
    const double C3[7] = {float(SH_C3[0]), float(SH_C3[1]), float(SH_C3[2]), float(SH_C3[3]),
                          float(SH_C3[4]), float(SH_C3[5]), float(SH_C3[6])};
    const double basis[15] = {-C1 * Y, C1 * Z, -C1 * X, C2[0] * X * Y, C2[1] * Y * Z,
                              C2[2] * (2 * ZZ - XX - YY), C2[3] * X * Z, C2[4] * (XX - YY),
                              C3[0] * Y * (3 * XX - YY), C3[1] * X * Y * Z, C3[2] * Y * (4 * ZZ - XX - YY),
                              C3[3] * Z * (2 * ZZ - 3 * XX - 3 * YY), C3[4] * X * (4 * ZZ - XX - YY),
                              C3[5] * Z * (XX - YY), C3[6] * X * (XX - 3 * YY)};
    const double dbx[15] = {0, 0, -C1, C2[0] * Y, 0, C2[2] * (-2 * X), C2[3] * Z, C2[4] * 2 * X,
                            C3[0] * Y * 6 * X, C3[1] * Y * Z, C3[2] * Y * (-2 * X), C3[3] * Z * (-6 * X),
                            C3[4] * (4 * ZZ - 3 * XX - YY), C3[5] * Z * 2 * X, C3[6] * (3 * XX - 3 * YY)};
    const double dby[15] = {-C1, 0, 0, C2[0] * X, C2[1] * Z, C2[2] * (-2 * Y), 0, C2[4] * (-2 * Y),
                            C3[0] * (3 * XX - 3 * YY), C3[1] * X * Z, C3[2] * (4 * ZZ - XX - 3 * YY),
                            C3[3] * Z * (-6 * Y), C3[4] * X * (-2 * Y), C3[5] * Z * (-2 * Y), C3[6] * X * (-6 * Y)};
    const double dbz[15] = {0, C1, 0, 0, C2[1] * Y, C2[2] * 4 * Z, C2[3] * X, 0, 0, C3[1] * X * Y,
                            C3[2] * Y * 8 * Z, C3[3] * (6 * ZZ - 3 * XX - 3 * YY), C3[4] * X * 8 * Z,
                            C3[5] * (XX - YY), 0};
    const double C0 = float(SH_C0);
    double gd[3] = {0, 0, 0};
    for (int ch = 0; ch < 3; ++ch) {
        double acc = C0 * P.dc(g, ch);
        for (int k = 0; k < 15; ++k) acc += basis[k] * P.rest(g, k, ch);
        const bool clamped = flags ? ((*flags >> ch) & 1) != 0 : (acc + 0.5 < 0.0);
        const double gch = clamped ? 0.0 : G[6 + ch];
        put(10 + ch, C0 * gch);
        for (int k = 0; k < 15; ++k) {
            const double f = P.rest(g, k, ch);
            put(14 + 15 * ch + k, basis[k] * gch);
            gd[0] += gch * dbx[k] * f;
            gd[1] += gch * dby[k] * f;
            gd[2] += gch * dbz[k] * f;
        }
    }
    const double dd = X * gd[0] + Y * gd[1] + Z * gd[2];
    gp[0] += (gd[0] - X * dd) / vl;
    gp[1] += (gd[1] - Y * dd) / vl;
    gp[2] += (gd[2] - Z * dd) / vl;
    for (int k = 0; k < 3; ++k) put(k, gp[k]);
}

// Whole-frame FP64 forward + adjoint.  decisions_f32 = 1 takes every
// discrete decision (tile rectangles, sort order, which splats a pixel
// blends, where it stops) from the FP32 pipeline and evaluates the values in
// FP64: the adjoint of the FP32 forward's branch.  decisions_f32 = 0 is the
// pure FP64 frame (what finite differences of orc_render_f64 see).
void backward_f64(const float* packed, uint32_t n, const Cam& cam, uint32_t tile, const double* dLdC,
                  int decisions_f32, double* grad, double* grad2d, double* image, double* grad2d_mass) {
    const uint32_t W = (uint32_t)cam.texture_size[0], H = (uint32_t)cam.texture_size[1];
    const uint32_t TXn = (W + tile - 1) / tile, TYn = (H + tile - 1) / tile;
    const size_t N = n;
    std::vector<uint32_t> tiles(N), meta(N * 4);
    std::vector<double> conic(N * 4), xy(N * 2), rgb(N * 4);
    std::vector<float> conicf(N * 4), xyf(N * 2), rgbf(N * 4);
    std::vector<uint8_t> flags32(decisions_f32 ? N : 0);
    {
        std::vector<uint32_t> t64(N), m64(N * 4);
        preprocess_all<double>(packed, n, cam, tile, t64.data(), m64.data(), conic.data(), xy.data(), rgb.data());
        if (decisions_f32) {
            preprocess_all<float>(packed, n, cam, tile, tiles.data(), meta.data(), conicf.data(), xyf.data(),
                                  rgbf.data(), flags32.data());
        } else {
            tiles = t64;
            meta = m64;
        }
    }
    std::vector<uint32_t> counts(tiles), offsets(N + 1);
    counts.push_back(0);
    const uint64_t total = scan_exclusive(counts.data(), n + 1, offsets.data());
    std::vector<uint64_t> keys(total);
    std::vector<uint32_t> vals(total);
    emit_all(offsets.data(), meta.data(), n, W, tile, ~0ull, keys.data(), vals.data());
    uint32_t tbits = 0;
    while ((1ull << tbits) < (uint64_t)TXn * TYn) ++tbits;
    sort_radix_parallel(keys.data(), vals.data(), total, 16 + tbits);
    std::vector<uint32_t> ranges((size_t)TXn * TYn * 2);
    tile_ranges(keys.data(), total, TXn * TYn, ranges.data());
    std::fill(grad2d, grad2d + N * G2, 0.0);
    if (grad2d_mass) std::fill(grad2d_mass, grad2d_mass + N * G2, 0.0);
    if (image) std::fill(image, image + (size_t)3 * W * H, 0.0);
    const int64_t ntile = (int64_t)TXn * TYn;
#pragma omp parallel
    {
        std::vector<double> acc, accm;
        std::vector<uint32_t> blend;
        std::vector<uint8_t> clampv;
        std::vector<double> al, Ts;
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < ntile; ++t) {
            const uint32_t tx = (uint32_t)(t % TXn), ty = (uint32_t)(t / TXn);
            const uint32_t start = ranges[t * 2], end = ranges[t * 2 + 1];
            if (end <= start) continue;
            acc.assign((size_t)(end - start) * G2, 0.0);
            accm.assign((size_t)(end - start) * G2, 0.0);
            for (uint32_t ly = 0; ly < tile; ++ly)
                for (uint32_t lx = 0; lx < tile; ++lx) {
                    const uint32_t x = tx * tile + lx, y = ty * tile + ly;
                    if (x >= W || y >= H) continue;
                    // discrete decisions: which range positions this pixel blends
                    blend.clear();
                    clampv.clear();
                    if (decisions_f32) {
                        float Tr = 1.0f;
                        for (uint32_t i = start; i < end; ++i) {
                            const uint32_t g = vals[i];
                            const float A = conicf[g * 4], B = conicf[g * 4 + 1], C = conicf[g * 4 + 2];
                            const float dx = xyf[g * 2] - (float)x, dy = xyf[g * 2 + 1] - (float)y;
                            const float power = power_f32(A, B, C, dx, dy);
                            if (!blend_range_f32(power)) continue;
                            const float oe = conicf[g * 4 + 3] * exp_blend_f32(power);
                            const float alpha = std::min(0.99f, oe);
                            if (alpha < 1.0f / 255.0f) continue;
                            const float tn = transmit_f32(Tr, alpha);
                            if (tn < 0.0001f) break;
                            Tr = tn;
                            blend.push_back(i);
                            clampv.push_back(oe > 0.99f ? 1 : 0);
                        }
                    } else {
                        double Tr = 1.0;
                        for (uint32_t i = start; i < end; ++i) {
                            const uint32_t g = vals[i];
                            const double dx = xy[g * 2] - x, dy = xy[g * 2 + 1] - y;
                            const double power = -0.5 * (conic[g * 4] * dx * dx + conic[g * 4 + 2] * dy * dy) -
                                                 conic[g * 4 + 1] * dx * dy;
                            if (power > 0.0) continue;
                            const double oe = conic[g * 4 + 3] * std::exp(power);
                            const double alpha = std::min(double(0.99f), oe);
                            if (alpha < double(1.0f / 255.0f)) continue;
                            const double tn = Tr * (1.0 - alpha);
                            if (tn < double(0.0001f)) break;
                            Tr = tn;
                            blend.push_back(i);
                            clampv.push_back(oe > double(0.99f) ? 1 : 0);
                        }
                    }
                    // FP64 forward over the blended list, then its exact adjoint
                    const size_t L = blend.size();
                    al.resize(L);
                    Ts.resize(L + 1);
                    Ts[0] = 1.0;
                    double col[3] = {0, 0, 0};
                    for (size_t k = 0; k < L; ++k) {
                        const uint32_t g = vals[blend[k]];
                        const double dx = xy[g * 2] - x, dy = xy[g * 2 + 1] - y;
                        const double power = -0.5 * (conic[g * 4] * dx * dx + conic[g * 4 + 2] * dy * dy) -
                                             conic[g * 4 + 1] * dx * dy;
                        al[k] = clampv[k] ? double(0.99f) : conic[g * 4 + 3] * std::exp(power);
                        for (int ch = 0; ch < 3; ++ch) col[ch] += rgb[g * 4 + ch] * al[k] * Ts[k];
                        Ts[k + 1] = Ts[k] * (1.0 - al[k]);
                    }
                    const size_t pix = (size_t)y * W + x;
                    if (image)
                        for (int ch = 0; ch < 3; ++ch) image[(size_t)ch * W * H + pix] = col[ch];
                    const double gC[3] = {dLdC[pix], dLdC[(size_t)W * H + pix], dLdC[(size_t)2 * W * H + pix]};
                    double Bk[3] = {0, 0, 0};  // colour behind splat k, normalised by T_{k+1}
                    for (size_t kk = L; kk-- > 0;) {
                        const uint32_t i = blend[kk];
                        const uint32_t g = vals[i];
                        double* a = &acc[(size_t)(i - start) * G2];
                        double* am = &accm[(size_t)(i - start) * G2];
                        const double alpha = al[kk], T = Ts[kk];
                        // mass: the summed magnitudes of every product inside each
                        // term (what FP32 rounding of the term is relative to)
                        double gal = 0, gal_m = 0;
                        for (int ch = 0; ch < 3; ++ch) {
                            a[6 + ch] += alpha * T * gC[ch];
                            am[6 + ch] += std::fabs(alpha * T * gC[ch]);
                            gal += (rgb[g * 4 + ch] - Bk[ch]) * gC[ch];
                            gal_m += (std::fabs(rgb[g * 4 + ch]) + std::fabs(Bk[ch])) * std::fabs(gC[ch]);
                        }
                        gal *= T;
                        gal_m *= T;
                        const double A = conic[g * 4], B = conic[g * 4 + 1], C = conic[g * 4 + 2];
                        const double dx = xy[g * 2] - x, dy = xy[g * 2 + 1] - y;
                        const double e = std::exp(-0.5 * (A * dx * dx + C * dy * dy) - B * dx * dy);
                        if (!clampv[kk]) {
                            const double gpow = alpha * gal, gpow_m = alpha * gal_m;
                            const double t[6] = {gpow * (-(A * dx + B * dy)), gpow * (-(C * dy + B * dx)),
                                                 gpow * (-0.5 * dx * dx), gpow * (-dx * dy), gpow * (-0.5 * dy * dy),
                                                 e * gal};
                            const double tm[6] = {gpow_m * (std::fabs(A * dx) + std::fabs(B * dy)),
                                                  gpow_m * (std::fabs(C * dy) + std::fabs(B * dx)),
                                                  gpow_m * 0.5 * dx * dx, gpow_m * std::fabs(dx * dy),
                                                  gpow_m * 0.5 * dy * dy, e * gal_m};
                            for (int q = 0; q < 6; ++q) {
                                a[q] += t[q];
                                am[q] += tm[q];
                            }
                        }
                        for (int ch = 0; ch < 3; ++ch) Bk[ch] = alpha * rgb[g * 4 + ch] + (1 - alpha) * Bk[ch];
                    }
                }
            for (uint32_t i = start; i < end; ++i) {
                const double* a = &acc[(size_t)(i - start) * G2];
                const double* am = &accm[(size_t)(i - start) * G2];
                double* dst = &grad2d[(size_t)vals[i] * G2];
                double* dm = grad2d_mass ? &grad2d_mass[(size_t)vals[i] * G2] : nullptr;
                for (int k = 0; k < 9; ++k)
                    if (a[k] != 0.0 || am[k] != 0.0) {
#pragma omp atomic
                        dst[k] += a[k];
                        if (dm) {
#pragma omp atomic
                            dm[k] += am[k];
                        }
                    }
            }
        }
    }
    Packed P{packed, n};
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < (int64_t)n; ++g)
        preprocess_backward_one(P, (uint32_t)g, cam, &grad2d[g * G2], grad, N, (size_t)g,
                                decisions_f32 ? &flags32[g] : nullptr);
}

// Central differences of L = sum(dLdC * C) of the FP64 frame for every
// parameter of the given splats, each perturbed frame rendered only over the
// pixels of the tiles the splat's rectangle covers (with the base parameters
// or either perturbation) where either perturbed splat can blend (power <= 0
// and alpha >= 1/255, tested with a widened threshold): at every other pixel
// both frames blend the same splats in the same order, so the difference of
// the rendered pixels' L is the whole frame's.  The window's instance lists are the base
// frame's (emit order, stable by depth16) with the perturbed splat removed
// and re-inserted at its key.  fd[k*59+ch]; same[k*59+ch] = 1 when both
// perturbed windows take the same discrete decisions (blends, stops, clamps,
// the splat's tile rectangle and depth key).
void fd_window_f64(const float* packed, uint32_t n, const Cam& cam, uint32_t tile, const double* dLdC,
                   const uint32_t* splats, uint32_t count, float eps, double* fd, uint8_t* same) {
    const uint32_t W = (uint32_t)cam.texture_size[0], H = (uint32_t)cam.texture_size[1];
    const uint32_t TXn = (W + tile - 1) / tile, TYn = (H + tile - 1) / tile;
    const size_t N = n, ntile = (size_t)TXn * TYn;
    std::vector<uint32_t> tiles(N), meta(N * 4);
    std::vector<double> conic(N * 4), xy(N * 2), rgb(N * 4);
    preprocess_all<double>(packed, n, cam, tile, tiles.data(), meta.data(), conic.data(), xy.data(), rgb.data());
    // base per-tile lists: splats in index order, then stable by depth16
    std::vector<uint64_t> start(ntile + 1, 0);
    auto each_tile = [&](const uint32_t* m, auto fn) {
        const uint32_t x0 = m[0] & 0xFFFF, y0 = m[0] >> 16, x1 = m[1] & 0xFFFF, y1 = m[1] >> 16;
        for (uint32_t ty = y0; ty <= y1; ++ty)
            for (uint32_t tx = x0; tx <= x1; ++tx) fn((size_t)ty * TXn + tx);
    };
    for (size_t g = 0; g < N; ++g)
        if (tiles[g]) each_tile(&meta[g * 4], [&](size_t t) { ++start[t + 1]; });
    for (size_t t = 0; t < ntile; ++t) start[t + 1] += start[t];
    std::vector<uint32_t> list(start[ntile]);
    {
        std::vector<uint64_t> cur(start.begin(), start.end() - 1);
        for (size_t g = 0; g < N; ++g)
            if (tiles[g]) each_tile(&meta[g * 4], [&](size_t t) { list[cur[t]++] = (uint32_t)g; });
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < (int64_t)ntile; ++t)
        std::stable_sort(list.begin() + start[t], list.begin() + start[t + 1],
                         [&](uint32_t a, uint32_t b) { return meta[a * 4 + 2] < meta[b * 4 + 2]; });
    // packed block offset of planar channel ch of splat g in an array of nn splats
    auto offset = [](uint32_t nn, int ch, uint32_t g) -> size_t {
        const size_t M = nn;
        if (ch < 3) return (size_t)g * 3 + ch;
        if (ch < 6) return M * 3 + (size_t)g * 3 + (ch - 3);
        if (ch < 10) return M * 6 + (size_t)g * 4 + (ch - 6);
        if (ch < 13) return M * 10 + (size_t)g * 3 + (ch - 10);
        if (ch < 14) return M * 13 + g;
        return M * 14 + (size_t)g * 45 + (ch - 14);
    };
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t job = 0; job < (int64_t)count * 59; ++job) {
        const uint32_t g = splats[job / 59];
        const int ch = (int)(job % 59);
        float one[59];
        for (int c = 0; c < 59; ++c) one[offset(1, c, 0)] = packed[offset(n, c, g)];
        const float v = one[offset(1, ch, 0)];
        const float vs[2] = {v + eps, v - eps};
        Pre<double> side[2];
        for (int k = 0; k < 2; ++k) {
            one[offset(1, ch, 0)] = vs[k];
            preprocess_one<double>(Packed{one, 1}, 0, cam, tile, side[k]);
        }
        // window: union of the base and perturbed tile rectangles
        uint32_t wx0 = ~0u, wy0 = ~0u, wx1 = 0, wy1 = 0;
        auto grow = [&](const uint32_t* m) {
            wx0 = std::min(wx0, m[0] & 0xFFFF); wy0 = std::min(wy0, m[0] >> 16);
            wx1 = std::max(wx1, m[1] & 0xFFFF); wy1 = std::max(wy1, m[1] >> 16);
        };
        if (tiles[g]) grow(&meta[(size_t)g * 4]);
        for (int k = 0; k < 2; ++k)
            if (side[k].tiles) grow(side[k].meta);
        if (wx0 == ~0u) {  // culled in all three frames
            fd[job] = 0.0;
            same[job] = 1;
            continue;
        }
        // pixels where either perturbed splat can blend: 0 <= q <= 2 ln(255 op), widened
        auto reach = [&](const Pre<double>& P, double x, double y) {
            if (!P.tiles) return false;
            const double op = P.conic[3];
            if (!(op * 255.0 > 1.0)) return false;
            const double dx = P.xy[0] - x, dy = P.xy[1] - y;
            const double q = P.conic[0] * dx * dx + 2.0 * P.conic[1] * dx * dy + P.conic[2] * dy * dy;
            return !(q > 1.02 * 2.0 * std::log(255.0 * op) + 0.1);
        };
        double L[2];
        uint64_t sig[2];
        std::vector<uint32_t> loc;
        for (int k = 0; k < 2; ++k) {
            const Pre<double>& P = side[k];
            const uint32_t px0 = P.meta[0] & 0xFFFF, py0 = P.meta[0] >> 16, px1 = P.meta[1] & 0xFFFF,
                           py1 = P.meta[1] >> 16;
            uint64_t h = 1469598103934665603ull;
            h = fnv(h, P.tiles ? ((uint64_t)P.meta[0] << 32 | P.meta[1]) : 0);
            h = fnv(h, P.tiles ? P.meta[2] : 0xFFFFFFFFull);
            double acc = 0.0;
            uint64_t ev = 0;
            for (uint32_t ty = wy0; ty <= wy1; ++ty)
                for (uint32_t tx = wx0; tx <= wx1; ++tx) {
                    const size_t t = (size_t)ty * TXn + tx;
                    loc.clear();
                    const bool in = P.tiles && tx >= px0 && tx <= px1 && ty >= py0 && ty <= py1;
                    bool placed = !in;
                    for (uint64_t i = start[t]; i < start[t + 1]; ++i) {
                        const uint32_t o = list[i];
                        if (o == g) continue;
                        if (!placed && (meta[o * 4 + 2] > P.meta[2] || (meta[o * 4 + 2] == P.meta[2] && o > g))) {
                            loc.push_back(g);
                            placed = true;
                        }
                        loc.push_back(o);
                    }
                    if (!placed) loc.push_back(g);
                    auto splat = [&](uint32_t q) {
                        const uint32_t o = loc[q];
                        if (o == g)
                            return SplatV<double>{P.conic[0], P.conic[1], P.conic[2], P.conic[3], P.xy[0], P.xy[1],
                                                  P.rgb[0], P.rgb[1], P.rgb[2]};
                        return SplatV<double>{conic[o * 4], conic[o * 4 + 1], conic[o * 4 + 2], conic[o * 4 + 3],
                                              xy[o * 2], xy[o * 2 + 1], rgb[o * 4], rgb[o * 4 + 1], rgb[o * 4 + 2]};
                    };
                    auto id = [&](uint32_t q) { return loc[q]; };
                    for (uint32_t ly = 0; ly < tile; ++ly)
                        for (uint32_t lx = 0; lx < tile; ++lx) {
                            const uint32_t x = tx * tile + lx, y = ty * tile + ly;
                            if (x >= W || y >= H) continue;
                            if (!reach(side[0], x, y) && !reach(side[1], x, y)) continue;
                            double Tr, c0, c1, c2;
                            uint32_t lastc;
                            blend_pixel<double>(double(x), double(y), (uint32_t)loc.size(), splat, id, c0, c1, c2, Tr,
                                                lastc, h, ev);
                            const size_t pix = (size_t)y * W + x;
                            acc += c0 * dLdC[pix] + c1 * dLdC[(size_t)W * H + pix] + c2 * dLdC[(size_t)2 * W * H + pix];
                        }
                }
            L[k] = acc;
            sig[k] = h;
        }
        fd[job] = (L[0] - L[1]) / ((double)vs[0] - (double)vs[1]);
        same[job] = sig[0] == sig[1] ? 1 : 0;
    }
}

// FP32 replay of the raster adjoint with the same steps as the GPU kernel
// (T recovered with 1 / (1 - alpha), the dL/dRGB-weighted light from behind
// GS = sum_ch g_ch sum_behind c_j alpha_j T_j, so that
// T_i (c_i - behind_i / T_{i+1}) . g = T_b (c . g) - GS / (1 - alpha),
// per-splat moments of dL/dpower turned into the 2-D gradient with the conic),
// per-splat sums in double.  The kernel takes a 1-ulp hardware reciprocal
// instead of the IEEE division and converts the moments in FP32; otherwise
// the arithmetic is the same, so the twin separates FP32 term error from
// summation order.
void raster_backward_f32(uint32_t W, uint32_t H, uint32_t tile, const uint32_t* ranges, const uint32_t* vals,
                         const float* conic, const float* xy, const float* rgb, const float* tfinal,
                         const uint32_t* last, const float* dLdC, uint32_t n, double* grad2d) {
    const uint32_t TXn = (W + tile - 1) / tile, TYn = (H + tile - 1) / tile;
    std::fill(grad2d, grad2d + (size_t)n * G2, 0.0);
    for (uint64_t t = 0; t < (uint64_t)TXn * TYn; ++t) {
        const uint32_t tx = (uint32_t)(t % TXn), ty = (uint32_t)(t / TXn);
        const uint32_t start = ranges[t * 2];
        for (uint32_t ly = 0; ly < tile; ++ly)
            for (uint32_t lx = 0; lx < tile; ++lx) {
                const uint32_t x = tx * tile + lx, y = ty * tile + ly;
                if (x >= W || y >= H) continue;
                const size_t o = (size_t)y * W + x;
                float T = tfinal[o];
                const uint32_t L = last[o];
                const float gr = dLdC[o], gg = dLdC[(size_t)W * H + o], gb = dLdC[(size_t)2 * W * H + o];
                float GS = 0;
                for (uint32_t pos = L; pos >= 1; --pos) {
                    const uint32_t g = vals[start + pos - 1];
                    const float A = conic[g * 4], B = conic[g * 4 + 1], C = conic[g * 4 + 2], op = conic[g * 4 + 3];
                    const float dx = xy[g * 2] - (float)x, dy = xy[g * 2 + 1] - (float)y;
                    const float power = power_f32(A, B, C, dx, dy);
                    if (!blend_range_f32(power)) continue;
                    const float e = exp_blend_f32(power);
                    const float oe = op * e;
                    const float alpha = std::min(0.99f, oe);
                    if (alpha < 1.0f / 255.0f) continue;
                    const float r = 1.0f / (1.0f - alpha);
                    const float Tb = T * r;
                    T = Tb;
                    const float cr = rgb[g * 4], cg = rgb[g * 4 + 1], cb = rgb[g * 4 + 2];
                    const float gc = fmaf(cr, gr, fmaf(cg, gg, cb * gb));
                    const float gal = fmaf(Tb, gc, -(r * GS));
                    const float w = alpha * Tb;
                    double* a = &grad2d[(size_t)g * G2];
                    a[6] += w * gr; a[7] += w * gg; a[8] += w * gb;
                    if (oe <= 0.99f) {
                        // moments of dL/dpower, turned into the 2-D gradient below
                        const float gpow = alpha * gal;
                        const float gdx = gpow * dx, gdy = gpow * dy;
                        a[0] += gdx;
                        a[1] += gdy;
                        a[2] += gdx * dx;
                        a[3] += gdx * dy;
                        a[4] += gdy * dy;
                        a[5] += e * gal;
                    }
                    GS = fmaf(w, gc, GS);
                }
            }
    }
    // per splat: d/dx = -(A Mx + B My), d/dy = -(C My + B Mx), d/dA = -Mxx/2,
    // d/dB = -Mxy, d/dC = -Myy/2 with the splat's conic
    for (uint32_t g = 0; g < n; ++g) {
        double* a = &grad2d[(size_t)g * G2];
        const double A = conic[g * 4], B = conic[g * 4 + 1], C = conic[g * 4 + 2];
        const double mx = a[0], my = a[1];
        a[0] = -(A * mx + B * my);
        a[1] = -(C * my + B * mx);
        a[2] *= -0.5;
        a[3] = -a[3];
        a[4] *= -0.5;
    }
}
}  // namespace

extern "C" {

float orc_exp_f32(float x) { return exp_f32(x); }

void orc_exp_blend_f32_array(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = exp_blend_f32(x[i]);
}

// The GPU kernels scale 2^f by 2^k by adding bits(t) << 23 to q's bits (the
// bits of 1.5 * 2^23 vanish in that shift) instead of ldexp.  Counts the
// floats p = -bits_as_float(b), b in [b_lo, b_hi], where that differs from
// exp_blend_f32's ldexp in any bit (tests: none on [-5.55, 0]; some where
// the result underflows, which shows the comparison sees a difference).
int64_t orc_exp_blend_shift_mismatches(uint32_t b_lo, uint32_t b_hi) {
    int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t b = b_lo; b <= (int64_t)b_hi; ++b) {
        const uint32_t pb = 0x80000000u | (uint32_t)b;
        float p, q;
        uint32_t tu;
        std::memcpy(&p, &pb, 4);
        exp_blend_parts(p, q, tu);
        const float a = ldexpf(q, (int32_t)(tu - 0x4B400000u));
        uint32_t qb, ab;
        std::memcpy(&qb, &q, 4);
        std::memcpy(&ab, &a, 4);
        bad += (ab != qb + (tu << 23)) ? 1 : 0;
    }
    return bad;
}

// The FP32 blend's alpha for (power, opacity) pairs, 0 where the pixel skips
// the splat (the raster's tests in its order, blend_range_f32 first).
void orc_blend_alpha_f32(const float* power, const float* op, float* alpha, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const float p = power[i];
        float a = 0.0f;
        if (blend_range_f32(p)) {
            a = std::min(0.99f, op[i] * exp_blend_f32(p));
            if (a < 1.0f / 255.0f) a = 0.0f;
        }
        alpha[i] = a;
    }
}

// The FP32 blend's power and transmittance update, elementwise (pinned
// against the FP64 formulas in tests/test_oracle.py).
void orc_power_f32_array(const float* A, const float* B, const float* C, const float* dx, const float* dy,
                         float* power, int64_t n) {
    for (int64_t i = 0; i < n; ++i) power[i] = power_f32(A[i], B[i], C[i], dx[i], dy[i]);
}

void orc_transmit_f32_array(const float* T, const float* alpha, float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = transmit_f32(T[i], alpha[i]);
}

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
                   int use_radix, uint64_t* evals) {
    return render<float>(packed, n, *(const Cam*)cam, tile, cap, out, tfinal, last, instances, stage_ms,
                         use_radix, nullptr, evals);
}

void orc_backward_f64(const float* packed, uint32_t n, const void* cam, uint32_t tile, const double* dLdC,
                      int decisions_f32, double* grad, double* grad2d, double* image, double* grad2d_mass) {
    backward_f64(packed, n, *(const Cam*)cam, tile, dLdC, decisions_f32, grad, grad2d, image, grad2d_mass);
}

// |J|^T m per splat: J the 59x9 Jacobian of the preprocess adjoint (columns
// from unit 2-D gradients), m the absolute mass of the 2-D gradient terms.
// Bounds the magnitude of the terms each parameter gradient sums.
void orc_param_mass_f64(const float* packed, uint32_t n, const void* cam, const double* grad2d_mass, double* out) {
    Packed P{packed, n};
    const size_t N = n;
#pragma omp parallel for schedule(static)
        for (int64_t g = 0; g < (int64_t)n; ++g) {
            double acc[59] = {0};
            for (int k = 0; k < 9; ++k) {
                const double mk = grad2d_mass[g * G2 + k];
                if (mk == 0.0) continue;
                double unit[G2] = {0};
                unit[k] = 1.0;
                double tmp[59];
                preprocess_backward_one(P, (uint32_t)g, *(const Cam*)cam, unit, tmp, 1, 0);
                for (int ch = 0; ch < 59; ++ch) acc[ch] += std::fabs(tmp[ch]) * mk;
            }
            for (int ch = 0; ch < 59; ++ch) out[(size_t)ch * N + g] = acc[ch];
        }
}

void orc_preprocess_backward_f64(const float* packed, uint32_t n, const void* cam, const double* grad2d,
                                 double* grad) {
    Packed P{packed, n};
    for (uint32_t g = 0; g < n; ++g)
        preprocess_backward_one(P, g, *(const Cam*)cam, &grad2d[(size_t)g * G2], grad, n, g);
}

// Central differences of L = sum(dLdC * C) of the FP64 frame for parameter
// `ch` (planar channel 0..58) of splats[k]: fd[k*59+ch], and same[k*59+ch] = 1
// when both perturbed frames take the same discrete decisions.
void orc_fd_f64(const float* packed, uint32_t n, const void* cam, uint32_t tile, const double* dLdC,
                const uint32_t* splats, uint32_t count, float eps, double* fd, uint8_t* same) {
    const Cam& C = *(const Cam*)cam;
    const uint32_t W = (uint32_t)C.texture_size[0], H = (uint32_t)C.texture_size[1];
    const size_t np = (size_t)n * 59;
    // planar channel -> packed offset for splat g
    auto offset = [n](int ch, uint32_t g) -> size_t {
        const size_t N = n;
        if (ch < 3) return (size_t)g * 3 + ch;
        if (ch < 6) return N * 3 + (size_t)g * 3 + (ch - 3);
        if (ch < 10) return N * 6 + (size_t)g * 4 + (ch - 6);
        if (ch < 13) return N * 10 + (size_t)g * 3 + (ch - 10);
        if (ch < 14) return N * 13 + g;
        return N * 14 + (size_t)g * 45 + (ch - 14);
    };
#pragma omp parallel
    {
        std::vector<float> pk(packed, packed + np);
        std::vector<double> img((size_t)3 * W * H), tf((size_t)W * H);
        std::vector<uint32_t> last((size_t)W * H);
#pragma omp for schedule(dynamic, 1)
        for (int64_t job = 0; job < (int64_t)count * 59; ++job) {
            const uint32_t g = splats[job / 59];
            const int ch = (int)(job % 59);
            const size_t o = offset(ch, g);
            const float v = packed[o];
            const float vp = v + eps, vm = v - eps;
            double L[2];
            uint64_t sig[2];
            for (int side = 0; side < 2; ++side) {
                pk[o] = side == 0 ? vp : vm;
                uint64_t inst = 0;
                render<double>(pk.data(), n, C, tile, 0, img.data(), tf.data(), last.data(), &inst, nullptr, 0,
                               &sig[side]);
                double acc = 0;
                for (size_t i = 0; i < img.size(); ++i) acc += img[i] * dLdC[i];
                L[side] = acc;
            }
            pk[o] = v;
            fd[job] = (L[0] - L[1]) / ((double)vp - (double)vm);
            same[job] = sig[0] == sig[1] ? 1 : 0;
        }
    }
}

void orc_fd_window_f64(const float* packed, uint32_t n, const void* cam, uint32_t tile, const double* dLdC,
                       const uint32_t* splats, uint32_t count, float eps, double* fd, uint8_t* same) {
    fd_window_f64(packed, n, *(const Cam*)cam, tile, dLdC, splats, count, eps, fd, same);
}

void orc_raster_backward_f32(uint32_t W, uint32_t H, uint32_t tile, const uint32_t* ranges, const uint32_t* vals,
                             const float* conic, const float* xy, const float* rgb, const float* tfinal,
                             const uint32_t* last, const float* dLdC, uint32_t n, double* grad2d) {
    raster_backward_f32(W, H, tile, ranges, vals, conic, xy, rgb, tfinal, last, dLdC, n, grad2d);
}

int orc_render_f64(const float* packed, uint32_t n, const void* cam, uint32_t tile, uint64_t cap,
                   double* out, double* tfinal, uint32_t* last, uint64_t* instances, uint64_t* signature) {
    return render<double>(packed, n, *(const Cam*)cam, tile, cap, out, tfinal, last, instances, nullptr,
                          0, signature);
}

}  // extern "C"
