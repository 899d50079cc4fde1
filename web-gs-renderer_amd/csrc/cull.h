// cull.h — conservative reach test of a splat against a pixel rectangle.
//
// A pixel receives a splat only if alpha = min(0.99, op * exp(power)) >= 1/255
// with power = -q/2, q = A dx^2 + 2 B dx dy + C dy^2 (raseter.wgsl:80-95,
// 113-117), i.e. only where q <= q_max = 2 ln(255 op).  The raster kernels
// skip a splat for a whole 8x8 pixel block when the minimum of q over the
// block's pixel rectangle exceeds q_max.  The threshold is widened
// (x1.02 + 0.1) so FP32 rounding of q, of the logarithm and of the minimum
// can never cull a splat some pixel would blend: skipping is exactly
// equivalent to evaluating.
// Any NaN in the test leaves the splat live.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gsr {

// q over the edge with dx fixed, dy in [lo, hi]: minimum at dy = -B dx / C
// (nbc = -B / C).  The stationary point comes from v_rcp_f32 (1 ulp): q at a
// point off the exact one by a relative 1e-7 exceeds the minimum by about
// C (1e-7 dy)^2, far inside the threshold's widening.
__device__ __forceinline__ float edge_min_dx(float A, float B, float C, float nbc, float dx, float lo, float hi) {
    const float dy = fminf(fmaxf(nbc * dx, lo), hi);
    return A * dx * dx + 2.0f * B * dx * dy + C * dy * dy;
}

__device__ __forceinline__ float edge_min_dy(float A, float B, float C, float nba, float dy, float lo, float hi) {
    const float dx = fminf(fmaxf(nba * dy, lo), hi);
    return A * dx * dx + 2.0f * B * dx * dy + C * dy * dy;
}

// Widened reach threshold on q; negative when the opacity alone keeps alpha
// below 1/255 everywhere.
__device__ __forceinline__ float reach_qmax(float op) {
    const float l = __logf(255.0f * op);   // v_log_f32 based; error absorbed by the widening
    return l > 0.0f ? 1.02f * (2.0f * l) + 0.1f : -1.0f;
}

// True unless the splat provably misses every pixel centre of [xa, xb] x [ya, yb].
// With A > 0 and C > 0, q is convex along every edge of the rectangle and
// has no interior minimum unless it is positive definite, whose minimum is
// the centre (tested first): the smallest edge minimum is the rectangle's.
// A conic with A <= 0 or C <= 0 (a covariance whose FP32 determinant came
// out <= 0: q is then indefinite and may still be in (0, q_max] off its
// axes) is never culled.
__device__ __forceinline__ bool reaches_rect(float gx, float gy, float A, float B, float C, float qmax, float xa,
                                             float xb, float ya, float yb) {
    if (!(qmax >= 0.0f)) return false;
    if (!(A > 0.0f) || !(C > 0.0f)) return true;
    if (gx >= xa && gx <= xb && gy >= ya && gy <= yb) return true;
    const float dxlo = gx - xb, dxhi = gx - xa, dylo = gy - yb, dyhi = gy - ya;
    const float nbc = -B * __builtin_amdgcn_rcpf(C), nba = -B * __builtin_amdgcn_rcpf(A);
    float q = edge_min_dx(A, B, C, nbc, dxhi, dylo, dyhi);
    q = fminf(q, edge_min_dx(A, B, C, nbc, dxlo, dylo, dyhi));
    q = fminf(q, edge_min_dy(A, B, C, nba, dyhi, dxlo, dxhi));
    q = fminf(q, edge_min_dy(A, B, C, nba, dylo, dxlo, dxhi));
    return !(q > qmax);
}

}  // namespace gsr
