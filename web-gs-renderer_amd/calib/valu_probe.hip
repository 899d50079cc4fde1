// valu_probe.hip — calibration kernels for reading rocprofv3's VALU counters.
//
// Not part of the renderer: bench.py --valu-probe launches them so that a
// counter run (SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, GRBM_GUI_ACTIVE) over
// kernels with a known VALU instruction count per wave tells how the counters
// scale and what VALU issue rates wave64 FP32 streams reach on this part.
// Built like the renderer (no packed FP32).  Each thread runs `iters`
// iterations of a loop body (the loop control on the scalar unit); the VALU
// count per wave of each mode is its body's count times iters plus the few
// instructions before and after the loop (read off the gfx950 assembly):
//   mode 0: 8 independent v_fma_f32 (a scalar and a vector source);
//   mode 1: 16 independent v_fma_f32, all sources in VGPRs;
//   mode 2: 16 independent v_add_f32;
//   mode 3: 8 v_fma_f32 and 8 selects, interleaved.
#include <hip/hip_runtime.h>

#include <cstdint>

template <int MODE>
__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int iters, float b, float c) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int K = MODE == 0 ? 8 : 16;
    float a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = (float)t + (float)k;
    float bv = b + (float)(t & 1) * 1e-30f, cv = c + (float)(t & 2) * 1e-30f;  // per-lane (VGPR) sources
#pragma unroll 1
    for (int i = 0; i < iters; ++i) {
        if (MODE == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = __builtin_fmaf(a[k], b, c);
        } else if (MODE == 1) {
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = __builtin_fmaf(a[k], bv, cv);
        } else if (MODE == 2) {
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = a[k] + bv;
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                a[k] = __builtin_fmaf(a[k], bv, cv);
                a[8 + k] = a[k] > a[8 + k] ? a[8 + k] : cv;
            }
        }
        asm volatile("" : "+v"(bv));  // keeps bv a per-lane register every iteration
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) s += a[k];
    out[t] = s;
}

// Launches `blocks` workgroups of 256 threads of the mode's kernel on the
// null stream (a warm-up, then one timed launch) and returns the timed
// launch's time in ms (HIP events); 0 on success.
extern "C" int gsr_calib_valu(int mode, int blocks, int iters, double* ms) {
    float* out = nullptr;
    if (mode < 0 || mode > 3) return 3;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    auto launch = [&]() {
        switch (mode) {
            case 0:
                hipLaunchKernelGGL(k_valu_probe<0>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0.001f);
                break;
            case 1:
                hipLaunchKernelGGL(k_valu_probe<1>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0.001f);
                break;
            case 2:
                hipLaunchKernelGGL(k_valu_probe<2>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0.001f);
                break;
            default:
                hipLaunchKernelGGL(k_valu_probe<3>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0.001f);
                break;
        }
    };
    launch();
    hipEventRecord(e0, 0);
    launch();
    hipEventRecord(e1, 0);
    const hipError_t e = hipEventSynchronize(e1);
    float t = 0.0f;
    hipEventElapsedTime(&t, e0, e1);
    *ms = t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(out);
    return e == hipSuccess ? 0 : 2;
}
