// context.hip — C ABI lifecycle, errors, scratch buffers, scenes, timings.
//
// Replaces the reference's device/context setup (src/gpu/device.js:3-40),
// scene residency (src/renderer/renderer.js:461-479), the per-pass timestamp
// profiler (src/gpu/profiler.js:52-123) and the error convention of
// src/utils/utils.js:3-47 (a status code plus a message instead of a throw).
#include <cstdarg>
#include <cstring>

#include "gsr_internal.h"

namespace gsr {

static thread_local std::string g_create_error = "no error";

int fail(Context* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->error = buf;
    else g_create_error = buf;
    return code;
}

void* buffer(Context* c, const char* name, size_t bytes) {
    if (bytes == 0) bytes = 256;
    DevBuf& b = c->bufs[name];
    if (b.bytes >= bytes) return b.ptr;
    if (b.ptr) {
        destroy_batch_graph(c);
        hipStreamSynchronize(c->stream);
        hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    const size_t rounded = (bytes + 255) & ~(size_t)255;
    if (hipMalloc(&b.ptr, rounded) != hipSuccess) {
        b.ptr = nullptr;
        fail(c, GSR_ERR_ALLOC, "device allocation of %zu bytes for '%s' failed", rounded, name);
        return nullptr;
    }
    b.bytes = rounded;
    return b.ptr;
}

void* scratch(Context* c, const char* name, size_t bytes) {
    return buffer(c, (std::string(c->scratch_prefix) + name).c_str(), bytes);
}

static hipEvent_t next_event(Context* c) {
    if (c->event_used == c->event_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->event_pool.push_back(e);
    }
    return c->event_pool[c->event_used++];
}

void mark(Context* c, const char* label) {
    if (!c->timing_on) return;
    // timing 2: only the raster stages and the frame / group ends, so that a
    // timed run carries two events per raster instead of one per stage
    if (c->timing_on == 2 && label[0] > 1 && !std::strstr(label, "raster")) return;
    hipEvent_t e = next_event(c);
    if (!e) return;
    hipEventRecord(e, c->stream);
    c->marks.emplace_back(label, e);
}

void mark_end(Context* c, bool frame_end) {
    if (!c->timing_on) return;
    mark(c, frame_end ? "" : "\x01");
}

namespace {

// Converts the recorded marks of the finished frames into rows averaged per
// frame (the stream is synced).  A frame's marks end with an empty label.
void collect_timings(Context* c) {
    if (c->marks.empty()) return;
    std::vector<gsr_timing_row> rows;
    int frames = 0;
    for (size_t i = 0; i < c->marks.size(); ++i) {
        const std::string& label = c->marks[i].first;
        if (label.empty()) { ++frames; continue; }   // end of a frame
        if (label[0] == '\x01') continue;            // end of a stage group inside a frame
        if (i + 1 >= c->marks.size()) break;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, c->marks[i].second, c->marks[i + 1].second);
        bool merged = false;
        for (auto& r : rows)
            if (label == r.label) { r.ms += ms; merged = true; }
        if (!merged) {
            gsr_timing_row r{};
            snprintf(r.label, sizeof(r.label), "%s", label.c_str());
            r.ms = ms;
            rows.push_back(r);
        }
    }
    c->marks.clear();
    c->event_used = 0;
    if (frames == 0) return;
    double total = 0.0;
    for (auto& r : rows) {
        r.ms /= frames;
        total += r.ms;
    }
    gsr_timing_row t{};
    snprintf(t.label, sizeof(t.label), "TOTAL");
    t.ms = total;
    rows.push_back(t);
    c->last_rows = rows;
    c->last_frames = frames;
}

__global__ void k_deinterleave(const float* __restrict__ src, uint32_t n, uint32_t k, float* __restrict__ dst) {
    const size_t total = (size_t)n * k;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const size_t g = i / k, j = i % k;
        dst[j * n + g] = src[i];
    }
}

__global__ void k_rgba8(const float* __restrict__ rgb, uint32_t W, uint32_t H, int flip, uint8_t* __restrict__ o) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const uint32_t x = i % W, y = i / W;
    const uint32_t sy = flip ? (H - 1 - y) : y;
    const size_t plane = (size_t)W * H;
    const size_t s = (size_t)sy * W + x;
    uint8_t px[4];
    for (int ch = 0; ch < 3; ++ch) {
        const float v = fminf(fmaxf(rgb[ch * plane + s], 0.0f), 1.0f);
        px[ch] = (uint8_t)__builtin_rintf(v * 255.0f);
    }
    px[3] = 255;
    *(uchar4*)(o + (size_t)i * 4) = make_uchar4(px[0], px[1], px[2], px[3]);
}

}  // namespace

void finish_frame_timings(Context* c) { collect_timings(c); }

}  // namespace gsr

using namespace gsr;

extern "C" {

int gsr_create(gsr_context** out, int device_ordinal) {
    if (!out) return fail(nullptr, GSR_ERR_ARGUMENT, "gsr_create: null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, GSR_ERR_DEVICE, "no HIP device available (%s)",
                    e != hipSuccess ? hipGetErrorString(e) : "device count is 0");
    if (device_ordinal < 0 || device_ordinal >= n)
        return fail(nullptr, GSR_ERR_ARGUMENT, "device ordinal %d out of range [0, %d)", device_ordinal, n);
    if ((e = hipSetDevice(device_ordinal)) != hipSuccess)
        return fail(nullptr, GSR_ERR_DEVICE, "hipSetDevice(%d): %s", device_ordinal, hipGetErrorString(e));
    Context* c = new Context();
    c->device = device_ordinal;
    if ((e = hipStreamCreate(&c->stream)) != hipSuccess) {
        delete c;
        return fail(nullptr, GSR_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess && prop.multiProcessorCount > 0)
        c->cu_count = prop.multiProcessorCount;
    noalloc_attach(c);
    if (hipHostMalloc((void**)&c->host_counters, sizeof(FrameCounters), hipHostMallocDefault) != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, GSR_ERR_ALLOC, "pinned host allocation failed");
    }
    *out = reinterpret_cast<gsr_context*>(c);
    return GSR_OK;
}

void gsr_destroy(gsr_context* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    destroy_batch_graph(c);
    if (c->staging) hipHostFree(c->staging);
    for (auto& kv : c->scenes) hipFree(kv.second.params);
    for (auto& kv : c->bufs) hipFree(kv.second.ptr);
    for (hipEvent_t e : c->event_pool) hipEventDestroy(e);
    if (c->host_counters) hipHostFree(c->host_counters);
    hipStreamDestroy(c->stream);
    delete c;
}

const char* gsr_last_error(const gsr_context* ctx) {
    const Context* c = reinterpret_cast<const Context*>(ctx);
    return c ? c->error.c_str() : g_create_error.c_str();
}

int gsr_synchronize(gsr_context* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    hipSetDevice(c->device);
    if (c->frame_unchecked) return check_frame(c, c->state_capacity);
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    return GSR_OK;
}

void* gsr_stream(gsr_context* ctx) {
    Context* c = reinterpret_cast<Context*>(ctx);
    return c ? (void*)c->stream : nullptr;
}

int gsr_scene_upload(gsr_context* ctx, const char* name, const float* packed, uint32_t count) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!name || (!packed && count > 0)) return fail(c, GSR_ERR_ARGUMENT, "gsr_scene_upload: null name or data");
    hipSetDevice(c->device);
    if (count == 0) {  // an empty scene renders black frames
        auto it = c->scenes.find(name);
        if (it != c->scenes.end()) hipFree(it->second.params);
        c->scenes[name] = Scene{nullptr, 0};
        if (c->current.empty()) c->current = name;
        return GSR_OK;
    }
    const size_t bytes = (size_t)count * kParams * sizeof(float);
    float* planar = nullptr;
    float* staging = nullptr;
    if (hipMalloc(&planar, bytes) != hipSuccess)
        return fail(c, GSR_ERR_ALLOC, "scene '%s': device allocation of %zu bytes failed", name, bytes);
    if (hipMalloc(&staging, bytes) != hipSuccess) {
        hipFree(planar);
        return fail(c, GSR_ERR_ALLOC, "scene '%s': staging allocation of %zu bytes failed", name, bytes);
    }
    hipError_t e = hipMemcpyAsync(staging, packed, bytes, hipMemcpyHostToDevice, c->stream);
    // Block layout sections (loader.js:75-85) -> planar channels.
    const uint32_t widths[6] = {3, 3, 4, 3, 1, 45};
    size_t off = 0;
    for (int s = 0; s < 6 && e == hipSuccess; ++s) {
        const size_t elems = (size_t)count * widths[s];
        const uint32_t blocks = (uint32_t)std::min<size_t>((elems + 255) / 256, 8192);
        hipLaunchKernelGGL(k_deinterleave, dim3(blocks), dim3(256), 0, c->stream, staging + off, count, widths[s],
                           planar + off);
        off += elems;
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(staging);
    if (e != hipSuccess) {
        hipFree(planar);
        return fail(c, GSR_ERR_DEVICE, "scene '%s' upload: %s", name, hipGetErrorString(e));
    }
    auto it = c->scenes.find(name);
    if (it != c->scenes.end()) hipFree(it->second.params);
    c->scenes[name] = Scene{planar, count};
    if (c->current.empty()) c->current = name;
    return GSR_OK;
}

int gsr_scene_select(gsr_context* ctx, const char* name) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!name || !c->scenes.count(name))
        return fail(c, GSR_ERR_NOT_FOUND, "Invalid Gs buffer for name \"%s\"", name ? name : "(null)");
    c->current = name;
    return GSR_OK;
}

int gsr_scene_free(gsr_context* ctx, const char* name) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    auto it = name ? c->scenes.find(name) : c->scenes.end();
    if (it == c->scenes.end()) return fail(c, GSR_ERR_NOT_FOUND, "unknown scene \"%s\"", name ? name : "(null)");
    hipStreamSynchronize(c->stream);
    hipFree(it->second.params);
    c->scenes.erase(it);
    if (c->current == name) c->current.clear();
    return GSR_OK;
}

int gsr_scene_params(gsr_context* ctx, const char* name, const float** params_dev, uint32_t* count) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    auto it = name ? c->scenes.find(name) : c->scenes.end();
    if (it == c->scenes.end()) return fail(c, GSR_ERR_NOT_FOUND, "unknown scene \"%s\"", name ? name : "(null)");
    if (params_dev) *params_dev = it->second.params;
    if (count) *count = it->second.count;
    return GSR_OK;
}

int gsr_get_timings(gsr_context* ctx, gsr_timing_row* rows, int capacity, int* count) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!count) return fail(c, GSR_ERR_ARGUMENT, "gsr_get_timings: null count");
    if (!c->marks.empty()) {
        GSR_HIP(c, hipStreamSynchronize(c->stream));
        collect_timings(c);
    }
    const int n = (int)c->last_rows.size();
    *count = n;
    if (rows)
        for (int i = 0; i < n && i < capacity; ++i) rows[i] = c->last_rows[i];
    return GSR_OK;
}

int gsr_to_rgba8(gsr_context* ctx, const float* rgb_dev, uint32_t width, uint32_t height, int flip_y,
                 uint8_t* out_rgba_host) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!out_rgba_host || width == 0 || height == 0) return fail(c, GSR_ERR_ARGUMENT, "gsr_to_rgba8: bad arguments");
    const size_t px = (size_t)width * height;
    if (!rgb_dev) {  // the context's last rendered frame
        if (c->last_frame_w != width || c->last_frame_h != height)
            return fail(c, GSR_ERR_STATE, "gsr_to_rgba8: no %ux%u frame was rendered last", width, height);
        rgb_dev = (const float*)buffer(c, "image", px * 12);
        if (!rgb_dev) return GSR_ERR_ALLOC;
    }
    uint8_t* d = (uint8_t*)buffer(c, "rgba8", px * 4);
    if (!d) return GSR_ERR_ALLOC;
    hipLaunchKernelGGL(k_rgba8, dim3((uint32_t)((px + 255) / 256)), dim3(256), 0, c->stream, rgb_dev, width, height,
                       flip_y, d);
    GSR_HIP(c, hipGetLastError());
    GSR_HIP(c, hipMemcpyAsync(out_rgba_host, d, px * 4, hipMemcpyDeviceToHost, c->stream));
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    return GSR_OK;
}

}  // extern "C"
