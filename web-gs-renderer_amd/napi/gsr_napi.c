/*
 * gsr_napi.c — Node N-API addon over the renderer's C ABI (include/gsr.h).
 *
 * The JS host (web-gs-renderer_amd/js) keeps the reference's surface
 * (src/renderer/renderer.js: uploadGsData / setGs / per-frame render,
 * src/gpu/profiler.js rows); this addon is the only native code it calls.
 * Every nonzero gsr status becomes a thrown JS Error carrying
 * gsr_last_error(), the observable convention of src/utils/utils.js:3-47.
 * Plain C N-API (node_api.h), so it loads in Node >= 12 without node-addon-api.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gsr.h"

typedef struct {
    gsr_context* ctx;
} handle_t;

#define CHECK(env, call)                                                        \
    do {                                                                        \
        if ((call) != napi_ok) {                                                \
            napi_throw_error((env), NULL, "gsr addon: N-API call failed: " #call); \
            return NULL;                                                        \
        }                                                                       \
    } while (0)

static napi_value throw_gsr(napi_env env, gsr_context* c, int rc) {
    char msg[640];
    snprintf(msg, sizeof(msg), "[gsr %d] %s", rc, gsr_last_error(c));
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static void finalize_handle(napi_env env, void* data, void* hint) {
    (void)env;
    (void)hint;
    handle_t* h = (handle_t*)data;
    if (h) {
        if (h->ctx) gsr_destroy(h->ctx);
        free(h);
    }
}

static handle_t* get_handle(napi_env env, napi_value v) {
    void* data = NULL;
    if (napi_get_value_external(env, v, &data) != napi_ok || !data) {
        napi_throw_type_error(env, NULL, "gsr addon: expected a context handle");
        return NULL;
    }
    handle_t* h = (handle_t*)data;
    if (!h->ctx) {
        napi_throw_error(env, NULL, "gsr addon: context already destroyed");
        return NULL;
    }
    return h;
}

static int get_string(napi_env env, napi_value v, char* buf, size_t cap) {
    size_t len = 0;
    return napi_get_value_string_utf8(env, v, buf, cap, &len) == napi_ok;
}

/* create(deviceOrdinal) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    int32_t device = 0;
    if (argc >= 1) napi_get_value_int32(env, argv[0], &device);
    gsr_context* c = NULL;
    int rc = gsr_create(&c, device);
    if (rc != GSR_OK) return throw_gsr(env, NULL, rc);
    handle_t* h = (handle_t*)calloc(1, sizeof(handle_t));
    h->ctx = c;
    napi_value out;
    CHECK(env, napi_create_external(env, h, finalize_handle, NULL, &out));
    return out;
}

/* destroy(handle) */
static napi_value js_destroy(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    void* data = NULL;
    if (argc >= 1 && napi_get_value_external(env, argv[0], &data) == napi_ok && data) {
        handle_t* h = (handle_t*)data;
        if (h->ctx) gsr_destroy(h->ctx);
        h->ctx = NULL;
    }
    return NULL;
}

/* uploadGsData(handle, name, Float32Array packed, count) */
static napi_value js_upload(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 4) {
        napi_throw_type_error(env, NULL, "uploadGsData(handle, name, packed, count)");
        return NULL;
    }
    handle_t* h = get_handle(env, argv[0]);
    if (!h) return NULL;
    char name[256];
    if (!get_string(env, argv[1], name, sizeof(name))) {
        napi_throw_type_error(env, NULL, "scene name must be a string");
        return NULL;
    }
    napi_typedarray_type type;
    size_t length = 0;
    void* data = NULL;
    napi_value ab;
    size_t off = 0;
    if (napi_get_typedarray_info(env, argv[2], &type, &length, &data, &ab, &off) != napi_ok ||
        type != napi_float32_array) {
        napi_throw_type_error(env, NULL, "packed scene data must be a Float32Array");
        return NULL;
    }
    uint32_t count = 0;
    CHECK(env, napi_get_value_uint32(env, argv[3], &count));
    if ((uint64_t)count * 59u != (uint64_t)length) {
        napi_throw_range_error(env, NULL, "packed scene data must hold 59 floats per splat");
        return NULL;
    }
    int rc = gsr_scene_upload(h->ctx, name, (const float*)data, count);
    if (rc != GSR_OK) return throw_gsr(env, h->ctx, rc);
    return NULL;
}

static napi_value name_call(napi_env env, napi_callback_info info, int (*fn)(gsr_context*, const char*)) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 2) {
        napi_throw_type_error(env, NULL, "expected (handle, name)");
        return NULL;
    }
    handle_t* h = get_handle(env, argv[0]);
    if (!h) return NULL;
    char name[256];
    if (!get_string(env, argv[1], name, sizeof(name))) {
        napi_throw_type_error(env, NULL, "scene name must be a string");
        return NULL;
    }
    int rc = fn(h->ctx, name);
    if (rc != GSR_OK) return throw_gsr(env, h->ctx, rc);
    return NULL;
}

static napi_value js_set_gs(napi_env env, napi_callback_info info) { return name_call(env, info, gsr_scene_select); }
static napi_value js_free_gs(napi_env env, napi_callback_info info) { return name_call(env, info, gsr_scene_free); }

static uint32_t opt_u32(napi_env env, napi_value obj, const char* key, uint32_t dflt) {
    bool has = false;
    napi_value v;
    uint32_t out = dflt;
    if (napi_has_named_property(env, obj, key, &has) == napi_ok && has &&
        napi_get_named_property(env, obj, key, &v) == napi_ok) {
        napi_valuetype t;
        napi_typeof(env, v, &t);
        if (t == napi_number) napi_get_value_uint32(env, v, &out);
        else if (t == napi_boolean) {
            bool b = false;
            napi_get_value_bool(env, v, &b);
            out = b ? 1u : 0u;
        }
    }
    return out;
}

/* render(handle, Float32Array(44) uniform, {tileSize, timing, rgba8, flipY})
 *   -> Float32Array planar RGB [3][H][W], or Uint8Array RGBA8 [H][W][4] */
static napi_value js_render(napi_env env, napi_callback_info info) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    if (argc < 2) {
        napi_throw_type_error(env, NULL, "render(handle, uniform, options)");
        return NULL;
    }
    handle_t* h = get_handle(env, argv[0]);
    if (!h) return NULL;
    napi_typedarray_type type;
    size_t length = 0;
    void* data = NULL;
    napi_value ab;
    size_t off = 0;
    if (napi_get_typedarray_info(env, argv[1], &type, &length, &data, &ab, &off) != napi_ok ||
        type != napi_float32_array || length != 44) {
        napi_throw_type_error(env, NULL, "the camera uniform must be a Float32Array of 44 slots (176 bytes)");
        return NULL;
    }
    gsr_camera cam;
    memcpy(&cam, data, sizeof(cam));
    gsr_options opts;
    memset(&opts, 0, sizeof(opts));
    uint32_t rgba8 = 0, flip = 0;
    if (argc >= 3) {
        napi_valuetype t;
        napi_typeof(env, argv[2], &t);
        if (t == napi_object) {
            opts.tile_size = opt_u32(env, argv[2], "tileSize", 16);
            opts.timing = opt_u32(env, argv[2], "timing", 0);
            rgba8 = opt_u32(env, argv[2], "rgba8", 0);
            flip = opt_u32(env, argv[2], "flipY", 0);
        }
    }
    const uint32_t W = (uint32_t)cam.texture_size[0], H = (uint32_t)cam.texture_size[1];
    if (W == 0 || H == 0) {
        napi_throw_range_error(env, NULL, "texture size must be positive");
        return NULL;
    }
    napi_value out_ab, out;
    void* out_data = NULL;
    if (rgba8) {
        int rc = gsr_render(h->ctx, &cam, &opts, NULL, NULL);
        if (rc != GSR_OK) return throw_gsr(env, h->ctx, rc);
        CHECK(env, napi_create_arraybuffer(env, (size_t)W * H * 4, &out_data, &out_ab));
        rc = gsr_to_rgba8(h->ctx, NULL, W, H, (int)flip, (uint8_t*)out_data);
        if (rc != GSR_OK) return throw_gsr(env, h->ctx, rc);
        CHECK(env, napi_create_typedarray(env, napi_uint8_array, (size_t)W * H * 4, out_ab, 0, &out));
    } else {
        CHECK(env, napi_create_arraybuffer(env, (size_t)W * H * 12, &out_data, &out_ab));
        int rc = gsr_render(h->ctx, &cam, &opts, (float*)out_data, NULL);
        if (rc != GSR_OK) return throw_gsr(env, h->ctx, rc);
        CHECK(env, napi_create_typedarray(env, napi_float32_array, (size_t)W * H * 3, out_ab, 0, &out));
    }
    return out;
}

/* timings(handle) -> [[label, ms], ...] of the frames since the last call */
static napi_value js_timings(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    handle_t* h = argc >= 1 ? get_handle(env, argv[0]) : NULL;
    if (!h) return NULL;
    int n = 0;
    int rc = gsr_get_timings(h->ctx, NULL, 0, &n);
    if (rc != GSR_OK) return throw_gsr(env, h->ctx, rc);
    gsr_timing_row* rows = (gsr_timing_row*)calloc(n > 0 ? (size_t)n : 1, sizeof(gsr_timing_row));
    rc = gsr_get_timings(h->ctx, rows, n, &n);
    if (rc != GSR_OK) {
        free(rows);
        return throw_gsr(env, h->ctx, rc);
    }
    napi_value arr;
    napi_create_array_with_length(env, (size_t)n, &arr);
    for (int i = 0; i < n; ++i) {
        napi_value pair, label, ms;
        napi_create_array_with_length(env, 2, &pair);
        napi_create_string_utf8(env, rows[i].label, NAPI_AUTO_LENGTH, &label);
        napi_create_double(env, rows[i].ms, &ms);
        napi_set_element(env, pair, 0, label);
        napi_set_element(env, pair, 1, ms);
        napi_set_element(env, arr, (uint32_t)i, pair);
    }
    free(rows);
    return arr;
}

/* instanceCount(handle) -> tile instances of the last frame */
static napi_value js_instances(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    handle_t* h = argc >= 1 ? get_handle(env, argv[0]) : NULL;
    if (!h) return NULL;
    uint64_t n = 0;
    int rc = gsr_last_instance_count(h->ctx, &n);
    if (rc != GSR_OK) return throw_gsr(env, h->ctx, rc);
    napi_value out;
    napi_create_double(env, (double)n, &out);
    return out;
}

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor props[] = {
        {"create", NULL, js_create, NULL, NULL, NULL, napi_enumerable, NULL},
        {"destroy", NULL, js_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
        {"uploadGsData", NULL, js_upload, NULL, NULL, NULL, napi_enumerable, NULL},
        {"setGs", NULL, js_set_gs, NULL, NULL, NULL, napi_enumerable, NULL},
        {"freeGs", NULL, js_free_gs, NULL, NULL, NULL, napi_enumerable, NULL},
        {"render", NULL, js_render, NULL, NULL, NULL, napi_enumerable, NULL},
        {"timings", NULL, js_timings, NULL, NULL, NULL, napi_enumerable, NULL},
        {"instanceCount", NULL, js_instances, NULL, NULL, NULL, napi_enumerable, NULL},
    };
    napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
