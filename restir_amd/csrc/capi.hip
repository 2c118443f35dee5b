// The C ABI (include/restir_amd.h): argument checks, error text, the
// context's scratch memory, and dispatch to the kernel files.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>

#include "internal.h"

// rgi_grid_insert_noalloc's scratch: the deterministic insert's scratch at
// its largest admitted inputs, one copy per GPU. In host code this name is a
// host-side handle, not a device address: kernels and copies get the address
// looked up per device by noalloc_scratch() below.
#define RGI_NOALLOC_SCRATCH_BYTES (88ull * 1024 * 1024 + 256 + 8448)
static __device__ __align__(256) unsigned char scratch[RGI_NOALLOC_SCRATCH_BYTES];

namespace rgi {

static_assert(RGI_NOALLOC_SCRATCH_BYTES == 6 * 4 * (size_t)RGI_NOALLOC_MAX_PIXELS + 4 * (size_t)RGI_NOALLOC_MAX_TABLE +
                                                3 * 4 * (size_t)RGI_NOALLOC_MAX_PIXELS + 256 +
                                                ((4 * ((size_t)RGI_NOALLOC_MAX_PIXELS / 1024 + 1) + 255) & ~(size_t)255),
              "the no-alloc scratch holds the deterministic insert's scratch at the largest admitted inputs");

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static void* noalloc_scratch(int& error)
{
    static std::mutex mu;
    static void* address[64] = {};
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) {
        error = fail(RGI_E_DEVICE, "grid_insert_noalloc: no current HIP device (or device index >= 64)");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!address[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || !p) {
            error = fail(RGI_E_DEVICE, "grid_insert_noalloc: scratch address lookup failed: %s", hipGetErrorString(e));
            return nullptr;
        }
        address[device] = p;
    }
    return address[device];
}

static thread_local char g_error[512] = "";

int fail(int code, const char* fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, args);
    va_end(args);
    return code;
}

int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(RGI_E_DEVICE, "%s: launch failed: %s", what, hipGetErrorString(e));
    return RGI_OK;
}

void* context_scratch(rgi_context* ctx, size_t bytes)
{
    if (ctx->scratch_bytes >= bytes) return ctx->scratch;
    if (ctx->scratch) {
        // Work queued on any stream may still use the old buffer.
        hipDeviceSynchronize();
        hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t grown = bytes + bytes / 4;
    if (hipMalloc(&ctx->scratch, grown) != hipSuccess) {
        ctx->scratch = nullptr;
        fail(RGI_E_NOMEM, "scratch allocation of %zu bytes failed", grown);
        return nullptr;
    }
    ctx->scratch_bytes = grown;
    return ctx->scratch;
}

void* context_buffer(rgi_context* ctx, ContextBuffer which, size_t bytes)
{
    if (ctx->buf_bytes[which] >= bytes) return ctx->bufs[which];
    if (ctx->bufs[which]) {
        hipDeviceSynchronize();
        hipFree(ctx->bufs[which]);
        ctx->bufs[which] = nullptr;
        ctx->buf_bytes[which] = 0;
    }
    if (hipMalloc(&ctx->bufs[which], bytes) != hipSuccess) {
        ctx->bufs[which] = nullptr;
        fail(RGI_E_NOMEM, "context buffer %d: allocation of %zu bytes failed", (int)which, bytes);
        return nullptr;
    }
    ctx->buf_bytes[which] = bytes;
    return ctx->bufs[which];
}

int context_compute_units(rgi_context* ctx)
{
    if (ctx->compute_units <= 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || n <= 0) n = 256;
        ctx->compute_units = n;
    }
    return ctx->compute_units;
}

static bool pow2(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }
static bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }

static int check_frame(const rgi_frame* f)
{
    if (!f) return fail(RGI_E_INVALID, "frame is NULL");
    if (!pow2(f->grid_table_size) || f->grid_table_size < 64u || f->grid_table_size > (1u << 26))
        return fail(RGI_E_INVALID, "grid_table_size %u is not a power of two in [64, 2^26]", f->grid_table_size);
    if (f->ray_mode < RGI_RAY_SCREEN || f->ray_mode > RGI_RAY_BVH)
        return fail(RGI_E_INVALID, "ray_mode %d is not one of RGI_RAY_*", f->ray_mode);
    if (f->initial_candidates < 0 || f->temporal_taps < 0 || f->spatial_taps < 0 || f->trace_steps < 0 ||
        f->di_candidates < 0 || f->di_spatial_taps < 0 || f->di_light_count < 0)
        return fail(RGI_E_INVALID, "negative sample counts in frame");
    return RGI_OK;
}

// The GI size (u_screenInfo.xy) of a pass over GI-size planes.
static int check_gi_size(const rgi_frame* f)
{
    if (f->gi_width <= 0 || f->gi_height <= 0 || (int64_t)f->gi_width * f->gi_height >= (1ll << 29))
        return fail(RGI_E_INVALID, "frame GI size %d x %d out of range", f->gi_width, f->gi_height);
    return RGI_OK;
}

// The GI passes read a window-size G-buffer at GI-pixel centres; the
// reference's GI size is the window size over a divisor (renderer.cpp:519-520),
// so a G-buffer smaller than the GI size is a caller error.
static int check_gi_gbuffer(const rgi_frame* f, const rgi_gbuffer* g)
{
    int e = check_gi_size(f);
    if (e != RGI_OK) return e;
    if (g->width < f->gi_width || g->height < f->gi_height)
        return fail(RGI_E_INVALID, "gbuffer %d x %d is smaller than the frame's GI size %d x %d", g->width, g->height,
                    f->gi_width, f->gi_height);
    return RGI_OK;
}

static int check_gbuffer(const rgi_gbuffer* g, bool need_albedo, bool need_shading)
{
    if (!g) return fail(RGI_E_INVALID, "gbuffer is NULL");
    if (g->width <= 0 || g->height <= 0 || (int64_t)g->width * g->height >= (1ll << 29))
        return fail(RGI_E_INVALID, "gbuffer size %d x %d out of range", g->width, g->height);
    if (!g->depth || !g->normal) return fail(RGI_E_INVALID, "gbuffer depth and normal planes are required");
    if (need_albedo && !g->albedo) return fail(RGI_E_INVALID, "gbuffer albedo plane is required");
    if (need_shading && !g->shading_normal) return fail(RGI_E_INVALID, "gbuffer shading_normal plane is required");
    return RGI_OK;
}

static int check_scene(const rgi_scene* s, bool need_lights)
{
    if (!s) return fail(RGI_E_INVALID, "scene is NULL");
    if (s->node_count > 0 && !s->bvh) return fail(RGI_E_INVALID, "scene has nodes but no bvh buffer");
    if (s->node_count > 0 && s->triangle_offset != 2u * s->node_count)
        return fail(RGI_E_INVALID, "triangle_offset %u != 2 * node_count %u", s->triangle_offset, s->node_count);
    if (s->bvh && !aligned(s->bvh, 16)) return fail(RGI_E_INVALID, "bvh buffer must be 16-byte aligned");
    if (need_lights && s->light_count > 0 && (!s->lights || !aligned(s->lights, 16)))
        return fail(RGI_E_INVALID, "lights buffer missing or not 16-byte aligned");
    return RGI_OK;
}

} // namespace rgi

using namespace rgi;

extern "C" {

const char* rgi_last_error(void) { return g_error; }
int rgi_abi_version(void) { return RGI_ABI_VERSION; }

uint32_t rgi_abi_sizeof(int which)
{
    switch (which) {
    case 0: return (uint32_t)sizeof(rgi_frame);
    case 1: return (uint32_t)sizeof(rgi_scene);
    case 2: return (uint32_t)sizeof(rgi_gbuffer);
    case 3: return (uint32_t)sizeof(rgi_gi_reservoirs);
    case 4: return (uint32_t)sizeof(rgi_gi_inputs);
    case 5: return (uint32_t)sizeof(rgi_di_reservoirs);
    case 6: return (uint32_t)sizeof(rgi_shadow);
    case 7: return (uint32_t)sizeof(rgi_composite_inputs);
    case 8: return (uint32_t)sizeof(rgi_nrc_config);
    case 9: return (uint32_t)sizeof(rgi_nrc_settings);
    default: return 0;
    }
}

int rgi_create(rgi_context** out)
{
    if (!out) return fail(RGI_E_INVALID, "out_context is NULL");
    rgi_context* ctx = new (std::nothrow) rgi_context();
    if (!ctx) return fail(RGI_E_NOMEM, "context allocation failed");
    if (hipGetDevice(&ctx->device) != hipSuccess) {
        delete ctx;
        return fail(RGI_E_DEVICE, "no HIP device: %s", hipGetErrorString(hipGetLastError()));
    }
    *out = ctx;
    return RGI_OK;
}

void rgi_destroy(rgi_context* ctx)
{
    if (!ctx) return;
    bool any = ctx->scratch || ctx->counter;
    for (void* b : ctx->bufs) any = any || b;
    if (any) hipDeviceSynchronize();
    if (ctx->scratch) hipFree(ctx->scratch);
    for (void* b : ctx->bufs)
        if (b) hipFree(b);
    if (ctx->counter) hipFree(ctx->counter);
    delete ctx;
}

int rgi_context_set_option(rgi_context* ctx, int option, int value)
{
    if (!ctx) return fail(RGI_E_INVALID, "context is NULL");
    switch (option) {
    case RGI_OPT_GI_TEMPORAL_WAVES:
        if (value != 0 && value != 4) return fail(RGI_E_INVALID, "GI temporal waves option takes 0 or 4, not %d", value);
        ctx->tuning.gi_temporal_waves = value;
        return RGI_OK;
    case RGI_OPT_GI_SPATIAL_WAVES:
        if (value != 0 && value != 5) return fail(RGI_E_INVALID, "GI spatial waves option takes 0 or 5, not %d", value);
        ctx->tuning.gi_spatial_waves = value;
        return RGI_OK;
    case RGI_OPT_TRACE_PERSISTENT:
        if (value != 0 && value != 1) return fail(RGI_E_INVALID, "persistent option takes 0 or 1, not %d", value);
        if (value && !ctx->counter && hipMalloc((void**)&ctx->counter, 1024) != hipSuccess)
            return fail(RGI_E_NOMEM, "work counter allocation failed");
        ctx->tuning.trace_persistent = value;
        return RGI_OK;
    case RGI_OPT_TILE_ORDER:
        if (value < 0 || value > 2) return fail(RGI_E_INVALID, "tile order takes 0, 1 or 2, not %d", value);
        ctx->tuning.tile_order = value;
        return RGI_OK;
    case RGI_OPT_TILE_SHAPE:
        if (value != 0 && value != 1) return fail(RGI_E_INVALID, "tile shape takes 0 or 1, not %d", value);
        ctx->tuning.tile_shape = value;
        return RGI_OK;
    case RGI_OPT_TRIANGLE_TEST:
        if (value != 0 && value != 1) return fail(RGI_E_INVALID, "triangle test takes 0 or 1, not %d", value);
        ctx->tuning.watertight = value;
        return RGI_OK;
    default:
        return fail(RGI_E_INVALID, "unknown option %d", option);
    }
}

int rgi_context_get_option(const rgi_context* ctx, int option)
{
    if (!ctx) return fail(RGI_E_INVALID, "context is NULL");
    switch (option) {
    case RGI_OPT_GI_TEMPORAL_WAVES: return ctx->tuning.gi_temporal_waves;
    case RGI_OPT_GI_SPATIAL_WAVES: return ctx->tuning.gi_spatial_waves;
    case RGI_OPT_TRACE_PERSISTENT: return ctx->tuning.trace_persistent;
    case RGI_OPT_TILE_ORDER: return ctx->tuning.tile_order;
    case RGI_OPT_TILE_SHAPE: return ctx->tuning.tile_shape;
    case RGI_OPT_TRIANGLE_TEST: return ctx->tuning.watertight;
    default: return fail(RGI_E_INVALID, "unknown option %d", option);
    }
}

int rgi_grid_clear(uint32_t* grid, uint32_t table_size, void* stream)
{
    if (!grid) return fail(RGI_E_INVALID, "grid is NULL");
    if (!aligned(grid, 16)) return fail(RGI_E_INVALID, "grid must be 16-byte aligned");
    if (table_size < 64u || (table_size & (table_size - 1u))) return fail(RGI_E_INVALID, "table_size must be a power of two >= 64");
    return grid_clear(grid, table_size, as_stream(stream));
}

int rgi_grid_insert(rgi_context* ctx, const rgi_frame* frame, const rgi_gbuffer* g, uint32_t* grid, int mode,
                    void* stream)
{
    int e;
    if ((e = check_frame(frame)) != RGI_OK) return e;
    if ((e = check_gbuffer(g, false, false)) != RGI_OK) return e;
    if ((e = check_gi_gbuffer(frame, g)) != RGI_OK) return e;
    if (!grid) return fail(RGI_E_INVALID, "grid is NULL");
    if (mode == RGI_GRID_ATOMIC) return grid_insert_atomic(*frame, *g, grid, as_stream(stream));
    if (mode == RGI_GRID_DETERMINISTIC) {
        if (!ctx) return fail(RGI_E_INVALID, "the deterministic insert needs a context");
        return grid_insert_deterministic(ctx, *frame, *g, grid, as_stream(stream));
    }
    return fail(RGI_E_INVALID, "unknown grid insert mode %d", mode);
}

int rgi_grid_insert_noalloc(const rgi_frame* frame, const rgi_gbuffer* g, uint32_t* grid, void* stream)
{
    int e;
    if ((e = check_frame(frame)) != RGI_OK) return e;
    if ((e = check_gbuffer(g, false, false)) != RGI_OK) return e;
    if ((e = check_gi_gbuffer(frame, g)) != RGI_OK) return e;
    if (!grid) return fail(RGI_E_INVALID, "grid is NULL");
    uint64_t pixels = (uint64_t)frame->gi_width * (uint64_t)frame->gi_height;
    if (pixels > RGI_NOALLOC_MAX_PIXELS || frame->grid_table_size > RGI_NOALLOC_MAX_TABLE)
        return fail(RGI_E_INVALID,
                    "grid_insert_noalloc: %llu GI pixels and a table of %u cells exceed its fixed scratch (at most %u "
                    "pixels and %u cells); use rgi_grid_insert",
                    (unsigned long long)pixels, frame->grid_table_size, RGI_NOALLOC_MAX_PIXELS, RGI_NOALLOC_MAX_TABLE);
    if (det_scratch_bytes((uint32_t)pixels, frame->grid_table_size) > RGI_NOALLOC_SCRATCH_BYTES)
        return fail(RGI_E_INVALID, "grid_insert_noalloc: scratch too small");
    void* base = noalloc_scratch(e);
    if (!base) return e;
    return grid_insert_deterministic_in(base, *frame, *g, grid, as_stream(stream));
}

int rgi_gi_temporal(rgi_context* ctx, const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* g,
                    const rgi_gi_inputs* in, const rgi_gi_reservoirs* out, void* stream)
{
    int e;
    if ((e = check_frame(frame)) != RGI_OK) return e;
    if ((e = check_gbuffer(g, true, false)) != RGI_OK) return e;
    if ((e = check_scene(scene, false)) != RGI_OK) return e;
    if (!in || !out) return fail(RGI_E_INVALID, "inputs and outputs are required");
    if (!in->history_sample || !in->history_receiver || !in->history_normal || !in->history_color || !in->surface_light ||
        !in->grid)
        return fail(RGI_E_INVALID, "gi temporal: history planes, surface light and previous grid are required");
    if (frame->shadows_enabled && !in->shadow_map) return fail(RGI_E_INVALID, "shadows enabled but shadow_map is NULL");
    if (!out->sample || !out->receiver || !out->normal) return fail(RGI_E_INVALID, "gi temporal: output planes are required");
    if ((e = check_gi_gbuffer(frame, g)) != RGI_OK) return e;
    return gi_temporal(ctx, *frame, *scene, *g, *in, *out, ctx ? ctx->tuning : Tuning{}, as_stream(stream));
}

int rgi_gi_spatial(rgi_context* ctx, const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* g,
                   const rgi_gi_inputs* in, const rgi_gi_reservoirs* res, uint16_t* out_color, uint16_t* out_moments,
                   void* stream)
{
    int e;
    if ((e = check_frame(frame)) != RGI_OK) return e;
    if ((e = check_gbuffer(g, true, false)) != RGI_OK) return e;
    if ((e = check_scene(scene, false)) != RGI_OK) return e;
    if (!in || !res || !out_color || !out_moments) return fail(RGI_E_INVALID, "gi spatial: inputs and outputs are required");
    if (!in->history_receiver || !in->history_normal || !in->history_color || !in->history_moments || !in->surface_light ||
        !in->grid)
        return fail(RGI_E_INVALID, "gi spatial: history planes, surface light and grid are required");
    if (frame->shadows_enabled && !in->shadow_map) return fail(RGI_E_INVALID, "shadows enabled but shadow_map is NULL");
    if (!res->sample || !res->receiver || !res->normal) return fail(RGI_E_INVALID, "gi spatial: reservoir planes are required");
    if ((e = check_gi_gbuffer(frame, g)) != RGI_OK) return e;
    return gi_spatial(*frame, *scene, *g, *in, *res, out_color, out_moments, ctx ? ctx->tuning : Tuning{},
                      as_stream(stream));
}

int rgi_di_temporal(rgi_context* ctx, const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* g,
                    const rgi_di_reservoirs* prev, const rgi_di_reservoirs* out, void* stream)
{
    (void)ctx;
    int e;
    if ((e = check_frame(frame)) != RGI_OK) return e;
    if ((e = check_gbuffer(g, true, true)) != RGI_OK) return e;
    if ((e = check_scene(scene, true)) != RGI_OK) return e;
    if ((uint32_t)frame->di_light_count > scene->light_count)
        return fail(RGI_E_INVALID, "di_light_count %d exceeds scene light_count %u", frame->di_light_count, scene->light_count);
    if (!prev || !out || !prev->sample || !prev->receiver || !prev->normal || !out->sample || !out->receiver || !out->normal)
        return fail(RGI_E_INVALID, "di temporal: reservoir planes are required");
    return di_temporal(ctx, *frame, *scene, *g, *prev, *out, ctx ? ctx->tuning : Tuning{}, as_stream(stream));
}

int rgi_di_spatial(rgi_context* ctx, const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* g,
                   const rgi_di_reservoirs* res, uint16_t* out_diffuse, uint16_t* out_specular, void* stream)
{
    (void)ctx;
    int e;
    if ((e = check_frame(frame)) != RGI_OK) return e;
    if ((e = check_gbuffer(g, true, true)) != RGI_OK) return e;
    if ((e = check_scene(scene, true)) != RGI_OK) return e;
    if ((uint32_t)frame->di_light_count > scene->light_count)
        return fail(RGI_E_INVALID, "di_light_count %d exceeds scene light_count %u", frame->di_light_count, scene->light_count);
    if (!res || !res->sample || !res->receiver || !res->normal || !out_diffuse || !out_specular)
        return fail(RGI_E_INVALID, "di spatial: reservoir and output planes are required");
    return di_spatial(*frame, *scene, *g, *res, out_diffuse, out_specular, ctx ? ctx->tuning : Tuning{},
                      as_stream(stream));
}

int rgi_trace(rgi_context* ctx, const rgi_scene* scene, const float* rays, uint32_t count, int32_t* hit_triangle,
              float* hit_distance, int any_hit, void* stream)
{
    int e;
    if ((e = check_scene(scene, false)) != RGI_OK) return e;
    if (count > 0 && (!rays || !hit_triangle || !hit_distance)) return fail(RGI_E_INVALID, "trace: buffers are required");
    if (rays && !aligned(rays, 16)) return fail(RGI_E_INVALID, "rays must be 16-byte aligned");
    return trace_rays(*scene, rays, count, hit_triangle, hit_distance, any_hit, ctx ? ctx->tuning : Tuning{},
                      ctx ? ctx->counter : nullptr, as_stream(stream));
}

int rgi_shadow_apply(const rgi_shadow* shadow, int enabled, rgi_frame* frame)
{
    if (!shadow || !frame) return fail(RGI_E_INVALID, "shadow_apply: NULL argument");
    memcpy(frame->shadow_view_proj, shadow->view_proj, sizeof(frame->shadow_view_proj));
    memcpy(frame->shadow_info, shadow->info, sizeof(frame->shadow_info));
    frame->shadow_width = 3 * shadow->face_size;
    frame->shadow_height = 2 * shadow->face_size;
    frame->shadows_enabled = enabled ? 1 : 0;
    return RGI_OK;
}

int rgi_shadow_atlas(rgi_context* ctx, const rgi_shadow* shadow, const rgi_scene* scene, float* atlas, void* stream)
{
    (void)ctx;
    int e;
    if (!shadow || !atlas) return fail(RGI_E_INVALID, "shadow_atlas: shadow and atlas are required");
    if ((e = check_scene(scene, false)) != RGI_OK) return e;
    int32_t F = shadow->face_size;
    if (F != 256 && F != 512 && F != 1024 && F != 2048)
        return fail(RGI_E_INVALID, "shadow face size must be 256, 512, 1024 or 2048, not %d", F);
    return shadow_atlas(*shadow, *scene, atlas, as_stream(stream));
}

int rgi_specular_gi(rgi_context* ctx, const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* g,
                    const rgi_gi_inputs* in, const uint16_t* previous_hdr, const uint16_t* history, uint16_t* out,
                    void* stream)
{
    int e;
    if ((e = check_frame(frame)) != RGI_OK) return e;
    if ((e = check_scene(scene, false)) != RGI_OK) return e;
    if ((e = check_gbuffer(g, true, false)) != RGI_OK) return e;
    if ((e = check_gi_gbuffer(frame, g)) != RGI_OK) return e;
    if (!in || !out) return fail(RGI_E_INVALID, "specular_gi: inputs and output are required");
    if (frame->history_valid && (!in->history_receiver || !in->history_normal || !in->history_color))
        return fail(RGI_E_INVALID, "specular_gi: history planes are required when history_valid");
    if (!in->surface_light) return fail(RGI_E_INVALID, "specular_gi: surface_light is required");
    if (history == out) return fail(RGI_E_INVALID, "specular_gi: out must not alias history");
    return specular_gi(*frame, *scene, *g, *in, previous_hdr, history, out, ctx ? ctx->tuning : Tuning{},
                       as_stream(stream));
}

int rgi_composite(rgi_context* ctx, const rgi_frame* frame, const rgi_gbuffer* g, const rgi_composite_inputs* in,
                  uint16_t* out, void* stream)
{
    (void)ctx;
    int e;
    if (!frame || !in || !out) return fail(RGI_E_INVALID, "composite: frame, inputs and output are required");
    if (frame->width <= 0 || frame->height <= 0) return fail(RGI_E_INVALID, "frame width/height must be positive");
    if ((e = check_gbuffer(g, true, true)) != RGI_OK) return e;
    if ((e = check_gi_gbuffer(frame, g)) != RGI_OK) return e;
    if (!g->emission) return fail(RGI_E_INVALID, "composite: the emission plane is required");
    if (!in->gi || !aligned(in->gi, 8)) return fail(RGI_E_INVALID, "composite: the GI plane is required (8-byte aligned)");
    if (frame->shadows_enabled && in->shadow_map && (frame->shadow_width <= 0 || frame->shadow_height <= 0))
        return fail(RGI_E_INVALID, "composite: shadows on without an atlas size");
    return composite(*frame, *g, *in, out, as_stream(stream));
}

int rgi_svgf_variance(rgi_context* ctx, const rgi_frame* frame, const uint16_t* color, const uint16_t* moments,
                      const float* receiver, const float* normal, uint16_t* out, void* stream)
{
    (void)ctx;
    if (!frame) return fail(RGI_E_INVALID, "frame is NULL");
    int e = check_gi_size(frame);
    if (e != RGI_OK) return e;
    if (!color || !moments || !receiver || !normal || !out) return fail(RGI_E_INVALID, "svgf_variance: NULL plane");
    if (!aligned(color, 8) || !aligned(moments, 8) || !aligned(out, 8) || !aligned(receiver, 16) || !aligned(normal, 16))
        return fail(RGI_E_INVALID, "svgf_variance: misaligned plane");
    return svgf_variance(*frame, color, moments, receiver, normal, out, as_stream(stream));
}

int rgi_svgf_atrous(rgi_context* ctx, const rgi_frame* frame, const uint16_t* input, const float* receiver,
                    const float* normal, int32_t step, uint16_t* out, void* stream)
{
    (void)ctx;
    if (!frame) return fail(RGI_E_INVALID, "frame is NULL");
    int e = check_gi_size(frame);
    if (e != RGI_OK) return e;
    if (!input || !receiver || !normal || !out) return fail(RGI_E_INVALID, "svgf_atrous: NULL plane");
    if (input == out) return fail(RGI_E_INVALID, "svgf_atrous: out must not alias input");
    if (step < 1 || step > 1 << 20) return fail(RGI_E_INVALID, "svgf_atrous: step %d out of range", step);
    if (!aligned(input, 8) || !aligned(out, 8) || !aligned(receiver, 16) || !aligned(normal, 16))
        return fail(RGI_E_INVALID, "svgf_atrous: misaligned plane");
    return svgf_atrous(*frame, input, receiver, normal, step, out, as_stream(stream));
}

int rgi_gbuffer_primary(rgi_context* ctx, const rgi_frame* frame, const rgi_scene* scene, const float* materials,
                        uint32_t material_count, float* depth, uint32_t* albedo, uint32_t* normal,
                        uint32_t* shading_normal, uint16_t* emission, void* stream)
{
    (void)ctx;
    int e;
    if (!frame) return fail(RGI_E_INVALID, "frame is NULL");
    if ((e = check_scene(scene, false)) != RGI_OK) return e;
    if (frame->width <= 0 || frame->height <= 0) return fail(RGI_E_INVALID, "frame width/height must be positive");
    if (!depth || !albedo || !normal || !shading_normal || !emission) return fail(RGI_E_INVALID, "gbuffer outputs are required");
    if (material_count > 0 && (!materials || !aligned(materials, 16))) return fail(RGI_E_INVALID, "materials missing or misaligned");
    return gbuffer_primary(*frame, *scene, materials, material_count, depth, albedo, normal, shading_normal, emission,
                           as_stream(stream));
}

} // extern "C"
