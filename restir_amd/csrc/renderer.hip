// The frame driver (include/restir_amd_renderer.h): Renderer::render's hot
// path (src/renderer.cpp:435-531, 614-833, 1119-1142) in C++ over the
// per-pass kernels, with the reference's ping-pong and validity flags.
#include <math.h>
#include <string.h>

#include <new>

#include "../../include/restir_amd_renderer.h"
#include "batch.h"
#include "internal.h"

#include <vector>

struct rgi_renderer {
    int32_t width = 0, height = 0;       // window (G-buffer, DI, composite)
    int32_t gi_width = 0, gi_height = 0; // GI planes: window / gi_divisor
    rgi_settings settings{};
    rgi_scene scene{};
    float radius = 1.0f;
    rgi_context* ctx = nullptr;
    // Per parity: GI reservoirs, colour, moments, grid, DI reservoirs.
    float* gi[2][3] = {};
    uint16_t* color[2] = {};
    uint16_t* moments[2] = {};
    uint32_t* grid[2] = {};
    float* di_sample[2] = {};
    float* di_receiver[2] = {};
    uint16_t* di_normal[2] = {};
    uint16_t* diffuse = nullptr;
    uint16_t* specular = nullptr;
    // The main light's shadow atlas: rendered on the first frame with
    // shadows on; scene and light are fixed for the renderer's lifetime, so
    // every later frame's shadow pass would write the same texels.
    float* atlas = nullptr;
    bool atlas_valid = false;
    // Denoiser targets (renderer.cpp:1424-1430): the filtered history per
    // parity and two ping-pong filter targets, RGBA16F.
    uint16_t* filtered[2] = {};
    uint16_t* filter[2] = {};
    const uint16_t* denoised = nullptr;
    uint16_t* hdr = nullptr; // the composite (m_hdrTexture, RGBA16F)
    uint16_t* black = nullptr; // m_blackTexture: DI light when ReSTIR DI is selected but has no lights
    uint16_t* specular_gi[2] = {}; // m_specularTextures, RGBA16F, own parity m_specularRead
    int specular_read = 0;
    // The side stream that traces GI temporal's candidates and builds the
    // frame's hash grid beside the DI passes (default priority), with its
    // fork event and the events that end its two parts.
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr, side_trace = nullptr, side_join = nullptr;
    hipStream_t grid_side = nullptr; // the hash grid's stream, beside the side stream

    // The cache's training step runs on its own stream (default priority):
    // `nrc_pending` marks a step that the caller's stream has not joined yet.
    hipStream_t nrc_stream = nullptr;
    hipEvent_t nrc_fork = nullptr, nrc_done = nullptr;
    bool nrc_pending = false;
    // The graph form of a frame (rgi_renderer_render_graph).
    hipGraphExec_t graph_exec = nullptr;
    bool graph_warm = false;
    // The neural radiance cache (rgi_renderer_set_nrc).
    rgi_nrc* nrc = nullptr;
    rgi_nrc_settings nrc_settings{};
    uint16_t* nrc_plane = nullptr;
    uint16_t* nrc_resolved = nullptr; // the query without emission, when the plane above carries it
    float* nrc_records = nullptr;
    float* nrc_targets = nullptr;
    float* nrc_loss = nullptr;
    uint32_t nrc_trained = 0;
    int read = 0;              // m_historyRead / m_diRead: the parity last frame wrote
    bool history_valid = false;
    bool di_valid = false;
    bool reference = false;
    uint32_t frame_number = 0;
    rgi_view previous{};
    rgi_frame last{};
    bool have_frame = false;
};

namespace {

using namespace rgi;

template <typename T>
bool alloc(T*& p, size_t count)
{
    void* q = nullptr;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return false;
    if (hipMemset(q, 0, count * sizeof(T)) != hipSuccess) return false;
    p = static_cast<T*>(q);
    return true;
}

template <typename T>
void release(T*& p)
{
    if (p) hipFree(p);
    p = nullptr;
}

void release_nrc(rgi_renderer* r)
{
    if (r->nrc) rgi_nrc_destroy(r->nrc);
    r->nrc = nullptr;
    release(r->nrc_plane);
    release(r->nrc_resolved);
    release(r->nrc_records);
    release(r->nrc_targets);
    release(r->nrc_loss);
    r->nrc_trained = 0;
    r->nrc_pending = false;
}

void free_all(rgi_renderer* r)
{
    for (int k = 0; k < 2; ++k) {
        for (int j = 0; j < 3; ++j) release(r->gi[k][j]);
        release(r->color[k]);
        release(r->moments[k]);
        release(r->grid[k]);
        release(r->di_sample[k]);
        release(r->di_receiver[k]);
        release(r->di_normal[k]);
    }
    release(r->diffuse);
    release(r->specular);
    release(r->atlas);
    release(r->hdr);
    release(r->black);
    release(r->specular_gi[0]);
    release(r->specular_gi[1]);
    for (int k = 0; k < 2; ++k) {
        release(r->filtered[k]);
        release(r->filter[k]);
    }
    release_nrc(r);
    if (r->nrc_fork) hipEventDestroy(r->nrc_fork);
    if (r->nrc_done) hipEventDestroy(r->nrc_done);
    if (r->nrc_stream) hipStreamDestroy(r->nrc_stream);
    r->nrc_fork = r->nrc_done = nullptr;
    r->nrc_stream = nullptr;
    r->nrc_pending = false;
    if (r->side_fork) hipEventDestroy(r->side_fork);
    if (r->side_trace) hipEventDestroy(r->side_trace);
    if (r->side_join) hipEventDestroy(r->side_join);
    if (r->side) hipStreamDestroy(r->side);
    if (r->grid_side) hipStreamDestroy(r->grid_side);
    r->grid_side = nullptr;
    r->side_fork = r->side_trace = r->side_join = nullptr;
    r->side = nullptr;
}

// kShadowFaceSizes (settings.h:26); 0 for an index out of range.
int32_t shadow_face_size(const rgi_settings* s)
{
    static const int32_t sizes[4] = {256, 512, 1024, 2048};
    return s->shadow_face_size >= 0 && s->shadow_face_size < 4 ? sizes[s->shadow_face_size] : 0;
}

// Effective settings while the reference accumulates (renderer.cpp:437-452).
rgi_settings effective(const rgi_renderer* r)
{
    rgi_settings s = r->settings;
    if (r->reference) {
        s.temporal_taps = 0;
        s.spatial_taps = 0;
        s.clamp_jacobian = 0;
        s.clamp_outliers = 0;
        s.boiling_filter = 0;
        s.di_spatial_taps = 0;
        s.denoiser = 0;
    }
    return s;
}

} // namespace

extern "C" {

void rgi_settings_default(rgi_settings* s)
{
    memset(s, 0, sizeof(*s));
    s->initial_candidates = 2;
    s->temporal_taps = 2;
    s->spatial_taps = 4;
    s->trace_steps = 32;
    s->di_candidates = 8;
    s->di_spatial_taps = 3;
    s->ray_mode = RGI_RAY_HYBRID;
    s->world_space_reuse = 1;
    s->multiple_bounces = 1;
    s->clamp_jacobian = 1;
    s->clamp_outliers = 1;
    s->boiling_filter = 1;
    s->restir_di = 1;
    s->emissive_lighting = 1;
    s->sky = 0;
    s->shadows = 0;
    s->grid_mode = RGI_GRID_ATOMIC;
    s->grid_table_size = 65536u;
    const float zenith[3] = {0.30f, 0.50f, 0.95f}, horizon[3] = {0.85f, 0.88f, 0.92f}, ground[3] = {0.32f, 0.29f, 0.26f};
    memcpy(s->sky_zenith, zenith, sizeof(zenith));
    memcpy(s->sky_horizon, horizon, sizeof(horizon));
    memcpy(s->sky_ground, ground, sizeof(ground));
    s->sky_intensity = 1.0f;
    s->light_position[1] = 1.7f;
    s->shadow_face_size = 1;
    s->denoiser = 1;
    s->atrous_iterations = 5;
    s->composite = 1;
    s->specular_gi = 1;
    s->gi_divisor = 2;
}

// The GI size of a window (renderer.cpp:519-520, 1329-1331).
static void gi_size(const rgi_settings* s, int32_t width, int32_t height, int32_t& gi_width, int32_t& gi_height)
{
    int32_t d = s->gi_divisor > 0 ? s->gi_divisor : 1;
    gi_width = width / d > 1 ? width / d : 1;
    gi_height = height / d > 1 ? height / d : 1;
}

uint32_t rgi_settings_sizeof(void) { return (uint32_t)sizeof(rgi_settings); }

int rgi_frame_pack(const rgi_settings* s, const rgi_scene* scene, float radius, uint32_t frame_number,
                   const rgi_view* view, const rgi_view* prev, int32_t width, int32_t height, int history_valid,
                   int di_history_valid, rgi_frame* out)
{
    if (!s || !view || !prev || !out) return fail(RGI_E_INVALID, "frame_pack: NULL argument");
    if (width <= 0 || height <= 0) return fail(RGI_E_INVALID, "frame_pack: size %d x %d", width, height);
    if (s->gi_divisor != 1 && s->gi_divisor != 2 && s->gi_divisor != 4)
        return fail(RGI_E_INVALID, "frame_pack: gi_divisor %d is not 1, 2 or 4", s->gi_divisor);
    rgi_frame f;
    memset(&f, 0, sizeof(f));
    memcpy(f.inv_view_proj, view->inv_view_proj, sizeof(f.inv_view_proj));
    memcpy(f.view_proj, view->view_proj, sizeof(f.view_proj));
    memcpy(f.prev_view_proj, prev->view_proj, sizeof(f.prev_view_proj));
    for (int i = 0; i < 3; ++i) {
        f.camera_position[i] = view->eye[i];
        f.prev_camera_position[i] = prev->eye[i];
        f.light_position[i] = s->light_position[i];
    }
    f.camera_position[3] = f.prev_camera_position[3] = f.light_position[3] = 1.0f;
    // sceneScale and the light's scale with the square of it (renderer.cpp:497-516).
    const float scale = radius / 1.75f;
    for (int i = 0; i < 3; ++i) f.light_color[i] = s->light_color[i] * (scale * scale);
    f.light_color[3] = 1.0f;
    f.frame_index = frame_number & 0xffffffu;
    f.history_valid = history_valid ? 1 : 0;
    f.origin_bottom_left = s->origin_bottom_left;
    f.multiple_bounces = s->multiple_bounces ? 1 : 0;
    f.homogeneous_depth = s->homogeneous_depth;
    gi_size(s, width, height, f.gi_width, f.gi_height);
    f.width = width;
    f.height = height;
    f.scene_scale = scale;
    f.boiling_filter = s->boiling_filter ? 4.0f : 0.0f;  // kBoilingFilter, settings.h:31
    f.jacobian_clamp = s->clamp_jacobian ? 10.0f : 0.0f; // kJacobianClamp, settings.h:29
    f.outlier_clamp = s->clamp_outliers ? 8.0f : 0.0f;   // kOutlierClamp, settings.h:30
    // u_lightingInfo (renderer.cpp:694-702).
    f.lighting_surface = s->restir_di ? 1.0f : (s->emissive_lighting ? 1.0f : 0.0f);
    f.lighting_emission = s->restir_di ? 0.0f : (s->emissive_lighting ? 1.0f : 0.0f);
    f.initial_candidates = s->initial_candidates;
    f.temporal_taps = s->temporal_taps;
    f.spatial_taps = s->spatial_taps;
    f.trace_steps = s->trace_steps;
    // u_gridInfo (renderer.cpp:714-720): kGridMinCellSize 0.02, kGridCellPixels 6, fovY 48.
    const float pi = 3.14159265358979323846f;
    f.grid_min_cell = 0.02f * scale;
    f.grid_pixel_size = 2.0f * tanf(0.5f * (48.0f * (pi / 180.0f))) / (float)f.gi_height;
    f.grid_cell_pixels = 6.0f;
    f.world_space_reuse = s->world_space_reuse ? 1 : 0;
    f.grid_table_size = s->grid_table_size;
    f.ray_mode = s->ray_mode;
    for (int i = 0; i < 3; ++i) {
        f.sky_zenith[i] = s->sky_zenith[i] * s->sky_intensity;
        f.sky_horizon[i] = s->sky_horizon[i] * s->sky_intensity;
        f.sky_ground[i] = s->sky_ground[i] * s->sky_intensity;
    }
    f.sky_zenith[3] = s->sky ? 1.0f : 0.0f;
    f.shadows_enabled = s->shadows ? 1 : 0;
    if (s->shadows) {
        // The shadow uniforms of this light (renderer.cpp:527-564, 782-783).
        rgi_shadow sh;
        int e = rgi_shadow_setup(s->light_position, radius, shadow_face_size(s), s->homogeneous_depth,
                                 s->origin_bottom_left, &sh);
        if (e != RGI_OK) return fail(e, "frame_pack: shadow setup failed (face size index %d)", s->shadow_face_size);
        rgi_shadow_apply(&sh, 1, &f);
    }
    f.di_light_count = s->restir_di && scene ? (int32_t)scene->light_count : 0;
    f.di_candidates = s->di_candidates;
    f.di_spatial_taps = s->di_spatial_taps;
    f.di_history_valid = di_history_valid ? 1 : 0;
    *out = f;
    return RGI_OK;
}

int rgi_renderer_create(rgi_renderer** out, int32_t width, int32_t height, const rgi_settings* settings,
                        const rgi_scene* scene, float radius)
{
    if (!out || !settings || !scene) return fail(RGI_E_INVALID, "renderer_create: NULL argument");
    if (width <= 0 || height <= 0 || (int64_t)width * height >= (1ll << 29))
        return fail(RGI_E_INVALID, "renderer_create: size %d x %d out of range", width, height);
    uint32_t T = settings->grid_table_size;
    if (T < 64u || (T & (T - 1u))) return fail(RGI_E_INVALID, "grid_table_size must be a power of two >= 64");
    if (settings->atrous_iterations < 0 || settings->atrous_iterations > 5)
        return fail(RGI_E_INVALID, "atrous_iterations %d is not in 0..5", settings->atrous_iterations);
    if (settings->shadows && shadow_face_size(settings) == 0)
        return fail(RGI_E_INVALID, "shadow_face_size index %d is not in 0..3", settings->shadow_face_size);
    if (settings->gi_divisor != 1 && settings->gi_divisor != 2 && settings->gi_divisor != 4)
        return fail(RGI_E_INVALID, "gi_divisor %d is not 1, 2 or 4", settings->gi_divisor);
    rgi_renderer* r = new (std::nothrow) rgi_renderer();
    if (!r) return fail(RGI_E_NOMEM, "renderer allocation failed");
    r->width = width;
    r->height = height;
    gi_size(settings, width, height, r->gi_width, r->gi_height);
    r->settings = *settings;
    r->scene = *scene;
    r->radius = radius;
    int e = rgi_create(&r->ctx);
    if (e != RGI_OK) {
        delete r;
        return e;
    }
    size_t n = (size_t)width * height;                    // window-size planes
    size_t ng = (size_t)r->gi_width * (size_t)r->gi_height; // GI-size planes
    bool ok = true;
    for (int k = 0; k < 2 && ok; ++k) {
        for (int j = 0; j < 3; ++j) ok = ok && alloc(r->gi[k][j], 4 * ng);
        ok = ok && alloc(r->color[k], 4 * ng) && alloc(r->moments[k], 4 * ng) && alloc(r->grid[k], (size_t)34 * T);
        ok = ok && alloc(r->di_sample[k], 4 * n) && alloc(r->di_receiver[k], 4 * n) && alloc(r->di_normal[k], 4 * n);
    }
    ok = ok && alloc(r->diffuse, 4 * n) && alloc(r->specular, 4 * n);
    for (int k = 0; k < 2 && ok; ++k) ok = alloc(r->filtered[k], 4 * ng) && alloc(r->filter[k], 4 * ng);
    if (settings->composite) ok = ok && alloc(r->hdr, 4 * n);
    if (settings->restir_di) ok = ok && alloc(r->black, 4 * n);
    if (settings->specular_gi) ok = ok && alloc(r->specular_gi[0], 4 * ng) && alloc(r->specular_gi[1], 4 * ng);
    if (settings->shadows) {
        size_t F = (size_t)shadow_face_size(settings);
        ok = ok && alloc(r->atlas, 6 * F * F);
    }
    ok = ok && hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking) == hipSuccess &&
         hipStreamCreateWithFlags(&r->grid_side, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&r->side_fork, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&r->side_trace, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&r->side_join, hipEventDisableTiming) == hipSuccess &&
         hipStreamCreateWithFlags(&r->nrc_stream, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&r->nrc_fork, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&r->nrc_done, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        free_all(r);
        rgi_destroy(r->ctx);
        delete r;
        return fail(RGI_E_NOMEM, "renderer buffers for %d x %d (grid %u) could not be allocated", width, height, T);
    }
    *out = r;
    return RGI_OK;
}

void rgi_renderer_destroy(rgi_renderer* r)
{
    if (!r) return;
    hipDeviceSynchronize();
    if (r->graph_exec) hipGraphExecDestroy(r->graph_exec);
    free_all(r);
    rgi_destroy(r->ctx);
    delete r;
}

int rgi_renderer_set_reference_mode(rgi_renderer* r, int on)
{
    if (!r) return fail(RGI_E_INVALID, "renderer is NULL");
    if ((bool)on != r->reference) {
        r->reference = on != 0;
        r->history_valid = false; // renderer.cpp:1140-1141
        r->di_valid = false;
    }
    return RGI_OK;
}

void rgi_renderer_invalidate_history(rgi_renderer* r)
{
    if (!r) return;
    r->history_valid = false;
    r->di_valid = false;
}

// What render_frame commits on the host (ping-pong parity, validity flags,
// frame count, cameras, the cache's step count): saved before a capture so
// that a capture, instantiation or launch that fails leaves the renderer as
// it was, its GPU buffers untouched by work that never ran.
struct HostState {
    int read, specular_read;
    bool history_valid, di_valid, have_frame, nrc_pending, atlas_valid;
    uint32_t frame_number, nrc_trained;
    rgi_view previous;
    rgi_frame last;
    const uint16_t* denoised;
};

static HostState save_state(const rgi_renderer* r)
{
    return HostState{r->read, r->specular_read, r->history_valid, r->di_valid, r->have_frame, r->nrc_pending,
                     r->atlas_valid, r->frame_number, r->nrc_trained, r->previous, r->last, r->denoised};
}

static void restore_state(rgi_renderer* r, const HostState& h)
{
    r->read = h.read;
    r->specular_read = h.specular_read;
    r->history_valid = h.history_valid;
    r->di_valid = h.di_valid;
    r->have_frame = h.have_frame;
    r->nrc_pending = h.nrc_pending;
    r->atlas_valid = h.atlas_valid;
    r->frame_number = h.frame_number;
    r->nrc_trained = h.nrc_trained;
    r->previous = h.previous;
    r->last = h.last;
    r->denoised = h.denoised;
}

// The caller's stream waits for a training step still running on the
// cache's stream.
static void join_nrc(rgi_renderer* r, hipStream_t stream)
{
    if (r->nrc_pending) hipStreamWaitEvent(stream, r->nrc_done, 0);
    r->nrc_pending = false;
}

// Everything one frame of a renderer settles on the host before its first
// launch: the packed frame, the ping-pong parity, and the planes that each
// pass reads and writes.
struct FrameJob {
    rgi_settings s;
    rgi_frame f;  // DI, grid, GI tracing, denoiser, composite
    rgi_frame fg; // GI resampling, spatial and specular: f with the cache's lighting
    int read, write;
    bool di_active, split, nrc_on, nrc_query, shadow_pass;
    const float* shadow_map;
    // The GI passes' surface light (renderer.cpp:690-695): the DI diffuse
    // light with ReSTIR DI (black when it has no lights), else the emission;
    // the cache's prediction once it has trained.
    const uint16_t* surface_light;
    // With the denoiser, the GI history colour is its first filtered
    // iteration (renderer.cpp:774-776).
    const uint16_t* history_color;
    const uint16_t* query_emission;   // the emission the cache's query adds, or NULL
    const uint16_t* resolved_diffuse; // the composite's diffuse light from the cache, or NULL
    rgi_di_reservoirs prev_di, next_di;
    rgi_gi_reservoirs gi_out;
    rgi_gi_inputs tin, sin;
};

// Checks the frame's arguments and plans it (no GPU work). The only state
// it changes is the reference mode's reset of DI history.
static int plan_frame(rgi_renderer* r, const rgi_view* view, const rgi_gbuffer* g, FrameJob& j)
{
    if (!r || !view || !g) return fail(RGI_E_INVALID, "renderer_render: NULL argument");
    if (g->width != r->width || g->height != r->height)
        return fail(RGI_E_INVALID, "G-buffer %d x %d does not match the renderer's %d x %d", g->width, g->height,
                    r->width, r->height);
    if (!g->depth || !g->normal || !g->albedo || !g->shading_normal || !g->emission)
        return fail(RGI_E_INVALID, "renderer_render: every G-buffer plane is required");
    j.s = effective(r);
    const rgi_settings& s = j.s;
    if (r->reference) r->di_valid = false; // no DI temporal reuse either (renderer.cpp:451)
    const rgi_view& prev = r->have_frame ? r->previous : *view;
    j.di_active = s.restir_di && r->scene.light_count > 0 && r->scene.node_count > 0;
    int e = rgi_frame_pack(&s, &r->scene, r->radius, r->frame_number, view, &prev, r->width, r->height,
                           r->history_valid, r->di_valid, &j.f);
    if (e != RGI_OK) return e;
    if (!j.di_active) j.f.di_light_count = 0;
    j.read = r->read;
    j.write = 1 - r->read;
    const int read = j.read, write = j.write;
    j.shadow_map = s.shadows ? r->atlas : nullptr;
    j.shadow_pass = s.shadows && !r->atlas_valid;
    j.split = gi_temporal_split(r->ctx, j.f);
    j.nrc_on = r->nrc && !r->reference;
    j.nrc_query = j.nrc_on && r->nrc_trained >= r->nrc_settings.warmup;
    j.fg = j.f;
    j.surface_light = s.restir_di ? r->black : g->emission;
    if (j.di_active) j.surface_light = r->diffuse;
    j.query_emission = nullptr;
    j.resolved_diffuse = nullptr;
    if (j.nrc_query) {
        j.query_emission = (!s.restir_di && s.emissive_lighting) ? g->emission : nullptr;
        j.surface_light = r->nrc_plane;
        j.fg.multiple_bounces = 0;
        j.fg.lighting_surface = 1.0f;
        // The composite's diffuse light: the cache's prediction itself,
        // which covers DI and GI only when the training target holds both.
        if (r->nrc_settings.resolve && s.multiple_bounces)
            j.resolved_diffuse = j.query_emission ? r->nrc_resolved : r->nrc_plane;
    }
    j.history_color = s.denoiser ? r->filtered[read] : r->color[read];
    j.prev_di = rgi_di_reservoirs{r->di_sample[read], r->di_receiver[read], r->di_normal[read]};
    j.next_di = rgi_di_reservoirs{r->di_sample[write], r->di_receiver[write], r->di_normal[write]};
    j.gi_out = rgi_gi_reservoirs{r->gi[write][0], r->gi[write][1], r->gi[write][2]};
    j.tin = rgi_gi_inputs{r->gi[read][0], r->gi[read][1], r->gi[read][2], j.history_color, r->moments[read],
                          j.surface_light, j.shadow_map, r->grid[read]};
    j.sin = rgi_gi_inputs{nullptr, r->gi[read][1], r->gi[read][2], j.history_color, r->moments[read],
                          j.surface_light, j.shadow_map, r->grid[write]};
    return RGI_OK;
}

// The shadow pass (renderer.cpp:527-596), on the first frame with shadows.
static int frame_shadow(rgi_renderer* r, const FrameJob& j, hipStream_t stream)
{
    if (!j.shadow_pass) return RGI_OK;
    rgi_shadow sh;
    int e = rgi_shadow_setup(j.s.light_position, r->radius, shadow_face_size(&j.s), j.s.homogeneous_depth,
                             j.s.origin_bottom_left, &sh);
    if (e != RGI_OK) return fail(e, "shadow setup failed");
    if ((e = shadow_atlas(sh, r->scene, r->atlas, stream)) != RGI_OK) return e;
    r->atlas_valid = true;
    return RGI_OK;
}

// This frame's hash grid (renderer.cpp:735-750).
static int frame_grid(rgi_renderer* r, const FrameJob& j, const rgi_gbuffer& g, hipStream_t stream)
{
    int e = grid_clear(r->grid[j.write], j.f.grid_table_size, stream);
    if (e != RGI_OK) return e;
    if (j.s.grid_mode == RGI_GRID_DETERMINISTIC) return grid_insert_deterministic(r->ctx, j.f, g, r->grid[j.write], stream);
    return grid_insert_atomic(j.f, g, r->grid[j.write], stream);
}

// The cache's query at every pixel once it has trained, on a stream that
// follows the cache's last training step.
static int nrc_query_planes(rgi_renderer* r, const FrameJob& j, const rgi_gbuffer& g, hipStream_t stream)
{
    if (!j.nrc_query) return RGI_OK;
    int e = nrc_infer_gbuffer(r->nrc, j.f, g, j.query_emission, r->nrc_settings.precision, r->nrc_plane, stream);
    if (e != RGI_OK || !j.query_emission || !j.resolved_diffuse) return e;
    return nrc_infer_gbuffer(r->nrc, j.f, g, nullptr, r->nrc_settings.precision, r->nrc_resolved, stream);
}

static int frame_nrc_query(rgi_renderer* r, const FrameJob& j, const rgi_gbuffer& g, hipStream_t stream)
{
    if (!j.nrc_query) return RGI_OK;
    join_nrc(r, stream);
    return nrc_query_planes(r, j, g, stream);
}

// After GI spatial: indirect specular (its plane for the composite goes to
// *specular_out, NULL without it) and the denoiser.
static int frame_specular(rgi_renderer* r, const FrameJob& j, const rgi_gbuffer& g, hipStream_t stream,
                          const uint16_t** specular_out)
{
    const rgi_settings& s = j.s;
    const int read = j.read, write = j.write;
    int e;
    // Indirect specular (renderer.cpp:836-855): last frame's composite is
    // still in the HDR target.
    const uint16_t* specular = nullptr;
    *specular_out = nullptr;
    if (s.specular_gi && r->specular_gi[0]) {
        const int sw = 1 - r->specular_read;
        rgi_gi_inputs spin{nullptr, r->gi[read][1], r->gi[read][2], j.history_color, r->moments[read],
                           j.surface_light, j.shadow_map, r->grid[write]};
        if ((e = specular_gi(j.fg, r->scene, g, spin, r->hdr, r->specular_gi[r->specular_read], r->specular_gi[sw],
                             r->ctx->tuning, stream)) != RGI_OK)
            return e;
        r->specular_read = sw;
        specular = r->specular_gi[sw];
        *specular_out = specular;
    }
    return RGI_OK;
}

// The indirect specular plane frame_specular will write for this frame
// (NULL without the pass).
static const uint16_t* planned_specular(const rgi_renderer* r, const FrameJob& j)
{
    return j.s.specular_gi && r->specular_gi[0] ? r->specular_gi[1 - r->specular_read] : nullptr;
}

// The denoiser's planes (renderer.cpp:857-907): the variance into filter
// target 0; iteration 0 writes the filtered history, iteration 1 target 1,
// then 1 -> 0, 0 -> 1, 1 -> 0. out[0] is the variance's output and out[it + 1]
// iteration it's; the denoised plane is out[iterations].
static void denoiser_planes(rgi_renderer* r, const FrameJob& j, uint16_t* out[6])
{
    out[0] = r->filter[0];
    for (int it = 0; it < 5; ++it) out[it + 1] = it == 0 ? r->filtered[j.write] : r->filter[it & 1];
}

static int frame_post(rgi_renderer* r, const FrameJob& j, const rgi_gbuffer& g, hipStream_t stream,
                      const uint16_t** specular_out)
{
    const rgi_settings& s = j.s;
    const rgi_frame& f = j.f;
    const int write = j.write;
    int e = frame_specular(r, j, g, stream, specular_out);
    if (e != RGI_OK) return e;
    r->denoised = r->color[write];
    if (s.denoiser) {
        const float* receiver = r->gi[write][1];
        const float* normal = r->gi[write][2];
        if ((e = svgf_variance(f, r->color[write], r->moments[write], receiver, normal, r->filter[0], stream)) !=
            RGI_OK)
            return e;
        for (int it = 0; it < s.atrous_iterations; ++it) {
            const uint16_t* input = it == 0 ? r->filter[0] : it == 1 ? r->filtered[write] : r->filter[(it - 1) & 1];
            uint16_t* output = it == 0 ? r->filtered[write] : r->filter[it & 1];
            if ((e = svgf_atrous(f, input, receiver, normal, 1 << it, output, stream)) != RGI_OK) return e;
            r->denoised = output;
        }
    }
    return RGI_OK;
}

// The cache's training records of this frame's surfaces: target = DI
// diffuse + albedo * next frame's GI history colour; after `stream` has
// joined the cache's last training step, which reads the same records.
static int frame_gather(rgi_renderer* r, const FrameJob& j, const rgi_gbuffer& g, hipStream_t stream)
{
    const rgi_settings& s = j.s;
    const rgi_frame& f = j.f;
    const int write = j.write;
    join_nrc(r, stream);
    const uint16_t* next_history = s.denoiser ? r->filtered[write] : r->color[write];
    // Records at the frame's ReSTIR GI path vertices (the temporal
    // reservoirs' sample points), at its surface pixels, or alternately.
    const int rec = r->nrc_settings.records;
    const bool vertices = rec == RGI_NRC_RECORDS_VERTICES || rec == RGI_NRC_RECORDS_MIXED;
    return nrc_gather(r->nrc, f, g, j.di_active ? r->diffuse : nullptr, s.multiple_bounces ? next_history : nullptr,
                      vertices ? r->gi[write][0] : nullptr, vertices ? r->gi[write][2] : nullptr,
                      rec == RGI_NRC_RECORDS_MIXED ? 2u : 1u, r->nrc_settings.batch, f.frame_index ^ 0x4e5243u,
                      r->nrc_records, r->nrc_targets, stream);
}

// The composite and the end of the frame; `cross_frame` leaves a training
// step of the cache running past it.
// The composite's inputs (renderer.cpp:967-990): the final image's weights.
static rgi_composite_inputs composite_inputs(const rgi_renderer* r, const FrameJob& j, const uint16_t* denoised,
                                             const uint16_t* specular)
{
    rgi_composite_inputs ci{};
    ci.gi = denoised;
    ci.di_diffuse = j.di_active ? r->diffuse : nullptr;
    ci.di_specular = j.di_active ? r->specular : nullptr;
    ci.specular = specular;
    ci.shadow_map = j.shadow_map;
    ci.resolved_diffuse = j.resolved_diffuse;
    for (int k = 0; k < 4; ++k) ci.options[k] = 1.0f;
    ci.specular_weight = specular ? 1.0f : 0.0f;
    return ci;
}

// End of frame (renderer.cpp:1119-1124).
static void frame_commit(rgi_renderer* r, const FrameJob& j, const rgi_view* view, hipStream_t stream, bool cross_frame)
{
    r->di_valid = j.di_active;
    r->previous = *view;
    r->have_frame = true;
    r->read = j.write;
    r->history_valid = true;
    ++r->frame_number;
    r->last = j.f;
    if (!cross_frame) join_nrc(r, stream);
}

static int frame_finish(rgi_renderer* r, const FrameJob& j, const rgi_view* view, const rgi_gbuffer& g,
                        const uint16_t* specular, hipStream_t stream, bool cross_frame)
{
    if (j.s.composite && r->hdr) {
        int e = composite(j.f, g, composite_inputs(r, j, r->denoised, specular), r->hdr, stream);
        if (e != RGI_OK) return e;
    }

    frame_commit(r, j, view, stream, cross_frame);
    return RGI_OK;
}

// Everything after GI spatial: indirect specular, the denoiser, the cache's
// training step, the composite, and the end of the frame. A non-null
// `train_on` runs the training step on that stream (forked from `stream`)
// beside the composite (and, with `cross_frame`, the next frame's DI passes
// and grid); `cross_frame` leaves it running past the end of the frame.
static int frame_tail(rgi_renderer* r, const FrameJob& j, const rgi_view* view, const rgi_gbuffer& g,
                      hipStream_t stream, hipStream_t train_on, bool cross_frame)
{
    const uint16_t* specular = nullptr;
    int e = frame_post(r, j, g, stream, &specular);
    if (e != RGI_OK) return e;
    if (j.nrc_on) {
        if ((e = frame_gather(r, j, g, stream)) != RGI_OK) return e;
        hipStream_t train_stream = stream;
        if (train_on) {
            hipEventRecord(r->nrc_fork, stream);
            hipStreamWaitEvent(train_on, r->nrc_fork, 0);
            train_stream = train_on;
        }
        e = rgi_nrc_train(r->nrc, r->nrc_records, r->nrc_targets, r->nrc_settings.batch, r->nrc_loss, train_stream);
        if (train_on) {
            hipEventRecord(r->nrc_done, train_on);
            r->nrc_pending = true;
        }
        if (e != RGI_OK) {
            join_nrc(r, stream);
            return e;
        }
        ++r->nrc_trained;
    }
    return frame_finish(r, j, view, g, specular, stream, cross_frame);
}

// One frame. `fork_grid` traces GI temporal's candidates and builds the hash
// grid on the renderer's side streams and trains the cache on the cache's
// stream; `cross_frame` leaves that training running past the end of the
// frame, joined where the next frame first touches the cache (eager frames
// only: a captured frame joins every branch before its capture ends).
static int render_frame(rgi_renderer* r, const rgi_view* view, const rgi_gbuffer* g, void* stream_ptr, bool fork_grid,
                        bool cross_frame)
{
    FrameJob j;
    int e = plan_frame(r, view, g, j);
    if (e != RGI_OK) return e;
    hipStream_t stream = as_stream(stream_ptr);
    if ((e = frame_shadow(r, j, stream)) != RGI_OK) return e;
    // GI temporal's tracing half reads only the G-buffer and the scene (and
    // writes the sky tiles' empty reservoirs into this frame's GI planes,
    // which nothing reads before GI temporal), so with `fork_grid` it runs
    // on the side stream beside the DI passes, joined before GI temporal's
    // resampling. This frame's hash grid is built on a third stream from the
    // same fork (beside both, not after the tracing, where it would slow the
    // resampling): only GI spatial and the specular pass read the grid (GI
    // temporal reads last frame's), so it is joined before GI spatial. The
    // buffers both write were last read by last frame's passes, which
    // precede the fork on `stream`.
    hipStream_t grid_stream = fork_grid ? r->grid_side : stream;
    if (fork_grid) {
        hipEventRecord(r->side_fork, stream);
        hipStreamWaitEvent(r->side, r->side_fork, 0);
        hipStreamWaitEvent(r->grid_side, r->side_fork, 0);
        if (j.split) {
            e = gi_temporal_trace(r->ctx, j.f, r->scene, *g, j.gi_out, r->side);
            hipEventRecord(r->side_trace, r->side);
        }
    }
    if (e == RGI_OK) e = frame_grid(r, j, *g, grid_stream);
    if (fork_grid) hipEventRecord(r->side_join, r->grid_side);
    // The DI passes, the cache's query and GI temporal, beside the grid.
    auto front = [&]() -> int {
        int fe;
        if (j.di_active) {
            if ((fe = di_temporal(r->ctx, j.f, r->scene, *g, j.prev_di, j.next_di, r->ctx->tuning, stream, true)) !=
                RGI_OK)
                return fe;
            if ((fe = di_spatial(j.f, r->scene, *g, j.next_di, r->diffuse, r->specular, r->ctx->tuning, stream)) !=
                RGI_OK)
                return fe;
        }
        if ((fe = frame_nrc_query(r, j, *g, stream)) != RGI_OK) return fe;
        if (fork_grid && j.split) {
            hipStreamWaitEvent(stream, r->side_trace, 0);
            return gi_temporal_resample(r->ctx, j.fg, r->scene, *g, j.tin, j.gi_out, r->ctx->tuning, stream);
        }
        return gi_temporal(r->ctx, j.fg, r->scene, *g, j.tin, j.gi_out, r->ctx->tuning, stream);
    };
    if (e == RGI_OK) e = front();
    if (fork_grid) {
        // Every branch of the side stream is joined, also when a pass failed.
        if (j.split) hipStreamWaitEvent(stream, r->side_trace, 0);
        hipStreamWaitEvent(stream, r->side_join, 0);
    }
    if (e != RGI_OK) return e;
    if ((e = gi_spatial(j.fg, r->scene, *g, j.sin, j.gi_out, r->color[j.write], r->moments[j.write], r->ctx->tuning,
                        stream)) != RGI_OK)
        return e;
    return frame_tail(r, j, view, *g, stream, fork_grid ? r->nrc_stream : nullptr, cross_frame);
}

int rgi_renderer_render(rgi_renderer* r, const rgi_view* view, const rgi_gbuffer* g, void* stream_ptr)
{
    return render_frame(r, view, g, stream_ptr, true, true);
}

int rgi_renderer_render_graph(rgi_renderer* r, const rgi_view* view, const rgi_gbuffer* g, void* stream_ptr)
{
    if (!r) return fail(RGI_E_INVALID, "renderer is NULL");
    if (!stream_ptr) return fail(RGI_E_INVALID, "render_graph: capture needs a stream other than the null stream");
    // The first frame runs eagerly: scratch buffers and the shadow atlas
    // reach their final state outside capture.
    if (!r->graph_warm) {
        int e = render_frame(r, view, g, stream_ptr, true, false);
        if (e == RGI_OK) r->graph_warm = true;
        return e;
    }
    hipStream_t stream = as_stream(stream_ptr);
    join_nrc(r, stream); // an eager frame's training, joined outside the capture
    const HostState before = save_state(r);
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(RGI_E_DEVICE, "render_graph: stream capture could not begin");
    int e = render_frame(r, view, g, stream_ptr, true, false);
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(stream, &graph);
    if (e != RGI_OK || ce != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        restore_state(r, before);
        return e != RGI_OK ? e : fail(RGI_E_DEVICE, "render_graph: capture failed: %s", hipGetErrorString(ce));
    }
    // Same launches as last frame with new arguments: update in place;
    // a changed topology (another pass set) instantiates afresh.
    if (r->graph_exec) {
        hipGraphNode_t error_node = nullptr;
        hipGraphExecUpdateResult result;
        if (hipGraphExecUpdate(r->graph_exec, graph, &error_node, &result) != hipSuccess) {
            hipGetLastError();
            hipGraphExecDestroy(r->graph_exec);
            r->graph_exec = nullptr;
        }
    }
    if (!r->graph_exec && hipGraphInstantiate(&r->graph_exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        r->graph_exec = nullptr;
        hipGraphDestroy(graph);
        restore_state(r, before);
        return fail(RGI_E_DEVICE, "render_graph: instantiation failed");
    }
    hipGraphDestroy(graph);
    if (hipGraphLaunch(r->graph_exec, stream) != hipSuccess) {
        restore_state(r, before);
        return fail(RGI_E_DEVICE, "render_graph: launch failed");
    }
    return RGI_OK;
}

int rgi_renderer_set_option(rgi_renderer* r, int option, int value)
{
    if (!r) return fail(RGI_E_INVALID, "renderer is NULL");
    return rgi_context_set_option(r->ctx, option, value);
}

int rgi_renderer_autotune(rgi_renderer* r, const rgi_view* views, const rgi_gbuffer* gbuffers, uint32_t count,
                          uint32_t frames, float* out_ms, void* stream_ptr)
{
    if (!r || !views || !gbuffers || count == 0 || frames == 0)
        return fail(RGI_E_INVALID, "autotune: views, gbuffers, count and frames are required");
    hipStream_t stream = as_stream(stream_ptr);
    hipEvent_t start, stop;
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess)
        return fail(RGI_E_DEVICE, "autotune: event creation failed");
    auto apply = [&](int config) {
        rgi::Tuning& t = r->ctx->tuning;
        t.gi_temporal_waves = (config & RGI_TUNE_TEMPORAL_WAVES4) ? 4 : 0;
        t.gi_spatial_waves = (config & RGI_TUNE_SPATIAL_WAVES5) ? 5 : 0;
        t.tile_order = (config & RGI_TUNE_TILE_ORDER) ? 2 : 0;
        t.tile_shape = (config & RGI_TUNE_TILE_SHAPE) ? 1 : 0;
    };
    int chosen = 0, e = RGI_OK;
    uint32_t k = 0;
    auto time_config = [&](int config, float& ms) {
        apply(config);
        // One untimed frame so each variant's code object is loaded and warm.
        e = rgi_renderer_render(r, &views[k % count], &gbuffers[k % count], stream_ptr);
        ++k;
        hipEventRecord(start, stream);
        for (uint32_t i = 0; i < frames && e == RGI_OK; ++i, ++k)
            e = rgi_renderer_render(r, &views[k % count], &gbuffers[k % count], stream_ptr);
        hipEventRecord(stop, stream);
        hipEventSynchronize(stop);
        ms = 0.0f;
        hipEventElapsedTime(&ms, start, stop);
        ms /= (float)frames;
    };
    // Knob t: the bits it owns and the values it tries.
    const int knob_mask[4] = {RGI_TUNE_TILE_ORDER, RGI_TUNE_TEMPORAL_WAVES4, RGI_TUNE_SPATIAL_WAVES5,
                              RGI_TUNE_TILE_SHAPE};
    const int knob_values[4][2] = {{0, RGI_TUNE_TILE_ORDER}, {0, RGI_TUNE_TEMPORAL_WAVES4}, {0, RGI_TUNE_SPATIAL_WAVES5},
                                   {0, RGI_TUNE_TILE_SHAPE}};
    int trial = 0;
    for (int knob = 0; knob < 4 && e == RGI_OK; ++knob) {
        float best = 3.4e38f;
        int best_config = chosen;
        for (int v = 0; v < 2 && e == RGI_OK; ++v, ++trial) {
            int config = (chosen & ~knob_mask[knob]) | knob_values[knob][v];
            float ms;
            time_config(config, ms);
            if (out_ms) out_ms[trial] = ms;
            if (ms < best) {
                best = ms;
                best_config = config;
            }
        }
        chosen = best_config;
    }
    hipEventDestroy(start);
    hipEventDestroy(stop);
    if (e != RGI_OK) return e;
    apply(chosen);
    return chosen;
}

int rgi_renderer_outputs_get(const rgi_renderer* r, rgi_renderer_outputs* out)
{
    if (!r || !out) return fail(RGI_E_INVALID, "renderer_outputs_get: NULL argument");
    int k = r->read;
    out->gi_sample = r->gi[k][0];
    out->gi_receiver = r->gi[k][1];
    out->gi_normal = r->gi[k][2];
    out->color = r->color[k];
    out->moments = r->moments[k];
    out->grid = r->grid[k];
    out->di_diffuse = r->diffuse;
    out->di_specular = r->specular;
    out->filtered_history = r->filtered[k];
    out->denoised = r->denoised ? r->denoised : r->color[k];
    out->hdr = r->hdr;
    out->specular_gi = r->specular_gi[0] ? r->specular_gi[r->specular_read] : nullptr;
    return RGI_OK;
}

void rgi_nrc_settings_default(rgi_nrc_settings* s)
{
    if (!s) return;
    memset(s, 0, sizeof(*s));
    rgi_nrc_config_default(&s->cache);
    s->batch = 16384;
    s->warmup = 4;
    s->precision = RGI_NRC_BF16;
    s->records = RGI_NRC_RECORDS_VERTICES;
}

int rgi_renderer_set_nrc(rgi_renderer* r, const rgi_nrc_settings* settings)
{
    if (!r) return fail(RGI_E_INVALID, "renderer is NULL");
    if (settings && (settings->batch == 0 || settings->batch > (1u << 24)))
        return fail(RGI_E_INVALID, "set_nrc: batch %u is not in 1..2^24", settings ? settings->batch : 0u);
    if (settings && (settings->precision < RGI_NRC_BF16 || settings->precision > RGI_NRC_FP8))
        return fail(RGI_E_INVALID, "set_nrc: unknown precision %d", settings->precision);
    if (settings && settings->resolve != 0 && settings->resolve != 1)
        return fail(RGI_E_INVALID, "set_nrc: resolve takes 0 or 1, not %d", settings->resolve);
    if (settings && (settings->records < RGI_NRC_RECORDS_PIXELS || settings->records > RGI_NRC_RECORDS_MIXED))
        return fail(RGI_E_INVALID, "set_nrc: unknown record source %d", settings->records);
    (void)hipDeviceSynchronize();
    release_nrc(r);
    if (!settings) return RGI_OK;
    int e = rgi_nrc_create(&r->nrc, &settings->cache);
    if (e != RGI_OK) return e;
    r->nrc_settings = *settings;
    size_t n = (size_t)r->width * r->height;
    if (!alloc(r->nrc_plane, 4 * n) || (settings->resolve && !alloc(r->nrc_resolved, 4 * n)) ||
        !alloc(r->nrc_records, (size_t)12 * settings->batch) ||
        !alloc(r->nrc_targets, (size_t)4 * settings->batch) || !alloc(r->nrc_loss, 1)) {
        release_nrc(r);
        return fail(RGI_E_NOMEM, "set_nrc: buffers could not be allocated");
    }
    // A changed pass set: the next graph frame instantiates afresh.
    r->graph_warm = false;
    return RGI_OK;
}

int rgi_renderer_nrc_state(const rgi_renderer* r, rgi_nrc** cache, const uint16_t** radiance, const float** loss,
                           uint32_t* trained)
{
    if (!r) return fail(RGI_E_INVALID, "renderer is NULL");
    // A training step still running on the cache's stream ends first, so the
    // state returned is the last frame's.
    if (r->nrc_pending) hipEventSynchronize(r->nrc_done);
    if (cache) *cache = r->nrc;
    if (radiance) *radiance = r->nrc ? r->nrc_plane : nullptr;
    if (loss) *loss = r->nrc ? r->nrc_loss : nullptr;
    if (trained) *trained = r->nrc_trained;
    return RGI_OK;
}

int rgi_renderer_last_frame(const rgi_renderer* r, rgi_frame* out)
{
    if (!r || !out) return fail(RGI_E_INVALID, "renderer_last_frame: NULL argument");
    if (!r->have_frame) return fail(RGI_E_INVALID, "no frame rendered yet");
    *out = r->last;
    return RGI_OK;
}

// ------------------------------------------------------------------ batch

// Renderers of any sizes and scenes whose frames go out together. The
// default (pass-major) form launches each screen pass once for every view:
// the kernels read each view's frame, scene and planes from a device table
// of BatchView entries, uploaded once per frame from a pinned staging slot,
// each view taking its own slice of the workgroups (screen passes) or its
// own tiles and rays from its own counters (the listed and persistent
// passes). Passes with per-renderer state (the cache's query and training,
// specular, denoiser, composite) stay per view. The graph form captures each
// renderer's frame on a stream of its own.
namespace {
// The view table of a batch of n views: the views, each pass's slice
// starts, each cache's training step, the training kernels' slice starts,
// each view's two cache queries (the query plane, and the resolved plane
// when the query adds emission) and their slice starts, each cache's
// training-record gather and its passes' slice starts.
struct TableLayout {
    size_t first, steps, nrc_first, queries, query_first, gathers, gather_first, bytes;
};
TableLayout table_layout(size_t n)
{
    TableLayout t;
    t.first = align256(n * sizeof(BatchView));
    t.steps = align256(t.first + (size_t)kBatchPasses * (n + 1) * sizeof(uint32_t));
    t.nrc_first = t.steps + n * kNrcStepBytes;
    t.queries = align256(t.nrc_first + (size_t)kNrcPasses * (n + 1) * sizeof(uint32_t));
    t.query_first = t.queries + 2 * n * kNrcQueryBytes;
    t.gathers = align256(t.query_first + (2 * n + 1) * sizeof(uint32_t));
    t.gather_first = t.gathers + n * kNrcGatherBytes;
    t.bytes = t.gather_first + 2 * (n + 1) * sizeof(uint32_t);
    return t;
}
} // namespace

struct rgi_batch {
    std::vector<rgi_renderer*> renderers;
    int mode = RGI_BATCH_PASSES;
    // Pass-major form: the side stream that traces and builds the grids, its
    // fork and join events, and the view table (two pinned staging slots,
    // each reused once the copy out of it has run, and the device copy).
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr, traced = nullptr, gridded = nullptr, gathered = nullptr;
    // The caches' stream: each frame's training chain, then the next frame's
    // queries (which need it), beside both other streams.
    hipStream_t cache_side = nullptr;
    hipEvent_t queried = nullptr;
    char* staging[2] = {};
    hipEvent_t copied[2] = {};
    bool copy_pending[2] = {};
    int slot = 0;
    char* table = nullptr;
    size_t table_bytes = 0;
    // Graph form.
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> joined;
    hipEvent_t fork = nullptr;
    hipGraphExec_t exec = nullptr;
    bool warm = false;
};

int rgi_batch_create(rgi_batch** out, rgi_renderer* const* renderers, uint32_t count)
{
    if (!out || !renderers || count == 0) return fail(RGI_E_INVALID, "batch_create: renderers are required");
    if (count > kMaxBatchViews) return fail(RGI_E_INVALID, "batch_create: %u renderers (at most %u)", count, kMaxBatchViews);
    for (uint32_t i = 0; i < count; ++i) {
        if (!renderers[i]) return fail(RGI_E_INVALID, "batch_create: renderer %u is NULL", i);
        for (uint32_t k = 0; k < i; ++k)
            if (renderers[k] == renderers[i])
                return fail(RGI_E_INVALID, "batch_create: renderer %u appears twice (also at %u)", i, k);
    }
    rgi_batch* b = new (std::nothrow) rgi_batch();
    if (!b) return fail(RGI_E_NOMEM, "batch allocation failed");
    b->renderers.assign(renderers, renderers + count);
    b->table_bytes = table_layout(count).bytes;
    bool ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
              hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&b->cache_side, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->queried, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->side_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->traced, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->gridded, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->gathered, hipEventDisableTiming) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&b->table), b->table_bytes) == hipSuccess;
    for (int k = 0; k < 2 && ok; ++k)
        ok = hipHostMalloc(reinterpret_cast<void**>(&b->staging[k]), b->table_bytes, hipHostMallocDefault) ==
                 hipSuccess &&
             hipEventCreateWithFlags(&b->copied[k], hipEventDisableTiming) == hipSuccess;
    for (uint32_t i = 0; i < count && ok; ++i) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
        if (st) b->streams.push_back(st);
        ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
        if (ev) b->joined.push_back(ev);
    }
    if (!ok) {
        rgi_batch_destroy(b);
        return fail(RGI_E_DEVICE, "batch_create: streams, events or the view table could not be created");
    }
    *out = b;
    return RGI_OK;
}

void rgi_batch_destroy(rgi_batch* b)
{
    if (!b) return;
    hipDeviceSynchronize();
    if (b->exec) hipGraphExecDestroy(b->exec);
    for (hipStream_t st : b->streams) hipStreamDestroy(st);
    for (hipEvent_t ev : b->joined) hipEventDestroy(ev);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->side) hipStreamDestroy(b->side);
    if (b->cache_side) hipStreamDestroy(b->cache_side);
    if (b->queried) hipEventDestroy(b->queried);
    if (b->side_fork) hipEventDestroy(b->side_fork);
    if (b->traced) hipEventDestroy(b->traced);
    if (b->gridded) hipEventDestroy(b->gridded);
    if (b->gathered) hipEventDestroy(b->gathered);
    for (int k = 0; k < 2; ++k) {
        if (b->staging[k]) hipHostFree(b->staging[k]);
        if (b->copied[k]) hipEventDestroy(b->copied[k]);
    }
    if (b->table) hipFree(b->table);
    delete b;
}

int rgi_batch_set_mode(rgi_batch* b, int mode)
{
    if (!b) return fail(RGI_E_INVALID, "batch_set_mode: NULL batch");
    if (mode != RGI_BATCH_PASSES && mode != RGI_BATCH_GRAPH)
        return fail(RGI_E_INVALID, "batch_set_mode: mode %d is not RGI_BATCH_PASSES or RGI_BATCH_GRAPH", mode);
    b->mode = mode;
    return RGI_OK;
}

} // extern "C"

namespace {

// Workgroups of a slice, padded to whole rounds of 8 so that every view's
// slice starts on XCD 0 (its workgroup b runs on XCD b mod 8, as in a launch
// of its own).
uint32_t slice(uint32_t blocks) { return (blocks + 7u) & ~7u; }

// The graph form: each renderer's frame captured on a stream of its own.
int render_graph_branches(rgi_batch* b, const rgi_view* views, const rgi_gbuffer* gbuffers, hipStream_t stream)
{
    if (!stream) return fail(RGI_E_INVALID, "batch_render: capture needs a stream other than the null stream");
    size_t n = b->renderers.size();
    int e = RGI_OK;
    // A cache that has not trained yet allocates its gather and training
    // scratch on its first step, which must not happen inside a capture.
    bool warm = b->warm;
    for (size_t i = 0; i < n; ++i)
        if (b->renderers[i]->nrc && b->renderers[i]->nrc_trained == 0) warm = false;
    if (!warm) {
        // First frame eagerly, on the caller's stream (allocations happen here).
        for (size_t i = 0; i < n && e == RGI_OK; ++i)
            e = render_frame(b->renderers[i], &views[i], &gbuffers[i], stream, false, false);
        if (e == RGI_OK) b->warm = true;
        return e;
    }
    for (size_t i = 0; i < n; ++i) join_nrc(b->renderers[i], stream); // eager frames' training, outside the capture
    std::vector<HostState> before;
    for (size_t i = 0; i < n; ++i) before.push_back(save_state(b->renderers[i]));
    auto restore_all = [&]() {
        for (size_t i = 0; i < n; ++i) restore_state(b->renderers[i], before[i]);
    };
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(RGI_E_DEVICE, "batch_render: stream capture could not begin");
    hipEventRecord(b->fork, stream);
    for (size_t i = 0; i < n && e == RGI_OK; ++i) {
        hipStreamWaitEvent(b->streams[i], b->fork, 0);
        // Each renderer's frame is already a branch of its own; its grid
        // stays on that branch (no nested fork inside the batch's capture).
        e = render_frame(b->renderers[i], &views[i], &gbuffers[i], b->streams[i], false, false);
        hipEventRecord(b->joined[i], b->streams[i]);
        hipStreamWaitEvent(stream, b->joined[i], 0);
    }
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(stream, &graph);
    if (e != RGI_OK || ce != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        restore_all();
        return e != RGI_OK ? e : fail(RGI_E_DEVICE, "batch_render: capture failed: %s", hipGetErrorString(ce));
    }
    if (b->exec) {
        hipGraphNode_t error_node = nullptr;
        hipGraphExecUpdateResult result;
        if (hipGraphExecUpdate(b->exec, graph, &error_node, &result) != hipSuccess) {
            hipGetLastError();
            hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
    }
    if (!b->exec && hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        b->exec = nullptr;
        hipGraphDestroy(graph);
        restore_all();
        return fail(RGI_E_DEVICE, "batch_render: instantiation failed");
    }
    hipGraphDestroy(graph);
    if (hipGraphLaunch(b->exec, stream) != hipSuccess) {
        restore_all();
        return fail(RGI_E_DEVICE, "batch_render: launch failed");
    }
    return RGI_OK;
}

// The pass-major form. Every view is checked and planned before the first
// launch, so a view that fails its checks leaves every renderer as it was.
int render_passes(rgi_batch* b, const rgi_view* views, const rgi_gbuffer* gbuffers, hipStream_t stream)
{
    const uint32_t n = (uint32_t)b->renderers.size();
    std::vector<FrameJob> jobs(n);
    std::vector<HostState> before;
    for (uint32_t i = 0; i < n; ++i) before.push_back(save_state(b->renderers[i]));
    auto restore_all = [&]() {
        for (uint32_t i = 0; i < n; ++i) restore_state(b->renderers[i], before[i]);
    };
    int e = RGI_OK;
    for (uint32_t i = 0; i < n && e == RGI_OK; ++i) e = plan_frame(b->renderers[i], &views[i], &gbuffers[i], jobs[i]);
    if (e != RGI_OK) {
        restore_all();
        return e;
    }
    for (uint32_t i = 0; i < n && e == RGI_OK; ++i) e = frame_shadow(b->renderers[i], jobs[i], stream);
    if (e != RGI_OK) {
        restore_all();
        return e;
    }

    // The view table, in the staging slot whose last copy has run.
    const int slot = b->slot;
    if (b->copy_pending[slot] && hipEventSynchronize(b->copied[slot]) != hipSuccess)
        return fail(RGI_E_DEVICE, "batch_render: the view table's last upload did not finish");
    b->copy_pending[slot] = false;
    char* host = b->staging[slot];
    memset(host, 0, b->table_bytes);
    const TableLayout T = table_layout(n);
    BatchView* hv = reinterpret_cast<BatchView*>(host);
    uint32_t* hfirst = reinterpret_cast<uint32_t*>(host + T.first);
    BatchLaunch L{};
    L.count = n;
    L.views = reinterpret_cast<const BatchView*>(b->table);
    for (int p = 0; p < kBatchPasses; ++p) {
        L.first[p] = reinterpret_cast<const uint32_t*>(b->table + T.first) + p * (n + 1);
        hfirst[p * (n + 1)] = 0u;
    }
    const Tuning& tuning0 = b->renderers[0]->ctx->tuning;
    L.gi_temporal_waves = tuning0.gi_temporal_waves;
    L.gi_spatial_waves = tuning0.gi_spatial_waves;
    L.compute_units = context_compute_units(b->renderers[0]->ctx);
    bool any_di = false;
    uint32_t max_atrous = 0u;
    for (uint32_t i = 0; i < n && e == RGI_OK; ++i) {
        rgi_renderer* r = b->renderers[i];
        const FrameJob& j = jobs[i];
        const rgi_gbuffer& g = gbuffers[i];
        BatchView& v = hv[i];
        v.f = j.f;
        v.fg = j.fg;
        v.s = r->scene;
        v.g = g;
        v.texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
        v.gi_texel = F2{1.0f / (float)j.f.gi_width, 1.0f / (float)j.f.gi_height};
        const Tuning& t = r->ctx->tuning;
        v.di_tiles = tile_map(t.tile_order, t.tile_shape, g.width, g.height);
        const dim3 dg = tile_grid(v.di_tiles);
        v.di_gx = dg.x;
        v.di_gy = dg.y;
        v.di_prev = j.prev_di;
        v.di_next = j.next_di;
        v.di_diffuse = r->diffuse;
        v.di_specular = r->specular;
        if (j.di_active) {
            uint32_t* guide = nullptr;
            e = di_light_guide(r->ctx, j.f, r->scene, &guide, &v.guide_buckets, stream, true);
            v.guide = guide;
            any_di = true;
        }
        v.grid = r->grid[j.write];
        v.grid_gx = ((uint32_t)j.f.gi_width + 15u) / 16u;
        v.grid_gy = ((uint32_t)j.f.gi_height + 15u) / 16u;
        v.tin = j.tin;
        v.sin = j.sin;
        v.res = j.gi_out;
        v.color = r->color[j.write];
        v.moments = r->moments[j.write];
        v.gi_tiles = tile_map(t.tile_order, t.tile_shape, j.f.gi_width, j.f.gi_height);
        const dim3 gg = tile_grid(v.gi_tiles);
        v.gi_gx = gg.x;
        v.gi_gy = gg.y;
        v.same = g.width == j.f.gi_width && g.height == j.f.gi_height ? 1u : 0u;
        v.split = j.split ? 1u : 0u;
        if (j.split && e == RGI_OK) {
            e = gi_listed_buffers(r->ctx, j.f, v.list, v.cand);
            L.split = true;
            L.tiles += v.list.tiles;
            if (j.f.initial_candidates > 0) {
                L.rounds += (v.list.tiles + 3u) / 4u;
                if (j.f.ray_mode != RGI_RAY_SCREEN) L.trace = true;
            }
        }
        (v.same ? L.same : L.larger) = true;
        // The denoiser's and the composite's planes (as frame_post and
        // frame_finish use them).
        const bool denoise = j.s.denoiser != 0;
        v.atrous = denoise ? (uint32_t)j.s.atrous_iterations : 0u;
        denoiser_planes(r, j, v.den_out);
        if (v.atrous > max_atrous) max_atrous = v.atrous;
        const bool composited = j.s.composite && r->hdr;
        if (composited) {
            v.ci = composite_inputs(r, j, v.atrous > 0u ? v.den_out[v.atrous] : r->color[j.write],
                                    planned_specular(r, j));
            v.hdr = r->hdr;
        }
        // Indirect specular (as frame_specular runs it).
        const bool specular_on = planned_specular(r, j) != nullptr;
        if (specular_on) {
            v.spec_previous = r->hdr;
            v.spec_history = r->specular_gi[r->specular_read];
            v.spec_out = r->specular_gi[1 - r->specular_read];
        }
        const uint32_t blocks[kBatchPasses] = {
            j.di_active ? slice(dg.x * dg.y) : 0u,
            slice(grid_clear_blocks(j.f.grid_table_size)),
            j.s.grid_mode == RGI_GRID_DETERMINISTIC ? 0u : slice(v.grid_gx * v.grid_gy),
            j.split ? slice((v.list.tiles + kListTilesPerGroup - 1u) / kListTilesPerGroup) : 0u,
            slice(gg.x * gg.y),
            denoise ? slice(svgf_blocks(j.f)) : 0u,
            composited ? slice(((uint32_t)g.width + 15u) / 16u * (((uint32_t)g.height + 15u) / 16u)) : 0u,
            specular_on ? slice(gg.x * gg.y) : 0u};
        for (int p = 0; p < kBatchPasses; ++p) hfirst[p * (n + 1) + i + 1] = hfirst[p * (n + 1) + i] + blocks[p];
    }
    // The caches' training steps of this frame, planned now (growing their
    // scratch before the first launch) and launched after the gathers.
    uint32_t* nfirst = reinterpret_cast<uint32_t*>(host + T.nrc_first);
    uint32_t nrc_blocks[kNrcPasses] = {};
    bool any_train = false;
    for (uint32_t i = 0; i < n && e == RGI_OK; ++i) {
        rgi_renderer* r = b->renderers[i];
        uint32_t blocks[kNrcPasses] = {};
        if (jobs[i].nrc_on) {
            e = nrc_step_plan(r->nrc, r->nrc_records, r->nrc_targets, r->nrc_settings.batch, r->nrc_loss,
                              host + T.steps + i * kNrcStepBytes, blocks);
            any_train = true;
        }
        for (int p = 0; p < kNrcPasses; ++p) nfirst[p * (n + 1) + i + 1] = nfirst[p * (n + 1) + i] + blocks[p];
    }
    for (int p = 0; p < kNrcPasses; ++p) nrc_blocks[p] = nfirst[p * (n + 1) + n];
    // The caches' training-record gathers (as frame_gather runs them).
    uint32_t* gfirst = reinterpret_cast<uint32_t*>(host + T.gather_first);
    for (uint32_t i = 0; i < n && e == RGI_OK; ++i) {
        rgi_renderer* r = b->renderers[i];
        const FrameJob& j = jobs[i];
        uint32_t blocks[2] = {0u, 0u};
        if (j.nrc_on) {
            const uint16_t* next_history = j.s.denoiser ? r->filtered[j.write] : r->color[j.write];
            const int rec = r->nrc_settings.records;
            const bool vertices = rec == RGI_NRC_RECORDS_VERTICES || rec == RGI_NRC_RECORDS_MIXED;
            e = nrc_gather_plan(r->nrc, j.f, gbuffers[i], j.di_active ? r->diffuse : nullptr,
                                j.s.multiple_bounces ? next_history : nullptr, vertices ? r->gi[j.write][0] : nullptr,
                                vertices ? r->gi[j.write][2] : nullptr, rec == RGI_NRC_RECORDS_MIXED ? 2u : 1u,
                                r->nrc_settings.batch, j.f.frame_index ^ 0x4e5243u, r->nrc_records, r->nrc_targets,
                                host + T.gathers + i * kNrcGatherBytes, blocks);
        }
        gfirst[i + 1] = gfirst[i] + blocks[0];
        gfirst[(n + 1) + i + 1] = gfirst[(n + 1) + i] + blocks[1];
    }
    const uint32_t gather_blocks[2] = {gfirst[n], gfirst[(n + 1) + n]};
    // The caches' queries: the query plane of every cache that has trained,
    // and its resolved plane where the query adds emission.
    uint32_t* qfirst = reinterpret_cast<uint32_t*>(host + T.query_first);
    bool precisions[3] = {false, false, false};
    bool any_query = false;
    for (uint32_t q = 0; q < 2 * n && e == RGI_OK; ++q) {
        const uint32_t i = q / 2;
        rgi_renderer* r = b->renderers[i];
        const FrameJob& j = jobs[i];
        const bool resolved = (q & 1u) != 0u;
        uint32_t blocks = 0;
        if (j.nrc_query && (!resolved || (j.query_emission && j.resolved_diffuse))) {
            e = nrc_query_plan(r->nrc, j.f, gbuffers[i], resolved ? nullptr : j.query_emission,
                               r->nrc_settings.precision, resolved ? r->nrc_resolved : r->nrc_plane,
                               host + T.queries + q * kNrcQueryBytes, &blocks);
            precisions[r->nrc_settings.precision] = true;
            any_query = true;
        }
        qfirst[q + 1] = qfirst[q] + blocks;
    }
    if (e != RGI_OK) {
        restore_all();
        return e;
    }
    for (int p = 0; p < kBatchPasses; ++p) L.blocks[p] = hfirst[p * (n + 1) + n];
    if (hipMemcpyAsync(b->table, host, b->table_bytes, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipEventRecord(b->copied[slot], stream) != hipSuccess) {
        restore_all();
        return fail(RGI_E_DEVICE, "batch_render: the view table could not be uploaded");
    }
    b->copy_pending[slot] = true;
    b->slot = 1 - slot;

    // Beside the DI passes on `stream`: on the side stream GI temporal's
    // tracing half, then the grids, joined before GI temporal's resampling
    // and GI spatial; on the caches' stream the caches' queries, after their
    // last training steps (which the previous frame left on that stream, or
    // a renderer's own frame on the cache's stream), joined before the
    // resampling.
    hipEventRecord(b->side_fork, stream);
    hipStreamWaitEvent(b->side, b->side_fork, 0);
    int se = RGI_OK;
    if (any_query) {
        hipStreamWaitEvent(b->cache_side, b->side_fork, 0);
        for (uint32_t i = 0; i < n; ++i)
            if (jobs[i].nrc_query && b->renderers[i]->nrc_pending)
                hipStreamWaitEvent(b->cache_side, b->renderers[i]->nrc_done, 0);
        se = nrc_query_batch(b->table + T.queries, reinterpret_cast<const uint32_t*>(b->table + T.query_first), 2 * n,
                             qfirst[2 * n], precisions, b->cache_side);
        hipEventRecord(b->queried, b->cache_side);
    }
    if (se == RGI_OK) se = gi_trace_batch(L, b->side);
    hipEventRecord(b->traced, b->side);
    if (se == RGI_OK) se = grid_clear_batch(L.views, L.first[kPassGridClear], n, L.blocks[kPassGridClear], b->side);
    if (se == RGI_OK) se = grid_insert_batch(L.views, L.first[kPassGrid], n, L.blocks[kPassGrid], b->side);
    for (uint32_t i = 0; i < n && se == RGI_OK; ++i)
        if (jobs[i].s.grid_mode == RGI_GRID_DETERMINISTIC)
            se = grid_insert_deterministic(b->renderers[i]->ctx, jobs[i].f, gbuffers[i], b->renderers[i]->grid[jobs[i].write],
                                           b->side);
    hipEventRecord(b->gridded, b->side);

    if (any_di) e = di_temporal_batch(L.views, L.first[kPassDi], n, L.blocks[kPassDi], stream);
    if (any_di && e == RGI_OK) e = di_spatial_batch(L.views, L.first[kPassDi], n, L.blocks[kPassDi], stream);
    hipStreamWaitEvent(stream, b->traced, 0);
    if (any_query) hipStreamWaitEvent(stream, b->queried, 0);
    // The caches' queries followed every pending training step.
    for (uint32_t i = 0; i < n; ++i)
        if (jobs[i].nrc_query) b->renderers[i]->nrc_pending = false;
    if (e == RGI_OK) e = gi_resample_batch(L, stream);
    for (uint32_t i = 0; i < n && e == RGI_OK; ++i) {
        rgi_renderer* r = b->renderers[i];
        if (!jobs[i].split)
            e = gi_temporal(r->ctx, jobs[i].fg, r->scene, gbuffers[i], jobs[i].tin, jobs[i].gi_out, r->ctx->tuning,
                            stream);
    }
    hipStreamWaitEvent(stream, b->gridded, 0);
    if (e == RGI_OK) e = se;
    if (e == RGI_OK) e = gi_spatial_batch(L, stream);
    // The rest of each view's frame. The caches' steps train in one chain of
    // launches on the caches' stream, past the end of the call (as
    // rgi_renderer_render leaves a training step): beside the composites and
    // the next frame's DI passes and tracing, ahead of its queries.
    if (e == RGI_OK) e = specular_batch(L.views, L.first[kPassSpecular], n, L.blocks[kPassSpecular], L.same, L.larger, stream);
    for (uint32_t i = 0; i < n; ++i)
        if (hv[i].spec_out) b->renderers[i]->specular_read = 1 - b->renderers[i]->specular_read;
    if (e == RGI_OK) e = svgf_batch(L.views, L.first[kPassDenoise], n, L.blocks[kPassDenoise], (int)max_atrous, stream);
    for (uint32_t i = 0; i < n; ++i)
        b->renderers[i]->denoised = hv[i].atrous > 0u ? hv[i].den_out[hv[i].atrous] : b->renderers[i]->color[jobs[i].write];
    // The gathers overwrite the records that each cache's last training step
    // reads: `stream` joins the steps still pending.
    for (uint32_t i = 0; i < n; ++i)
        if (jobs[i].nrc_on) join_nrc(b->renderers[i], stream);
    if (any_train && e == RGI_OK)
        e = nrc_gather_batch(b->table + T.gathers, reinterpret_cast<const uint32_t*>(b->table + T.gather_first), n,
                             gather_blocks, stream);
    if (any_train && e == RGI_OK) {
        hipEventRecord(b->gathered, stream);
        hipStreamWaitEvent(b->cache_side, b->gathered, 0);
        e = nrc_train_batch(b->table + T.steps, reinterpret_cast<const uint32_t*>(b->table + T.nrc_first), n,
                            nrc_blocks, b->cache_side);
        for (uint32_t i = 0; i < n; ++i) {
            rgi_renderer* r = b->renderers[i];
            if (!jobs[i].nrc_on) continue;
            hipEventRecord(r->nrc_done, b->cache_side);
            r->nrc_pending = true;
            if (e == RGI_OK) {
                nrc_step_taken(r->nrc);
                ++r->nrc_trained;
            }
        }
    }
    if (e == RGI_OK) e = composite_batch(L.views, L.first[kPassComposite], n, L.blocks[kPassComposite], stream);
    for (uint32_t i = 0; i < n && e == RGI_OK; ++i) frame_commit(b->renderers[i], jobs[i], &views[i], stream, true);
    return e;
}

} // namespace

extern "C" {

int rgi_batch_render(rgi_batch* b, const rgi_view* views, const rgi_gbuffer* gbuffers, void* stream_ptr)
{
    if (!b || !views || !gbuffers) return fail(RGI_E_INVALID, "batch_render: NULL argument");
    hipStream_t stream = as_stream(stream_ptr);
    if (b->mode == RGI_BATCH_GRAPH) return render_graph_branches(b, views, gbuffers, stream);
    return render_passes(b, views, gbuffers, stream);
}

} // extern "C"
