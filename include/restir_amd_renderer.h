/*
 * restir_amd frame driver — C ABI.
 *
 * The counterpart of the reference's Renderer::render for the hot path
 * (src/renderer.cpp:435-531, 614-833, 1119-1142): it owns the ping-pong
 * buffers (GI reservoirs, accumulated colour and moments, hash grids, DI
 * reservoirs), packs one rgi_frame per frame from the settings and the
 * camera, and runs DI temporal -> DI spatial -> grid clear -> grid insert ->
 * GI temporal -> GI spatial on one stream. It reproduces the reference's
 * parity swaps, the previous camera for temporal grid lookups and this
 * camera for insert and spatial reuse, the history and DI validity flags,
 * and the effective-settings override while a reference accumulates.
 */
#ifndef RESTIR_AMD_RENDERER_H
#define RESTIR_AMD_RENDERER_H

#include <stdint.h>

#include "restir_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The hot path's part of the reference's Settings (src/settings.h:183-227);
 * rgi_settings_default fills the reference defaults. */
typedef struct rgi_settings {
    int32_t initial_candidates;   /* 2 */
    int32_t temporal_taps;        /* 2 */
    int32_t spatial_taps;         /* 4 */
    int32_t trace_steps;          /* 32 */
    int32_t di_candidates;        /* 8 */
    int32_t di_spatial_taps;      /* 3 */
    int32_t ray_mode;             /* RGI_RAY_HYBRID */
    int32_t world_space_reuse;    /* 1 */
    int32_t multiple_bounces;     /* 1 */
    int32_t clamp_jacobian;       /* 1 */
    int32_t clamp_outliers;       /* 1 */
    int32_t boiling_filter;       /* 1 */
    int32_t restir_di;            /* 1: DirectLighting::ReStirDi, 0: GiOnly */
    int32_t emissive_lighting;    /* 1 */
    int32_t sky;                  /* 0 */
    int32_t shadows;              /* 0: the main light's shadow atlas (rgi_shadow_atlas) and its lookups */
    int32_t grid_mode;            /* RGI_GRID_ATOMIC */
    uint32_t grid_table_size;     /* 65536 */
    float sky_zenith[3];
    float sky_horizon[3];
    float sky_ground[3];
    float sky_intensity;          /* 1 */
    float light_position[3];      /* main point light */
    float light_color[3];         /* 0: off */
    int32_t origin_bottom_left;   /* 0 */
    int32_t homogeneous_depth;    /* 0 */
    int32_t shadow_face_size;     /* 1: index into 256, 512, 1024, 2048 (kShadowFaceSizes) */
    int32_t denoiser;             /* 1: variance + a-trous passes; GI history colour = filtered history */
    int32_t atrous_iterations;    /* 5 (kMaxAtrousIterations), 0..5 */
    int32_t composite;            /* 1: the composite pass into an RGBA16F HDR image every frame */
    int32_t specular_gi;          /* 1: indirect specular (fs_specular.sc), composited with weight 1 */
    /* 2: the GI size is the window size over this divisor, floored, at least
     * 1 (kGiResolutions, settings.h:18-23; giResolution = 1 -> Half,
     * settings.h:218; renderer.cpp:519-520). 1, 2 or 4. The GI passes, the
     * grid, the denoiser and indirect specular run at the GI size over the
     * window-size G-buffer; DI and the composite at the window size. */
    int32_t gi_divisor;
    int32_t reserved[2];
} rgi_settings;

/* One frame's camera: eye and (jittered) view-projection and its inverse. */
typedef struct rgi_view {
    float eye[4];
    float view_proj[16];
    float inv_view_proj[16];
} rgi_view;

/* Device buffers of the most recent frame, for the caller to read. */
typedef struct rgi_renderer_outputs {
    const float* gi_sample;       /* RGBA32F reservoirs */
    const float* gi_receiver;
    const float* gi_normal;
    const uint16_t* color;        /* RGBA16F accumulated demodulated GI, history length */
    const uint16_t* moments;      /* RGBA16F luminance moments */
    const uint32_t* grid;         /* this frame's hash grid */
    const uint16_t* di_diffuse;   /* RGBA16F (NULL when DI is off) */
    const uint16_t* di_specular;
    const uint16_t* filtered_history; /* RGBA16F first a-trous iteration (next frame's GI history colour) */
    const uint16_t* denoised;     /* RGBA16F the composite's GI input: the last a-trous output, or `color` */
    const uint16_t* hdr;          /* RGBA16F composite (NULL when the composite is off) */
    const uint16_t* specular_gi;  /* RGBA16F demodulated indirect specular, history length (NULL when off) */
} rgi_renderer_outputs;

typedef struct rgi_renderer rgi_renderer;

void rgi_settings_default(rgi_settings* out);
uint32_t rgi_settings_sizeof(void);

/* Allocates every per-frame buffer for a width x height window: G-buffer, DI,
 * surface light and composite planes at that size, GI reservoirs, colour,
 * moments, denoiser and indirect specular planes at the GI size
 * (settings->gi_divisor). `scene_radius` is describeScene's radius
 * (scene_scale = radius / 1.75, renderer.cpp:52, 501). */
int rgi_renderer_create(rgi_renderer** out, int32_t width, int32_t height, const rgi_settings* settings,
                        const rgi_scene* scene, float scene_radius);
void rgi_renderer_destroy(rgi_renderer* renderer);

/* Renders one frame of the hot path on `stream` from a G-buffer of the
 * renderer's window size (device planes, including shading_normal and
 * emission). */
int rgi_renderer_render(rgi_renderer* renderer, const rgi_view* view, const rgi_gbuffer* gbuffer, void* stream);

/* setReferenceAccumulation (renderer.cpp:1127-1142): while on, reuse,
 * clamps and DI reuse are off, and history restarts. */
int rgi_renderer_set_reference_mode(rgi_renderer* renderer, int on);

/* Drops the history (scene, size or settings change: renderer.cpp:416-418, 1448-1451). */
void rgi_renderer_invalidate_history(rgi_renderer* renderer);

int rgi_renderer_outputs_get(const rgi_renderer* renderer, rgi_renderer_outputs* out);

/* rgi_context_set_option on the renderer's context (launch variants). */
int rgi_renderer_set_option(rgi_renderer* renderer, int option, int value);

/* rgi_renderer_render as one HIP graph launch: the frame's launches are
 * captured from `stream` (not the null stream), the renderer's executable
 * graph is updated in place with them (instantiated when the set of passes
 * changed), then launched on `stream`. The first call renders eagerly, so
 * that scratch buffers and the shadow atlas are in their final state before
 * any capture. Results are those of rgi_renderer_render. */
int rgi_renderer_render_graph(rgi_renderer* renderer, const rgi_view* view, const rgi_gbuffer* gbuffer, void* stream);

/* A batch of up to 64 distinct renderers (any sizes and scenes) rendered
 * together; views[i] and gbuffers[i] go to renderers[i], and results are
 * those of rgi_renderer_render on each renderer. The renderers stay owned by
 * the caller and must outlive the batch; consecutive frames of a batch go to
 * one stream (or the caller orders them).
 *
 * RGI_BATCH_PASSES (the default): each screen pass of the frame (DI temporal
 * and spatial, the hash grids' clear and atomic insert, GI temporal's tile
 * list, candidate rays, BVH walks and resampling, GI spatial, indirect
 * specular, the denoiser's variance and a-trous iterations, the composite)
 * runs once for every view,
 * its kernels reading each view's frame, scene and planes from a device
 * table that the call uploads from pinned host memory; the caches' queries
 * run as one launch per precision and their training steps as one chain of
 * launches for all of them (as rgi_nrc_train_batch), their training-record
 * gathers as one chain; only the deterministic insert runs per view.
 * GI temporal's tracing and the grids run on the batch's side stream, the
 * caches' queries and training on a third stream, forked from and joined
 * back to `stream` (which may be the null stream); as with
 * rgi_renderer_render, a training step may still run when the call returns,
 * and the renderer's next use of its cache waits for it. Every view is
 * checked before the first launch: a view that fails its checks leaves
 * every renderer as it was.
 *
 * RGI_BATCH_GRAPH: one HIP graph per call holds every renderer's frame, each
 * on a stream of its own forked from and joined to `stream` (not the null
 * stream). The first call renders eagerly, one renderer after another; later
 * calls capture, update the batch's executable graph in place and launch it;
 * a capture that fails leaves every renderer as it was. */
#define RGI_BATCH_PASSES 0
#define RGI_BATCH_GRAPH 1
typedef struct rgi_batch rgi_batch;
int rgi_batch_create(rgi_batch** out, rgi_renderer* const* renderers, uint32_t count);
int rgi_batch_set_mode(rgi_batch* batch, int mode);
int rgi_batch_render(rgi_batch* batch, const rgi_view* views, const rgi_gbuffer* gbuffers, void* stream);
void rgi_batch_destroy(rgi_batch* batch);

/* The autotuner: renders `frames` frames per launch variant (cycling through
 * `views` and their G-buffers), times each variant with HIP events on
 * `stream`, and keeps the fastest. Variants are bit-exact with each other,
 * so tuning never changes an image. The search is coordinate descent over
 * four knobs, each trial keeping the best value found so far for the
 * others: trials 0-1 tile order 0/2; trials 2-3 GI temporal waves 0/4;
 * trials 4-5 GI spatial waves 0/4; trials 6-7 tile shape 16 x 16 / 32 x 8. `out_ms` (may be NULL) receives the mean
 * frame time of each of the `RGI_TUNE_TRIALS` trials. Returns the chosen
 * configuration as a bit set (RGI_TUNE_* below), which stays set on the
 * renderer, or a negative code. */
#define RGI_TUNE_TRIALS 8
#define RGI_TUNE_TEMPORAL_WAVES4 1
#define RGI_TUNE_SPATIAL_WAVES5 2
#define RGI_TUNE_TILE_ORDER 4 /* RGI_OPT_TILE_ORDER 2 */
#define RGI_TUNE_TILE_SHAPE 8 /* RGI_OPT_TILE_SHAPE 1 */
int rgi_renderer_autotune(rgi_renderer* renderer, const rgi_view* views, const rgi_gbuffer* gbuffers, uint32_t count,
                          uint32_t frames, float* out_ms, void* stream);

/* Packs the uniforms of one frame as Renderer::render does
 * (renderer.cpp:497-531, 614-621, 632-640, 693-731). `width` x `height` is
 * the window (u_screenSize); the GI size (u_screenInfo, and u_gridInfo.y's
 * pixel size) is that over effective->gi_divisor. `scene` supplies the light
 * count; `effective` settings are used as given (apply the reference
 * override before calling). */
int rgi_frame_pack(const rgi_settings* effective, const rgi_scene* scene, float scene_radius, uint32_t frame_number,
                   const rgi_view* view, const rgi_view* previous_view, int32_t width, int32_t height,
                   int history_valid, int di_history_valid, rgi_frame* out);

/* The rgi_frame the last render used (for tests and tools). */
int rgi_renderer_last_frame(const rgi_renderer* renderer, rgi_frame* out);

#ifdef __cplusplus
}
#endif

#endif /* RESTIR_AMD_RENDERER_H */
