/*
 * restir_amd — C ABI of the MI355X (gfx950) ReSTIR GI/DI hot path.
 *
 * The reference renderer (pkonecki/restir) has no plugin API: each pass is a
 * bgfx compute program whose inputs are bound by stage number and dispatched
 * from Renderer::render (src/renderer.cpp:626-833). Every entry point below
 * replaces one such dispatch, keeping its inputs and outputs (G-buffer planes,
 * reservoir planes, the hash grid buffer and the uniforms) as caller-owned
 * device buffers with explicit formats. Calls are stream ordered and
 * asynchronous; `stream` is a hipStream_t (NULL: the legacy default stream).
 * Nothing crosses the ABI but plain pointers, sizes and POD structs.
 *
 * Conventions (what the reference leaves to bgfx's shader library and GPU):
 *   - Matrices are 16 floats, column major (element [col * 4 + row]), and
 *     applied as M * v.
 *   - A texel fetch at texture coordinate uv of a W x H plane reads texel
 *     (clamp(floor(uv.x * W), 0, W - 1), clamp(floor(uv.y * H), 0, H - 1)),
 *     row-major, row 0 first: point sampling with clamp, as the reference's
 *     samplers are set up (renderer.cpp:1324-1328).
 *   - RGBA8 planes are uint32 words, red in the low byte; a channel reads as
 *     byte / 255. RGBA16F planes are 4 IEEE binary16 values per texel, written
 *     with round-to-nearest-even. RGBA32F planes are 4 floats per texel.
 *   - toClipSpaceDepth(d) is d * 2 - 1 with homogeneous_depth, else d;
 *     toClipSpaceY(y) is y with origin_bottom_left, else -y.
 *
 * Return value of every int function: RGI_OK, or a negative RGI_E_* code
 * whose text rgi_last_error() returns (per thread).
 */
#ifndef RESTIR_AMD_H
#define RESTIR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RGI_ABI_VERSION 1

#define RGI_OK 0
#define RGI_E_INVALID (-1)   /* an argument is out of range or inconsistent */
#define RGI_E_DEVICE (-2)    /* a HIP call failed */
#define RGI_E_NOMEM (-3)     /* a device allocation failed */

/* Ray modes, RAY_MODE_* in shaders/bvh.sh:17-19 (RayMode, settings.h:58-65). */
#define RGI_RAY_SCREEN 0
#define RGI_RAY_HYBRID 1
#define RGI_RAY_BVH 2

/* Grid insert modes (gridInsert, shaders/hashgrid.sh:182-200). */
#define RGI_GRID_ATOMIC 0        /* the reference's atomics, in whatever order the GPU runs them */
#define RGI_GRID_DETERMINISTIC 1 /* one fixed linearization of those atomics (see rgi_grid_insert) */

/* One frame's uniforms: the u_* vec4/mat4 uniforms of the hot-path passes,
 * one field each (restir.sh:57-71, hashgrid.sh:49-50, bvh.sh:14, di.sh:35-39,
 * shadow.sh:23-25, sky.sh:11-13, cs_di_temporal.sc:28-30), packed by the
 * reference at renderer.cpp:497-531, 614-621, 632-640, 693-731. Integers the
 * reference passes as floats are integers here. */
typedef struct rgi_frame {
    float inv_view_proj[16];      /* u_sceneInvViewProj */
    float view_proj[16];          /* u_sceneViewProj */
    float prev_view_proj[16];     /* u_previousViewProj */
    float camera_position[4];     /* u_cameraPosition; u_gridCamera of insert and GI spatial */
    float prev_camera_position[4];/* u_diPrevious; u_gridCamera of GI temporal (renderer.cpp:813) */
    float light_position[4];      /* u_lightPosition (main point light) */
    float light_color[4];         /* u_lightColor */
    uint32_t frame_index;         /* u_frameInfo.x (frame number & 0xffffff) */
    int32_t history_valid;        /* u_frameInfo.y */
    int32_t origin_bottom_left;   /* u_frameInfo.z */
    int32_t multiple_bounces;     /* u_frameInfo.w */
    int32_t homogeneous_depth;    /* bgfx caps homogeneousDepth (toClipSpaceDepth) */
    int32_t gi_width;             /* u_screenInfo.x: GI target size */
    int32_t gi_height;            /* u_screenInfo.y */
    int32_t width;                /* u_screenSize.x: full-resolution size (DI passes) */
    int32_t height;               /* u_screenSize.y */
    float scene_scale;            /* u_sceneInfo.x */
    float boiling_filter;         /* u_sceneInfo.y (0: off) */
    float jacobian_clamp;         /* u_sceneInfo.z (0: off) */
    float outlier_clamp;          /* u_sceneInfo.w (0: off) */
    float lighting_surface;       /* u_lightingInfo.x: scale of the surface light plane */
    float lighting_emission;      /* u_lightingInfo.y: scale of off-screen triangle emission */
    int32_t initial_candidates;   /* u_restirOptions.x */
    int32_t temporal_taps;        /* u_restirOptions.y */
    int32_t spatial_taps;         /* u_restirOptions.z */
    int32_t trace_steps;          /* u_restirOptions.w */
    float grid_min_cell;          /* u_gridInfo.x: smallest cell size */
    float grid_pixel_size;        /* u_gridInfo.y: world size of a GI pixel at unit distance */
    float grid_cell_pixels;       /* u_gridInfo.z: cell size in GI pixels */
    int32_t world_space_reuse;    /* u_gridInfo.w */
    uint32_t grid_table_size;     /* kGridTableSize (hashgrid.sh:53); any power of two >= 64 */
    int32_t ray_mode;             /* u_bvhInfo.w: RGI_RAY_* */
    float sky_zenith[4];          /* u_skyZenith (w: 1 when the sky is on) */
    float sky_horizon[4];         /* u_skyHorizon */
    float sky_ground[4];          /* u_skyGround */
    float shadow_view_proj[96];   /* u_shadowViewProj[6] */
    float shadow_info[4];         /* u_shadowInfo: far plane, tan(half fov), 1 / face size, origin bottom-left */
    int32_t shadows_enabled;      /* u_shadowOptions.x */
    int32_t shadow_width;         /* shadow atlas size in texels (3 x 2 faces) */
    int32_t shadow_height;
    int32_t di_light_count;       /* u_diInfo.x */
    int32_t di_candidates;        /* u_diInfo.y */
    int32_t di_spatial_taps;      /* u_diInfo.z */
    int32_t di_history_valid;     /* u_diInfo.w */
    int32_t reserved[3];
} rgi_frame;

/* The scene: the BVH buffer in the layout of src/bvh.h:9-22 (nodes of 2 vec4,
 * then triangle records of 4 vec4, indices stored as float values) and the
 * light list in the layout of src/lights.h:7-14 (4 vec4 per light). */
typedef struct rgi_scene {
    const float* bvh;             /* b_scene: (2 * node_count + 4 * triangle_count) float4 */
    uint32_t node_count;          /* u_bvhInfo.z (0: no BVH) */
    uint32_t triangle_offset;     /* u_bvhInfo.x, in float4 (= 2 * node_count) */
    uint32_t triangle_count;
    uint32_t light_count;
    const float* lights;          /* b_lights: 4 * light_count float4 */
} rgi_scene;

/* G-buffer planes at one resolution (the GI size for the GI passes and the
 * grid insert, the full size for the DI passes). Planes a pass does not read
 * may be NULL. */
typedef struct rgi_gbuffer {
    int32_t width;
    int32_t height;
    const float* depth;           /* D32F window depth; >= 0.99999 is the sky */
    const uint32_t* albedo;       /* RGBA8: base colour, a: roughness */
    const uint32_t* normal;       /* RGBA8: normal * 0.5 + 0.5, a: metallic */
    const uint32_t* shading_normal; /* RGBA8: shading normal * 0.5 + 0.5 (DI) */
    const uint16_t* emission;     /* RGBA16F: emitted radiance (GI-only lighting's surface light) */
} rgi_gbuffer;

/* One frame's GI reservoirs: three RGBA32F planes at the GI size
 * (restir.sh:11-15): sample position + M, receiver position + W, receiver
 * normal + sample id. */
typedef struct rgi_gi_reservoirs {
    float* sample;
    float* receiver;
    float* normal;
} rgi_gi_reservoirs;

/* What the GI passes read besides the G-buffer (setGiPassInputs,
 * renderer.cpp:769-798). */
typedef struct rgi_gi_inputs {
    const float* history_sample;   /* last frame's reservoirs (stage 3 of the temporal pass) */
    const float* history_receiver; /* stage 9 */
    const float* history_normal;   /* stage 6 */
    const uint16_t* history_color; /* RGBA16F, stage 5: last frame's (filtered) demodulated GI */
    const uint16_t* history_moments; /* RGBA16F, stage 8 (spatial pass) */
    const uint16_t* surface_light; /* RGBA16F, stage 14: DI diffuse, or the G-buffer emission */
    const float* shadow_map;       /* R32F atlas, stage 7 (NULL when shadows are off) */
    const uint32_t* grid;          /* stage 11: last frame's grid (temporal), this frame's (spatial) */
} rgi_gi_inputs;

/* DI reservoirs at full resolution (di.sh:12-16): sample RGBA32F (point on
 * the light, light index), receiver RGBA32F (position, W), normal RGBA16F
 * (receiver normal, M). */
typedef struct rgi_di_reservoirs {
    float* sample;
    float* receiver;
    uint16_t* normal;
} rgi_di_reservoirs;

const char* rgi_last_error(void);
int rgi_abi_version(void);
/* sizeof of the ABI structs (0: rgi_frame, 1: rgi_scene, 2: rgi_gbuffer,
 * 3: rgi_gi_reservoirs, 4: rgi_gi_inputs, 5: rgi_di_reservoirs, 6: rgi_shadow,
 * 7: rgi_composite_inputs, 8: rgi_nrc_config, 9: rgi_nrc_settings), so that
 * bindings can check their mirrors. */
uint32_t rgi_abi_sizeof(int which);

/* Library state: scratch buffers of the deterministic grid insert and of the
 * traversal queue, and GI temporal's surface-tile list and candidate records
 * (about 56 bytes per GI pixel and candidate), grown on demand. One context
 * per host thread; calls that share a context run in stream order. */
typedef struct rgi_context rgi_context;
int rgi_create(rgi_context** out_context);
void rgi_destroy(rgi_context* context);

/* Launch variants of the kernels, chosen per context (0 = default). They
 * change how work maps onto the GPU, never a result: every variant is
 * bit-exact with every other. rgi_autotune-style callers time them on their
 * workload and keep the fastest. */
#define RGI_OPT_GI_TEMPORAL_WAVES 1 /* occupancy target of GI temporal: 0 (compiler's choice) or 4 waves per SIMD */
#define RGI_OPT_GI_SPATIAL_WAVES 2  /* occupancy target of GI spatial: 0 (compiler's choice) or 5 waves per SIMD (default) */
#define RGI_OPT_TRACE_PERSISTENT 3  /* rgi_trace: 0 one ray per lane, 1 persistent 64-ray batches per wave */
/* Screen passes' 16 x 16 tiles: 0 row-major; 1 a contiguous band of tiles per
 * XCD; 2 64 x 64-pixel blocks dealt to the 8 XCDs in turn. */
#define RGI_OPT_TILE_ORDER 4
/* Not a launch variant: the triangle test of rgi_trace. 0 (default) the
 * reference's Moeller-Trumbore test (bvh.sh:44-71), 1 the watertight test of
 * Woop, Benthin and Wald 2013 (FP32, with an FP64 recomputation of an edge
 * function that comes out exactly zero), which never lets a ray through the
 * shared edge or vertex of two triangles. The resampling passes and the
 * G-buffer and shadow producers always use the reference's test. */
#define RGI_OPT_TRIANGLE_TEST 5
/* Screen passes' workgroup shape: 0 16 x 16 pixels, 1 32 x 8 pixels (both 256
 * threads; GI keeps the reference's 4 x 8 groups either way). */
#define RGI_OPT_TILE_SHAPE 6
int rgi_context_set_option(rgi_context* context, int option, int value);
int rgi_context_get_option(const rgi_context* context, int option);

/* cs_grid_clear (shaders/cs_grid_clear.sc:12-21, dispatched at
 * renderer.cpp:735-738): zeroes the checksums and counts of `grid`, a buffer
 * of (2 + 32) * table_size uint32 laid out [checksums | counts | entries]. */
int rgi_grid_clear(uint32_t* grid, uint32_t table_size, void* stream);

/* cs_grid_insert (shaders/cs_grid_insert.sc:40-73, dispatched at
 * renderer.cpp:740-750): lists every GI pixel of `gbuffer` (depth and normal,
 * at the GI size) in the cell of its jittered receiver. The grid camera is
 * frame->camera_position. RGI_GRID_ATOMIC runs gridInsert's atomics as the GPU
 * schedules them. RGI_GRID_DETERMINISTIC gives the result of one legal order
 * of them: probe round by probe round, pixels in ascending index within a
 * round; a cell's count is every pixel resolved to it, and its entries are
 * its 32 lowest pixel indices, ascending. Both modes expect `grid` cleared by
 * rgi_grid_clear first, as the reference always clears it before inserting. */
int rgi_grid_insert(rgi_context* context, const rgi_frame* frame, const rgi_gbuffer* gbuffer,
                    uint32_t* grid, int mode, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). rgi_grid_insert in RGI_GRID_DETERMINISTIC mode, the
 * same kernels and the same result, with its scratch in a fixed device array
 * of the library (RGI_NOALLOC_SCRATCH_BYTES, 88 MiB plus 8,704 bytes, held by
 * each GPU the library runs on until the process ends) instead of a
 * context's: it needs no context and calls no allocator. GI sizes up to
 * RGI_NOALLOC_MAX_PIXELS pixels (1920 x 1080 fits) and tables up to
 * RGI_NOALLOC_MAX_TABLE cells; larger inputs return RGI_E_INVALID. Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialized by the caller.
 * The first call on a device looks up the array's address on it. */
#define RGI_NOALLOC_MAX_PIXELS (1u << 21)
#define RGI_NOALLOC_MAX_TABLE (1u << 22)
int rgi_grid_insert_noalloc(const rgi_frame* frame, const rgi_gbuffer* gbuffer, uint32_t* grid, void* stream);

/* cs_restir_temporal (shaders/cs_restir_temporal.sc:57-297, dispatched at
 * renderer.cpp:807-817): fresh candidates, temporal reuse through last
 * frame's grid (inputs->grid), boiling filter; writes this frame's GI
 * reservoirs. With a context (and at most 8 candidates) the pass lists its
 * surface tiles, marches the candidates and walks the rays the march leaves
 * open in kernels of their own before resampling; without one it runs as
 * one kernel. Both give the same result. */
int rgi_gi_temporal(rgi_context* context, const rgi_frame* frame, const rgi_scene* scene,
                    const rgi_gbuffer* gbuffer, const rgi_gi_inputs* inputs,
                    const rgi_gi_reservoirs* out, void* stream);

/* cs_restir_spatial (shaders/cs_restir_spatial.sc:95-246, dispatched at
 * renderer.cpp:821-833): spatial reuse through this frame's grid
 * (inputs->grid) over this frame's reservoirs, shading, temporal
 * accumulation. Writes RGBA16F color (demodulated estimate, history length)
 * and moments (mean luminance, mean squared luminance, history length, 0). */
int rgi_gi_spatial(rgi_context* context, const rgi_frame* frame, const rgi_scene* scene,
                   const rgi_gbuffer* gbuffer, const rgi_gi_inputs* inputs,
                   const rgi_gi_reservoirs* reservoirs, uint16_t* out_color, uint16_t* out_moments,
                   void* stream);

/* cs_di_temporal (shaders/cs_di_temporal.sc:32-180, dispatched at
 * renderer.cpp:660-671), at full resolution. */
int rgi_di_temporal(rgi_context* context, const rgi_frame* frame, const rgi_scene* scene,
                    const rgi_gbuffer* gbuffer, const rgi_di_reservoirs* previous,
                    const rgi_di_reservoirs* out, void* stream);

/* cs_di_spatial (shaders/cs_di_spatial.sc:30-169, dispatched at
 * renderer.cpp:673-681): writes RGBA16F diffuse and specular radiance. */
int rgi_di_spatial(rgi_context* context, const rgi_frame* frame, const rgi_scene* scene,
                   const rgi_gbuffer* gbuffer, const rgi_di_reservoirs* reservoirs,
                   uint16_t* out_diffuse, uint16_t* out_specular, void* stream);

/* traceBvh (shaders/bvh.sh:298-319) over `count` rays, each 8 floats:
 * origin.xyz, max distance, direction.xyz, unused. Closest hit replays
 * traceBvhSharedStack's visit order (bvh.sh:216-295); any hit is
 * traceBvhStackless (bvh.sh:140-174). Writes the triangle (BVH order, -1 for
 * a miss) and the distance (the max distance on a miss). */
int rgi_trace(rgi_context* context, const rgi_scene* scene, const float* rays, uint32_t count,
              int32_t* hit_triangle, float* hit_distance, int any_hit, void* stream);

/* The G-buffer of the scene seen through `frame` (inv_view_proj,
 * camera_position, homogeneous_depth, origin_bottom_left, width x height), one
 * primary ray per pixel centre through the BVH: the producer of the GI
 * passes' inputs, standing in for fs_gbuffer.sc. `materials` holds one float4
 * per material index (roughness, metallic, 0, 0). Writes depth, albedo,
 * normal, shading normal (= normal) and emission planes. */
int rgi_gbuffer_primary(rgi_context* context, const rgi_frame* frame, const rgi_scene* scene,
                        const float* materials, uint32_t material_count, float* depth,
                        uint32_t* albedo, uint32_t* normal, uint32_t* shading_normal,
                        uint16_t* emission, void* stream);

/* The main point light's omnidirectional shadow map (renderer.cpp:527-564,
 * fs_shadow.sc, shadow.sh): six cube faces (+X, -X, +Y, -Y, +Z, -Z) in a
 * 3 x 2 atlas of face_size texels per face side, face f at column f % 3,
 * row f / 3 (rows counted from the top; from the bottom when
 * origin_bottom_left). Each face is a right-handed look-at from the light
 * (up +Z for the Y faces, +Y otherwise) with a square perspective of
 * tan(half fov) = F / (F - 8), near 0.01 radius and far 2.5 radius. */
typedef struct rgi_shadow {
    float view_proj[96];      /* u_shadowViewProj[6], column-major */
    float inv_view_proj[96];  /* their inverses (computed in double, rounded once) */
    float info[4];            /* u_shadowInfo: far, tan(half fov), 1 / face size, origin bottom-left */
    float light_position[4];
    float near_plane;
    int32_t face_size;
    int32_t homogeneous_depth;
    int32_t reserved;
} rgi_shadow;

/* Fills `out` for a light at `light_position` in a scene of `scene_radius`;
 * face_size is one of 256, 512, 1024, 2048 (kShadowFaceSizes, settings.h:26). */
int rgi_shadow_setup(const float* light_position, float scene_radius, int32_t face_size, int homogeneous_depth,
                     int origin_bottom_left, rgi_shadow* out);

/* Copies the shadow uniforms into a frame (u_shadowViewProj, u_shadowInfo,
 * atlas size) and sets shadows_enabled, as renderer.cpp:782-783 binds them. */
int rgi_shadow_apply(const rgi_shadow* shadow, int enabled, rgi_frame* frame);

/* The shadow pass (renderer.cpp:566-596 with fs_shadow.sc): for every atlas
 * texel, the distance from the light to the nearest surface along the
 * texel centre's ray, divided by the far plane, every triangle two-sided,
 * surfaces before the near plane or past the far plane clipped; texels that
 * see nothing keep the clear value 1. One closest-hit BVH ray per texel takes
 * the place of rasterizing the scene six times. `atlas` holds
 * 3 face_size x 2 face_size floats, row-major. */
int rgi_shadow_atlas(rgi_context* context, const rgi_shadow* shadow, const rgi_scene* scene, float* atlas,
                     void* stream);

/* The denoiser (renderer.cpp:857-907), over the GI planes of frame->width x
 * frame->height texels. Step 1, fs_svgf_variance.sc: from this frame's
 * accumulated colour and moments (rgi_gi_spatial's outputs) and the GI
 * receiver and normal planes (rgi_gi_reservoirs.receiver, .normal), writes
 * rgb = colour, a = luminance variance. */
int rgi_svgf_variance(rgi_context* context, const rgi_frame* frame, const uint16_t* color, const uint16_t* moments,
                      const float* receiver, const float* normal, uint16_t* out, void* stream);

/* Step 2, fs_svgf_atrous.sc: one edge-avoiding a-trous iteration with taps
 * `step` texels apart (1, 2, 4, 8, 16 for iterations 0-4), input and output
 * as rgi_svgf_variance writes them; `out` must not alias `input`. The
 * reference keeps iteration 0's output as next frame's GI history colour. */
int rgi_svgf_atrous(rgi_context* context, const rgi_frame* frame, const uint16_t* input, const float* receiver,
                    const float* normal, int32_t step, uint16_t* out, void* stream);

/* Indirect specular (fs_specular.sc, dispatched at renderer.cpp:836-855):
 * one GGX visible-normal ray per GI pixel, traced like the GI rays (depth
 * buffer, then BVH, then sky) and shaded with last frame's composite where
 * the hit is a visible surface that was on screen. `in` supplies last frame's
 * GI receiver, normal and colour planes (the same-surface test and bounce
 * light), the surface light and the shadow atlas (history_sample and grid
 * are not read). `previous_hdr`: last frame's composite (NULL: none);
 * `history`: last frame's output of this pass (NULL: none). Writes RGBA16F
 * rgb = demodulated specular radiance, a = history length. */
int rgi_specular_gi(rgi_context* context, const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* gbuffer,
                    const rgi_gi_inputs* in, const uint16_t* previous_hdr, const uint16_t* history, uint16_t* out,
                    void* stream);

/* The composite pass's inputs besides the G-buffer (fs_composite.sc:10-29).
 * GI planes are frame->width x frame->height (the GI resolution); the DI
 * planes, the G-buffer and the output are gbuffer->width x height. */
typedef struct rgi_composite_inputs {
    const uint16_t* gi;          /* RGBA16F demodulated GI: the denoiser's output, or the accumulated colour */
    const uint16_t* di_diffuse;  /* RGBA16F ReSTIR DI light (NULL: none) */
    const uint16_t* di_specular; /* RGBA16F (NULL: none) */
    const uint16_t* specular;    /* RGBA16F demodulated indirect specular (NULL: none) */
    const float* shadow_map;     /* R32F shadow atlas (NULL: every light sample visible) */
    /* RGBA16F, NULL for the reference's composite. When given, the pixel's
     * whole diffuse light, taken as is in place of di_diffuse plus the
     * re-modulated GI irradiance (the neural radiance cache's per-pixel
     * prediction, include/restir_amd_nrc.h); the indirect specular still
     * comes from the GI planes. No counterpart in the reference. */
    const uint16_t* resolved_diffuse;
    float options[4];            /* u_compositeOptions: direct, indirect, albedo modulation, emission (final image: 1, 1, 1, 1) */
    float specular_weight;       /* u_compositeSpecular.x: 1 with indirect specular */
    float reserved[3];
} rgi_composite_inputs;

/* fs_composite.sc:82-143: the scene-referred HDR image, RGBA16F. Per pixel:
 * the main light through the BRDF with the 4 x 4 soft shadow-atlas compare,
 * plus DI diffuse and specular; the GI irradiance upsampled over the four
 * nearest same-surface GI texels (fs_composite.sc:48-80) times albedo,
 * (1 - metallic) and ambient occlusion (emission.a); indirect specular times
 * the specular albedo; emission. Sky pixels: the sky radiance with alpha 1,
 * or 0 when the sky is off. */
int rgi_composite(rgi_context* context, const rgi_frame* frame, const rgi_gbuffer* gbuffer,
                  const rgi_composite_inputs* inputs, uint16_t* out_hdr, void* stream);

/* The image the error metric compares, per pixel (linear RGB): the GI
 * estimate re-modulated by the diffuse albedo, color.rgb * albedo.rgb *
 * (1 - metallic), plus the DI diffuse and specular radiance when those planes
 * are given (the light the reference composites, fs_composite.sc). Pixels
 * whose depth is the sky are not part of the image. */

/* fs_reference.sc's running mean, in FP64: accum[3 * i + c] += image_c(i) for
 * every pixel of a width x height frame. Divide by the frame count to read the
 * reference. NaN contributions are skipped, as fs_reference.sc:21-23 skips
 * them. */
int rgi_reference_accumulate(const rgi_gbuffer* gbuffer, const uint16_t* color, const uint16_t* di_diffuse,
                             const uint16_t* di_specular, double* accum, void* stream);

/* cs_error.sc's measurement against accum / frames, over geometry pixels, in
 * FP64 with a fixed reduction order: out[0] relative MSE in luminance,
 * mean((L - L_ref)^2 / (L_ref^2 + 0.01)) (cs_error.sc:59-63); out[1] MSE in
 * linear RGB (mean over pixels and channels); out[2] mean image luminance;
 * out[3] mean reference luminance; out[4] pixels compared. `out` is a device
 * buffer of 5 doubles. */
int rgi_reference_error(rgi_context* context, const rgi_gbuffer* gbuffer, const uint16_t* color,
                        const uint16_t* di_diffuse, const uint16_t* di_specular, const double* accum,
                        uint32_t frames, double* out, void* stream);

/* The same two on the composite's HDR image (rgi_composite), the image the
 * reference accumulates and measures (fs_reference.sc samples the composite,
 * cs_error.sc the display before tone mapping; this build has no TAA). The
 * G-buffer supplies only the depth that selects geometry pixels. */
int rgi_reference_accumulate_composite(const rgi_gbuffer* gbuffer, const uint16_t* hdr, double* accum, void* stream);
int rgi_reference_error_composite(rgi_context* context, const rgi_gbuffer* gbuffer, const uint16_t* hdr,
                                  const double* accum, uint32_t frames, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RESTIR_AMD_H */
