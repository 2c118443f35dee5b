/*
 * restir_amd neural radiance cache — C ABI.
 *
 * The reference has no counterpart: this is the learned extension of the GI
 * path. A small MLP over a multiresolution hash encoding of position (Mueller
 * et al. 2022, "Instant Neural Graphics Primitives") predicts the radiance
 * that a diffuse surface point reflects, trained online on the surface points
 * that ReSTIR GI and DI shade each frame (Mueller et al. 2021, "Real-time
 * Neural Radiance Caching for Path Tracing": relative L2 loss, Adam). The
 * frame driver queries it at every visible surface and hands the result to
 * the GI passes in place of the reference's surface light plus last frame's
 * indirect light (restir.sh:265-298, evaluateSample 445-530), so it stands in
 * for the rest of the path and becomes part of the GI target function.
 *
 * Model (fixed shape):
 *   encoding  16 levels x 2 features, 2^log2_table entries per level, level l
 *             with res_l = floor(base * b^l) cells per axis, b chosen so that
 *             res_15 = max_resolution. A level whose (res+1)^3 corners fit the
 *             table is indexed densely, x + y (res+1) + z (res+1)^2; finer
 *             levels hash the corner, (x * 1 ^ y * 2654435761 ^ z * 805459861)
 *             mod 2^log2_table (Teschner et al. 2003). Trilinear weights.
 *   input     64 features: the 32 encoded values, normal * 0.5 + 0.5 (3),
 *             albedo (3), a constant 1, zeros.
 *   MLP       5 hidden layers of 64 with ReLU, a linear output of 3, no
 *             biases (the constant input feeds the first layer's).
 *   loss      relative L2 per sample, sum_c (y_c - t_c)^2 / (sg(mean_c y)^2
 *             + 0.01), times the sample's weight, averaged over the batch.
 *   optimiser Adam with bias correction on FP32 master weights; a table entry
 *             whose gradient is exactly zero keeps its value and moments.
 *   average   after each step, every parameter's moving average e = d e +
 *             (1 - d) w with d = min(ema_decay, (1 + t) / (10 + t)) at step
 *             t, in FP32 (both products rounded, then their sum); inference
 *             reads the averaged weights and table, training the parameters.
 *             Creation and rgi_nrc_set_params restart it at the parameters.
 *
 * Precision: training runs the MLP in BF16 on the matrix cores (products of
 * BF16 operands, FP32 accumulation); activations and back-propagated
 * gradients are rounded to BF16 between layers, table values to FP16. The
 * table gradient is accumulated with FP32 atomics, so its last bits depend on
 * the order the GPU runs them. Inference runs in BF16, FP16 or FP8 (OCP E4M3,
 * per-layer scales for weights and activations).
 *
 * Buffers are caller-owned device pointers; calls are stream ordered.
 * Records are 12 floats per point: position xyz, 0, unit normal xyz, 0,
 * albedo rgb, 0. Targets are 4 floats: radiance rgb and a weight (0 drops the
 * sample, 1 keeps it). Outputs are 4 floats: radiance rgb and 0.
 */
#ifndef RESTIR_AMD_NRC_H
#define RESTIR_AMD_NRC_H

#include <stdint.h>

#include "restir_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RGI_NRC_LEVELS 16
#define RGI_NRC_FEATURES 2
#define RGI_NRC_WIDTH 64
#define RGI_NRC_HIDDEN 5
#define RGI_NRC_OUTPUTS 3
/* Logical MLP parameters: W1..W5 (64 x 64, [output][input]) then W6 (3 x 64). */
#define RGI_NRC_MLP_PARAMS (5 * 64 * 64 + 3 * 64)
#define RGI_NRC_RECORD_FLOATS 12
#define RGI_NRC_TARGET_FLOATS 4

/* Inference precision. */
#define RGI_NRC_BF16 0
#define RGI_NRC_FP16 1
#define RGI_NRC_FP8 2

typedef struct rgi_nrc rgi_nrc;

typedef struct rgi_nrc_config {
    uint32_t log2_table;      /* 19: entries per level = 2^log2_table (10..22) */
    uint32_t base_resolution; /* 16: cells per axis of level 0 */
    uint32_t max_resolution;  /* 512: cells per axis of level 15 */
    uint32_t seed;            /* 1: initialisation of the weights and the table */
    float learning_rate;      /* 0.005 */
    float beta1;              /* 0.9 */
    float beta2;              /* 0.99 */
    float epsilon;            /* 1e-8 */
    float aabb_min[3];        /* positions map linearly onto [0, 1]^3, clamped */
    float aabb_max[3];
    float ema_decay;          /* 0.95: inference reads a moving average of the parameters (0: the parameters) */
    int32_t reserved[3];
} rgi_nrc_config;

void rgi_nrc_config_default(rgi_nrc_config* config);

/* Cells per axis of `level` for `config` (the host computes them in double). */
int rgi_nrc_level_resolution(const rgi_nrc_config* config, uint32_t level, uint32_t* resolution);

/* Allocates the parameters (FP32 master table and weights, their FP16, BF16
 * and FP8 copies), the Adam moments and the gradient buffers on the current
 * device and initialises them from config->seed: weights uniform in
 * +-sqrt(6 / (fan_in + fan_out)), table values uniform in +-1e-4. */
int rgi_nrc_create(rgi_nrc** out, const rgi_nrc_config* config);
void rgi_nrc_destroy(rgi_nrc* nrc);

/* One Adam step on `count` records and targets (any count >= 1). `loss`, if
 * not NULL, is a device float that receives the batch loss. */
int rgi_nrc_train(rgi_nrc* nrc, const float* records, const float* targets, uint32_t count, float* loss,
                  void* stream);

/* One Adam step of each of `count` distinct caches, caches[i] on counts[i]
 * records and targets (losses may be NULL, or hold NULL entries): each
 * kernel of rgi_nrc_train's chain runs once for all the caches, each cache
 * taking its own slice of the workgroups, so that small steps fill the chip
 * together. Each cache ends where rgi_nrc_train would leave it (the table
 * gradient accumulates with FP32 atomics in either, so the table's last bits
 * follow the order the GPU runs them). Returns after the steps have run (its
 * argument table is allocated and freed per call); a renderer batch trains
 * its caches the same way from its own table without that wait. */
int rgi_nrc_train_batch(rgi_nrc* const* caches, const float* const* records, const float* const* targets,
                        const uint32_t* counts, float* const* losses, uint32_t count, void* stream);

/* Radiance predictions for `count` records into out[4 * count]. */
int rgi_nrc_infer(rgi_nrc* nrc, const float* records, uint32_t count, int precision, float* out, void* stream);

/* The loss and its gradients without a step: mlp_grad[RGI_NRC_MLP_PARAMS]
 * (logical layout) and, if not NULL, table_grad[16 * 2^log2_table * 2]
 * ([level][entry][feature]). Both are device buffers that the call
 * overwrites. `loss` as for rgi_nrc_train. */
int rgi_nrc_gradients(rgi_nrc* nrc, const float* records, const float* targets, uint32_t count, float* mlp_grad,
                      float* table_grad, float* loss, void* stream);

/* The FP32 master parameters: mlp[RGI_NRC_MLP_PARAMS] and table[16 *
 * 2^log2_table * 2] (device buffers; either may be NULL). Setting them
 * clears the Adam moments and the step count. */
int rgi_nrc_get_params(const rgi_nrc* nrc, float* mlp, float* table, void* stream);
int rgi_nrc_set_params(rgi_nrc* nrc, const float* mlp, const float* table, void* stream);

/* The moving average of the parameters, laid out as rgi_nrc_get_params (the
 * parameters themselves when ema_decay is 0). */
int rgi_nrc_get_average(const rgi_nrc* nrc, float* mlp, float* table, void* stream);

/* Adam steps taken since creation or the last rgi_nrc_set_params. */
int rgi_nrc_steps(const rgi_nrc* nrc, uint32_t* steps);

/* The cache at every pixel of a G-buffer: the radiance of each visible
 * surface (position reconstructed from the depth with frame->inv_view_proj,
 * surface normal, albedo * (1 - metallic)), clamped to >= 0, plus the
 * G-buffer emission when `emission` is not NULL, into out_radiance as
 * RGBA16F with alpha 1; sky pixels get 0. This is the frame driver's query. */
int rgi_nrc_query_gbuffer(rgi_nrc* nrc, const rgi_frame* frame, const rgi_gbuffer* gbuffer, const uint16_t* emission,
                          int precision, uint16_t* out_radiance, void* stream);

/* Training records of a G-buffer, as the frame driver gathers them: `batch`
 * records stratified over the N surface pixels (depth below the far plane)
 * in row-major order, record i from surface pixel floor((i * N + u_i) /
 * batch) with u_i in [0, N) from `seed`, each with its surface inputs and
 * the target di_diffuse + albedo * color (either plane may be NULL), weight
 * 1; when the G-buffer has no surface pixel every record has weight 0.
 * records[12 * batch], targets[4 * batch]. `nrc` owns the scratch (the
 * surface list, grown on the first call at a larger size, after a device
 * synchronization). */
int rgi_nrc_gather_gbuffer(rgi_nrc* nrc, const rgi_frame* frame, const rgi_gbuffer* gbuffer,
                           const uint16_t* di_diffuse, const uint16_t* color, uint32_t batch, uint32_t seed,
                           float* records, float* targets, void* stream);

/* The same records at the frame's ReSTIR GI path vertices: record i is drawn
 * from the same stratified surface pixel, which is then replaced by the
 * pixel where that pixel's GI reservoir sample lies (reservoirs: the frame's
 * temporal reservoirs, GI-size planes; sample = point xyz and M, normal.w =
 * sample id), when the sample is a surface seen on screen (it projects onto
 * the screen and the depth buffer there reconstructs it within 0.1 scene
 * units, evaluateSample's test, restir.sh:454-461); a pixel whose sample is
 * hidden, the sky or absent (M = 0) keeps its own record. The target is taken
 * at the record's pixel as for rgi_nrc_gather_gbuffer. */
int rgi_nrc_gather_vertices(rgi_nrc* nrc, const rgi_frame* frame, const rgi_gbuffer* gbuffer,
                            const rgi_gi_reservoirs* reservoirs, const uint16_t* di_diffuse, const uint16_t* color,
                            uint32_t batch, uint32_t seed, float* records, float* targets, void* stream);

/* ---------------------------------------------------------------------
 * The cache in the frame driver (include/restir_amd_renderer.h).
 *
 * With a cache set, each frame (outside reference accumulation):
 *  1. after the DI passes and the grid, once the cache has trained `warmup`
 *     frames, queries it at every G-buffer pixel (one query per pixel) and
 *     gives the GI temporal, GI spatial and specular passes its radiance as
 *     their surface light, with multiple bounces off: a sample point seen on
 *     screen then reflects the point light plus the cache's radiance (plus
 *     its emission when the G-buffer emission is the surface light), where
 *     the reference reads the DI light plus last frame's indirect light. The
 *     GI target function and the shaded samples both use it;
 *  2. after the denoiser, gathers `batch` training records at the ReSTIR GI
 *     path vertices that the reservoirs of surface pixels stratified over
 *     the image hold (RGI_NRC_RECORDS_VERTICES, see rgi_nrc_gather_vertices;
 *     _PIXELS keeps the pixels themselves, _MIXED moves every other record),
 *     each with target DI diffuse +
 *     albedo * the next frame's GI history colour (the ReSTIR estimate of the light the
 *     surface reflects, which already contains the cache's own prediction at
 *     the sample points: self-training), and takes one Adam step. The step
 *     runs on a stream of the renderer's own, forked from the caller's: in
 *     rgi_renderer_render it may still run when the call's work on `stream`
 *     is done (beside the next frame's DI passes), and the next frame waits
 *     for it before it reads the cache; rgi_renderer_render_graph and
 *     batches join it within the frame. rgi_renderer_nrc_state waits for it
 *     on the host, so the state it returns is the last frame's.
 * With `resolve` set the cache is queried at every pixel, so records at the
 * pixels (RGI_NRC_RECORDS_PIXELS) suit it better than path vertices, which
 * crowd where GI samples land.
 * With `resolve` set (and multiple bounces on, so that the target is the
 * surface's whole diffuse light), the composite also takes the query of
 * step 1 at each pixel as the pixel's diffuse light, in place of the DI
 * diffuse light plus the re-modulated GI (rgi_composite_inputs::
 * resolved_diffuse): the cache's average over every frame it has learnt
 * from replaces one frame's DI and GI estimates, so their noise leaves the
 * image, and the cache's error (spatial detail finer than its finest
 * level, lag behind changed lighting) takes its place. Ambient occlusion
 * (emission.a) does not scale the resolved light. The GI frame's own
 * outputs (reservoirs, colour, history) are unchanged by it.
 * Until `warmup` frames have trained, and whenever the reference
 * accumulates, the frame is the plain ReSTIR path. */
typedef struct rgi_renderer rgi_renderer;

/* Where the frame driver draws its training records. */
#define RGI_NRC_RECORDS_PIXELS 0   /* surface pixels (rgi_nrc_gather_gbuffer) */
#define RGI_NRC_RECORDS_VERTICES 1 /* ReSTIR GI path vertices (rgi_nrc_gather_vertices) */
#define RGI_NRC_RECORDS_MIXED 2    /* the two in turn: even records at path vertices, odd ones at pixels */

typedef struct rgi_nrc_settings {
    rgi_nrc_config cache;     /* aabb: the scene's bounds */
    uint32_t batch;           /* 16384: training records per frame */
    uint32_t warmup;          /* 4: training frames before the GI passes read the cache */
    int32_t precision;        /* RGI_NRC_BF16: precision of the per-pixel queries */
    int32_t resolve;          /* 0; 1: the composite takes the cache's prediction as each pixel's diffuse light */
    int32_t records;          /* RGI_NRC_RECORDS_VERTICES (default), _PIXELS (better with resolve) or _MIXED */
    int32_t reserved[3];
} rgi_nrc_settings;

void rgi_nrc_settings_default(rgi_nrc_settings* settings);

/* Gives the renderer a new cache (NULL: none; the previous one is freed). */
int rgi_renderer_set_nrc(rgi_renderer* renderer, const rgi_nrc_settings* settings);

/* The renderer's cache (NULL when none), its per-pixel radiance plane
 * (RGBA16F, valid after a frame that queried it), the last training loss (a
 * device float) and the number of frames it has trained. Any output may be
 * NULL. */
int rgi_renderer_nrc_state(const rgi_renderer* renderer, rgi_nrc** cache, const uint16_t** radiance,
                           const float** loss, uint32_t* trained_frames);

#ifdef __cplusplus
}
#endif

#endif /* RESTIR_AMD_NRC_H */
