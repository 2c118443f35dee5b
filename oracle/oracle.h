/*
 * Host oracle of the ReSTIR GI/DI hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room C++ restatement of the reference's shader logic
 * (shaders/hashgrid.sh, cs_grid_clear.sc, cs_grid_insert.sc, compute.sh,
 * bvh.sh, restir.sh, cs_restir_temporal.sc, cs_restir_spatial.sc, di.sh,
 * cs_di_temporal.sc, cs_di_spatial.sc, brdf.sh, shadow.sh, sky.sh), run on the
 * CPU over host copies of the same buffers the GPU passes take. It is the
 * checker of the HIP kernels and the CPU baseline of the benchmark; nothing
 * in the product links or calls it. Every function cites the reference lines
 * it follows.
 *
 * All pointers in the rgi_* structs are HOST pointers here. Deterministic
 * only: the grid insert runs the fixed linearization of the reference's
 * atomics documented at rgi_grid_insert (include/restir_amd.h).
 */
#ifndef RESTIR_ORACLE_H
#define RESTIR_ORACLE_H

#include <stdint.h>

#include "../include/restir_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of OpenMP threads the oracle uses (0: the OpenMP default). */
void orc_set_threads(int threads);
int orc_get_threads(void);

/* For the oracle's own pins only: commit planted mistake `mistake` (see
 * enum Plant in oracle.cpp; 0 restores the faithful oracle), so a pin can
 * show that it fails on it. */
void orc_plant(int mistake);

/* traceScreenSpaceRay (restir.sh:364-436) of one ray over a G-buffer at the
 * frame's GI size; returns 1 on a hit and writes the hit surface point. */
int orc_screen_march(const rgi_frame* frame, const rgi_gbuffer* gbuffer, const float* origin3, const float* dir3,
                     float max_distance, float* hit3);

/* reconnectionJacobian (restir.sh:623-649): a sample point (position,
 * normal) shifted from receiver `there` to receiver `here`. */
float orc_reconnection_jacobian(const rgi_frame* frame, const float* here3, const float* there3,
                                const float* sample3, const float* normal3);

/* The pairwise MIS bookkeeping of restir.sh:789-931 for one sample point held
 * by the canonical reservoir and `taps` neighbours (rows of receiver
 * position, normal, count; a negative count is a tap with no reservoir):
 * the canonical share m_c(x) M_total, each neighbour's m_i(x) M_total, and
 * M_total. Visibility goes through `scene`'s BVH (frame ray mode BVH). */
void orc_gi_pairwise_weights(const rgi_frame* frame, const rgi_scene* scene, const float* receiver6,
                             float canonical_count, const float* sample6, int taps, const float* neighbours,
                             float* out_canonical_share, float* out_neighbour_mis, float* out_total);

/* Scalar primitives, for known-answer tests. */
uint32_t orc_pcg(uint32_t value);                       /* gridHash / pcg, hashgrid.sh:61-67 */
uint32_t orc_checksum_hash(uint32_t value);             /* gridChecksumHash, hashgrid.sh:69-78 */
float orc_next_random(uint32_t* state);                 /* nextRandom, restir.sh:131-135 */
float orc_grid_cell_size(const rgi_frame* frame, const float* position3, const float* camera3); /* hashgrid.sh:82-87 */
uint32_t orc_grid_normal_bucket(const float* normal3);  /* hashgrid.sh:89-103 */
void orc_grid_key(const rgi_frame* frame, const float* position3, const float* normal3, float cell_size,
                  uint32_t* slot, uint32_t* checksum);  /* hashgrid.sh:116-127 */
void orc_cosine_hemisphere(const float* normal3, float r0, float r1, float* out3); /* restir.sh:201-212 */
void orc_sincos(float x, float* s, float* c);
float orc_expf(float x);
float orc_pow64(float x);
float orc_exp2f(float x);
void orc_specular_albedo(const float* f0, float roughness, float nv, float* out);

/* The indirect specular pass of rgi_specular_gi (fs_specular.sc), and its
 * sampled direction with weight G1(l) for pins. */
void orc_specular_gi(const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* gbuffer,
                     const rgi_gi_inputs* in, const uint16_t* previous_hdr, const uint16_t* history, uint16_t* out);
void orc_specular_sample(const float* n3, const float* v3, float roughness, float r0, float r1, float* out_dir,
                         float* out_g1);

/* The composite pass of rgi_composite (fs_composite.sc), on the CPU. */
void orc_composite(const rgi_frame* frame, const rgi_gbuffer* gbuffer, const rgi_composite_inputs* inputs,
                   uint16_t* out_hdr);

/* The SVGF denoiser passes (fs_svgf_variance.sc, fs_svgf_atrous.sc) over
 * frame->width x frame->height half4 planes. */
void orc_svgf_variance(const rgi_frame* frame, const uint16_t* color, const uint16_t* moments,
                       const float* receiver, const float* normal, uint16_t* out);
void orc_svgf_atrous(const rgi_frame* frame, const uint16_t* input, const float* receiver, const float* normal,
                     int step, uint16_t* out);
/* evaluateBrdf (brdf.sh:38-62): diffuse and specular (each times n.l). */
void orc_brdf(const float* n3, const float* v3, const float* l3, const float* albedo3, float roughness,
              float metallic, float* diffuse3, float* specular3);
uint16_t orc_half_bits(float value);

/* Per-pixel keys of cs_grid_insert (cs_grid_insert.sc:40-73): slot, checksum
 * (0 for sky pixels, which insert nothing) and the jittered cell size. */
void orc_grid_keys(const rgi_frame* frame, const rgi_gbuffer* gbuffer, uint32_t* out_slot,
                   uint32_t* out_checksum, float* out_cell_size);

void orc_grid_clear(uint32_t* grid, uint32_t table_size);          /* cs_grid_clear.sc:12-21 */
void orc_grid_insert(const rgi_frame* frame, const rgi_gbuffer* gbuffer, uint32_t* grid); /* deterministic */
uint32_t orc_grid_pick_entry(const rgi_frame* frame, const uint32_t* grid, const float* camera3,
                             const float* position3, const float* normal3, float r0, float r1,
                             float r2);                               /* hashgrid.sh:155-177 */

/* traceBvh over rays of 8 floats (origin, tmax, direction, unused), as
 * rgi_trace. */
void orc_trace(const rgi_scene* scene, const float* rays, uint32_t count, int32_t* hit_triangle,
               float* hit_distance, int any_hit);
/* The same with the watertight triangle test (rgi_trace with
 * RGI_OPT_TRIANGLE_TEST = 1). */
void orc_trace_watertight(const rgi_scene* scene, const float* rays, uint32_t count, int32_t* hit_triangle,
                          float* hit_distance, int any_hit);

void orc_gi_temporal(const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* gbuffer,
                     const rgi_gi_inputs* inputs, const rgi_gi_reservoirs* out);
void orc_gi_spatial(const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* gbuffer,
                    const rgi_gi_inputs* inputs, const rgi_gi_reservoirs* reservoirs,
                    uint16_t* out_color, uint16_t* out_moments);
void orc_di_temporal(const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* gbuffer,
                     const rgi_di_reservoirs* previous, const rgi_di_reservoirs* out);
void orc_di_spatial(const rgi_frame* frame, const rgi_scene* scene, const rgi_gbuffer* gbuffer,
                    const rgi_di_reservoirs* reservoirs, uint16_t* out_diffuse, uint16_t* out_specular);

/* fs_reference.sc / cs_error.sc as rgi_reference_accumulate and
 * rgi_reference_error compute them, with the same FP64 reduction order. */
void orc_reference_accumulate(const rgi_gbuffer* gbuffer, const uint16_t* color, const uint16_t* di_diffuse,
                              const uint16_t* di_specular, double* accum);
void orc_reference_error(const rgi_gbuffer* gbuffer, const uint16_t* color, const uint16_t* di_diffuse,
                         const uint16_t* di_specular, const double* accum, uint32_t frames, double* out5);

/* The shadow pass of rgi_shadow_atlas (renderer.cpp:566-596, fs_shadow.sc),
 * on the CPU: one closest-hit ray per atlas texel. */
void orc_shadow_atlas(const rgi_shadow* shadow, const rgi_scene* scene, float* atlas);

/* The same two over the composite's HDR image. */
void orc_reference_accumulate_composite(const rgi_gbuffer* gbuffer, const uint16_t* hdr, double* accum);
void orc_reference_error_composite(const rgi_gbuffer* gbuffer, const uint16_t* hdr, const double* accum,
                                   uint32_t frames, double* out5);

/* The G-buffer producer of rgi_gbuffer_primary, on the CPU. */
void orc_gbuffer_primary(const rgi_frame* frame, const rgi_scene* scene, const float* materials,
                         uint32_t material_count, float* depth, uint32_t* albedo, uint32_t* normal,
                         uint32_t* shading_normal, uint16_t* emission);

/* The neural radiance cache of include/restir_amd_nrc.h (oracle/nrc.cpp).
 * res: the 16 level resolutions; lo/hi: the bounds; mlp: the logical weights
 * (RGI_NRC_MLP_PARAMS floats); table: [16][2^log2_table][2]; records and
 * targets as in the header. emulate = 1 rounds where the GPU rounds, 0
 * nowhere. */
void orc_nrc_forward(const uint32_t* res, uint32_t log2_table, const float* lo, const float* hi, const float* mlp,
                     const float* table, const float* records, uint32_t count, int emulate, double* out);
void orc_nrc_gradients(const uint32_t* res, uint32_t log2_table, const float* lo, const float* hi, const float* mlp,
                       const float* table, const float* records, const float* targets, uint32_t count, int emulate,
                       double* loss, double* mlp_grad, double* table_grad, double* denominators, int frozen);
void orc_nrc_adam(float* p, float* m, float* v, const float* g, uint32_t count, float lr, float beta1, float beta2,
                  float eps, float c1, float c2, int skip_zero);
void orc_nrc_corner(const uint32_t* res, uint32_t log2_table, int level, const float* p01, int k, uint32_t* idx,
                    float* w);
float orc_bf16_round(float x);

/* Floats (bit patterns lo, lo + stride, ... < hi, non-negative) for which the
 * kernels' constant division, fma(fma(-c, q, x), RN(1/c), q) with
 * q = x * RN(1/c), differs from IEEE x / c. */
uint64_t orc_div_const_mismatches(float c, uint32_t lo, uint32_t hi, uint32_t stride);

#ifdef __cplusplus
}
#endif

#endif
