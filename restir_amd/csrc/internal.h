// Library internals shared by the kernel files: the context (scratch
// buffers), error reporting and launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/restir_amd.h"
#include "../../include/restir_amd_nrc.h"

namespace rgi {
// Launch variants (RGI_OPT_*), all bit-exact with each other.
struct Tuning {
    int gi_temporal_waves = 0;
    int gi_spatial_waves = 5; // faster on every measured workload (RGI_OPT_GI_SPATIAL_WAVES)
    int trace_persistent = 0;
    int tile_order = 0;
    int tile_shape = 0; // 1: 32 x 8-pixel tiles (RGI_OPT_TILE_SHAPE)
    int watertight = 0; // rgi_trace's triangle test (RGI_OPT_TRIANGLE_TEST)
};
} // namespace rgi

struct rgi_context {
    int device = 0;
    rgi::Tuning tuning;
    uint32_t* counter = nullptr;   // persistent-launch work counter
    void* scratch = nullptr;       // grid insert (deterministic) and other per-call scratch
    size_t scratch_bytes = 0;
    // Buffers of single passes, apart from `scratch` (the renderer's grid
    // insert runs on a side stream beside GI temporal and the DI passes):
    // rgi::kBuf* name them.
    void* bufs[3] = {};
    size_t buf_bytes[3] = {};
    int compute_units = 0;
    // The light table and size that the guide buffer last held a guide for
    // (a renderer's lights are fixed for its lifetime, so it builds it once).
    const void* guide_lights = nullptr;
    int guide_count = 0;
    const void* guide_buffer = nullptr;
};

namespace rgi {

int fail(int code, const char* fmt, ...);
int check_launch(const char* what);
// Scratch of at least `bytes`, grown (after a device sync) when too small.
void* context_scratch(rgi_context* ctx, size_t bytes);
// A context buffer of at least `bytes`, grown (after a device sync) when too
// small; NULL (with the error set) when it cannot be allocated.
enum ContextBuffer { kBufTileList = 0, kBufCandidates = 1, kBufLightGuide = 2 };
void* context_buffer(rgi_context* ctx, ContextBuffer which, size_t bytes);
// The device's compute units, looked up once per context.
int context_compute_units(rgi_context* ctx);
struct CandRays;
// Points `c` (with c.pixels set) at the context's candidate records and rays
// for `candidates` candidates per pixel, grown when too small.
int context_cand_rays(rgi_context* ctx, uint32_t candidates, CandRays& c);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Kernel-file entry points, called by the C ABI after argument checks.
int grid_clear(uint32_t* grid, uint32_t table_size, hipStream_t stream);
int grid_insert_atomic(const rgi_frame& f, const rgi_gbuffer& g, uint32_t* grid, hipStream_t stream);
int grid_insert_deterministic(rgi_context* ctx, const rgi_frame& f, const rgi_gbuffer& g, uint32_t* grid,
                              hipStream_t stream);
// The deterministic insert's scratch: its size for a GI-pixel count and a
// table size, and the insert over caller-provided scratch of that size (every
// word of it is written by the insert's own kernels before it is read).
size_t det_scratch_bytes(uint32_t pixels, uint32_t table);
int grid_insert_deterministic_in(void* scratch_base, const rgi_frame& f, const rgi_gbuffer& g, uint32_t* grid,
                                 hipStream_t stream);
int trace_rays(const rgi_scene& s, const float* rays, uint32_t count, int32_t* hit, float* dist, int any_hit,
               const Tuning& tuning, uint32_t* counter, hipStream_t stream);
int gbuffer_primary(const rgi_frame& f, const rgi_scene& s, const float* materials, uint32_t material_count,
                    float* depth, uint32_t* albedo, uint32_t* normal, uint32_t* shading_normal, uint16_t* emission,
                    hipStream_t stream);
int shadow_atlas(const rgi_shadow& sh, const rgi_scene& s, float* atlas, hipStream_t stream);
int specular_gi(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                const uint16_t* previous_hdr, const uint16_t* history, uint16_t* out, const Tuning& tuning,
                hipStream_t stream);
int composite(const rgi_frame& f, const rgi_gbuffer& g, const rgi_composite_inputs& in, uint16_t* out,
              hipStream_t stream);
int svgf_variance(const rgi_frame& f, const uint16_t* color, const uint16_t* moments, const float* receiver,
                  const float* normal, uint16_t* out, hipStream_t stream);
int svgf_atrous(const rgi_frame& f, const uint16_t* input, const float* receiver, const float* normal, int step,
                uint16_t* out, hipStream_t stream);
// GI temporal. With a context (and at most 8 candidates) in two halves:
// gi_temporal_trace lists the surface tiles (writing the sky tiles' empty
// reservoirs), marches the candidates' rays and walks the open ones through
// the BVH, reading only the G-buffer and the scene; gi_temporal_resample
// then resamples the listed tiles. A caller may run other work between the
// two (the DI passes), on other streams. Without a context, one pass over
// every tile of the screen.
bool gi_temporal_split(const rgi_context* ctx, const rgi_frame& f);
int gi_temporal_trace(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                      const rgi_gi_reservoirs& out, hipStream_t stream);
int gi_temporal_resample(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                         const rgi_gi_inputs& in, const rgi_gi_reservoirs& out, const Tuning& tuning,
                         hipStream_t stream);
int gi_temporal(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                const rgi_gi_inputs& in, const rgi_gi_reservoirs& out, const Tuning& tuning, hipStream_t stream);
int gi_spatial(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
               const rgi_gi_reservoirs& res, uint16_t* color, uint16_t* moments, const Tuning& tuning,
               hipStream_t stream);
// With a context, scenes with more lights than fit in LDS search the CDF
// from a guide table that the context holds, built on every call unless
// `fixed_lights` (the caller's light table never changes, as a renderer's):
// then once per table.
int di_temporal(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                const rgi_di_reservoirs& prev, const rgi_di_reservoirs& out, const Tuning& tuning, hipStream_t stream,
                bool fixed_lights = false);
int di_spatial(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_di_reservoirs& res,
               uint16_t* diffuse, uint16_t* specular, const Tuning& tuning, hipStream_t stream);

// Batches (batch.h): each screen pass over every view of a batch in one
// launch. The guide table of DI temporal's light choice for frame f (no
// table, and 0 buckets, when the scene's lights fit in LDS or there is no
// context), the listed GI temporal's buffers, and the passes.
struct TileList;
struct BatchView;
struct BatchLaunch;
int di_light_guide(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, uint32_t** guide, uint32_t* buckets,
                   hipStream_t stream, bool fixed_lights = false);
int gi_listed_buffers(rgi_context* ctx, const rgi_frame& f, TileList& list, CandRays& c);
uint32_t grid_clear_blocks(uint32_t table_size);
int grid_clear_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, hipStream_t stream);
int grid_insert_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks,
                      hipStream_t stream);
int di_temporal_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks,
                      hipStream_t stream);
int di_spatial_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks,
                     hipStream_t stream);
int gi_trace_batch(const BatchLaunch& b, hipStream_t stream);
int gi_resample_batch(const BatchLaunch& b, hipStream_t stream);
int gi_spatial_batch(const BatchLaunch& b, hipStream_t stream);
// The denoiser (variance, then `iterations` a-trous steps) and the composite
// over every view of a batch; svgf_blocks is a view's denoiser grid.
uint32_t svgf_blocks(const rgi_frame& f);
int svgf_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, int iterations,
               hipStream_t stream);
int composite_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, hipStream_t stream);
// Indirect specular over every view of a batch (`same`, `larger`: views whose
// G-buffer has the GI size, or is larger).
int specular_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, bool same,
                   bool larger, hipStream_t stream);
// A batch of caches' training steps: nrc_step_plan writes the next step of
// cache n on `count` records into a kNrcStepBytes slot (growing the cache's
// scratch; nothing is launched) with each kernel's workgroups in blocks[];
// nrc_train_batch runs the steps of a device table of such slots, each
// kernel of rgi_nrc_train's chain once for all of them (first: per kernel,
// count + 1 slice starts; blocks: per kernel, the launch's workgroups), with
// results identical to rgi_nrc_train on each cache; nrc_step_taken then
// advances each cache's step count.
int nrc_step_plan(rgi_nrc* n, const float* records, const float* targets, uint32_t count, float* loss, void* slot,
                  uint32_t* blocks);
int nrc_train_batch(const void* steps, const uint32_t* first, uint32_t count, const uint32_t* blocks,
                    hipStream_t stream);
void nrc_step_taken(rgi_nrc* n);
// A batch of caches' G-buffer queries: nrc_query_plan writes the query of
// nrc_infer_gbuffer's arguments into a kNrcQueryBytes slot (nothing is
// launched), with its workgroups in *blocks; nrc_query_batch runs a device
// table of them, one launch per precision present (precisions[p] for
// RGI_NRC_BF16..FP8), each query on its own slice (first: count + 1 slice
// starts; blocks: their total), with results identical to
// nrc_infer_gbuffer's.
int nrc_query_plan(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* emission, int precision,
                   uint16_t* out, void* slot, uint32_t* blocks);
int nrc_query_batch(const void* queries, const uint32_t* first, uint32_t count, uint32_t blocks,
                    const bool* precisions, hipStream_t stream);
// A batch of caches' training-record gathers: nrc_gather_plan writes
// nrc_gather's arguments into a kNrcGatherBytes slot (growing the cache's
// surface list; nothing is launched), with the workgroups of its surface
// passes in blocks[0] and of its gather in blocks[1]; nrc_gather_batch runs
// a device table of them (first: count + 1 slice starts for the surface
// passes, then count + 1 for the gather; blocks: the two totals), with
// results identical to nrc_gather's.
int nrc_gather_plan(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* di_diffuse,
                    const uint16_t* color, const float* gi_sample, const float* gi_normal, uint32_t vertex_every,
                    uint32_t batch, uint32_t seed, float* records, float* targets, void* slot, uint32_t* blocks);
int nrc_gather_batch(const void* gathers, const uint32_t* first, uint32_t count, const uint32_t* blocks,
                     hipStream_t stream);

// The neural radiance cache's frame-driver passes (nrc.hip): the cache's
// radiance at every G-buffer pixel, and one frame's training records.
int nrc_infer_gbuffer(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* emission, int precision,
                      uint16_t* out, hipStream_t stream);
// With gi_sample and gi_normal, every vertex_every-th record is taken at
// the ReSTIR GI path vertex of its pixel (rgi_nrc_gather_vertices).
int nrc_gather(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* di_diffuse,
               const uint16_t* color, const float* gi_sample, const float* gi_normal, uint32_t vertex_every,
               uint32_t batch, uint32_t seed, float* records, float* targets, hipStream_t stream);

} // namespace rgi
