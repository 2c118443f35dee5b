// Batches of views in one launch per pass: the per-view descriptor that a
// batch's kernels read from device memory, and the map from a launch's
// workgroups to views.
#pragma once

#include "device.h"
#include "tiles.h"

namespace rgi {

constexpr uint32_t kMaxBatchViews = 64;

// Everything a batched pass reads of one view: the frames, scene and
// G-buffer, each pass's planes, launch maps and context buffers. The host
// fills one per view each frame and copies the array to the device before
// the frame's first batched launch.
struct BatchView {
    rgi_frame f;                    // DI, grid and GI tracing passes
    rgi_frame fg;                   // GI resampling passes (f with the cache's lighting)
    rgi_scene s;
    rgi_gbuffer g;
    F2 texel;                       // 1 / window size
    F2 gi_texel;                    // 1 / GI size
    // DI temporal and spatial.
    TileMap di_tiles;
    uint32_t di_gx, di_gy;          // the view's DI grid
    rgi_di_reservoirs di_prev, di_next;
    uint16_t* di_diffuse;
    uint16_t* di_specular;
    const uint32_t* guide;          // the light CDF's guide table, or NULL
    uint32_t guide_buckets;
    // The hash grid: clear, then the atomic insert.
    uint32_t* grid;
    uint32_t grid_gx, grid_gy;      // the insert's 16 x 16-pixel grid
    // GI temporal (listed form) and GI spatial.
    rgi_gi_inputs tin, sin;
    rgi_gi_reservoirs res;
    uint16_t* color;
    uint16_t* moments;
    TileMap gi_tiles;
    uint32_t gi_gx, gi_gy;          // GI spatial's grid
    TileList list;
    CandRays cand;
    uint32_t same;                  // the G-buffer has the GI size
    uint32_t split;                 // GI temporal in its listed form (list and cand set)
    // The denoiser: the variance into den_out[0], a-trous iteration it from
    // den_out[it] into den_out[it + 1] (planes of the GI size).
    uint32_t atrous;                // iterations, 0 without the denoiser
    uint16_t* den_out[6];
    // The composite into hdr (the window size).
    rgi_composite_inputs ci;
    uint16_t* hdr;
    // Indirect specular: last frame's composite and the history in, the
    // other history plane out (GI spatial's inputs and tile map).
    const uint16_t* spec_previous;
    const uint16_t* spec_history;
    uint16_t* spec_out;
};

// A cache's training step in a batch's table (nrc.hip fills it), and the
// kernels of the training chain that give each cache a slice of a launch.
constexpr size_t kNrcStepBytes = 512;
// A cache's query of one G-buffer plane in a batch's table (nrc.hip fills it).
constexpr size_t kNrcQueryBytes = 2048;
// A cache's training-record gather in a batch's table (nrc.hip fills it).
constexpr size_t kNrcGatherBytes = 1024;
enum NrcPass {
    kNrcTrain = 0, kNrcWgrad, kNrcReduce, kNrcAdamMlp, kNrcCopies, kNrcAdamTable, kNrcEmaMlp, kNrcEmaCopies,
    kNrcEmaTable, kNrcPasses
};

// The batched screen passes, each with its own slices of a launch.
enum BatchPass {
    kPassDi = 0, kPassGridClear, kPassGrid, kPassTiles, kPassGiSpatial, kPassDenoise, kPassComposite, kPassSpecular,
    kBatchPasses
};

// Host side: a batch frame's device table and the sizes of its launches.
struct BatchLaunch {
    const BatchView* views;              // the device copy of the table
    uint32_t count;
    const uint32_t* first[kBatchPasses]; // device: per pass, count + 1 slice starts
    uint32_t blocks[kBatchPasses];       // per pass, the workgroups of its launch
    bool same, larger;                   // views whose G-buffer has the GI size / is larger
    bool split, trace;                   // views with the listed GI temporal / with BVH walks
    uint32_t tiles, rounds;              // the listed views' GI tiles, and their rounds of 4 tiles
    int compute_units;
    int gi_temporal_waves, gi_spatial_waves;
};

// Workgroups [first[v], first[v + 1]) of a batched launch work on view v
// (first[0] = 0, nondecreasing, views entries after first[0]); returns v and
// the workgroup's index within the view's grid. Every lane of the workgroup
// gets the same answer.
RGI_DEV uint32_t batch_view(const uint32_t* first, uint32_t views, uint32_t block, uint32_t& local)
{
    uint32_t lo = 0u, hi = views - 1u; // the last view whose range starts at or before `block`
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (first[mid] <= block) lo = mid;
        else hi = mid - 1u;
    }
    local = block - first[lo];
    return lo;
}

} // namespace rgi
