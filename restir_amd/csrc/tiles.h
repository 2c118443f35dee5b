// The surface-tile lists of the GI passes and the temporal pass's candidate
// records: the device-side layouts that the listed kernels and their callers
// share.
#pragma once

#include "device.h"

namespace rgi {

// The surface tiles of a GI pass: 8 x 8-pixel tiles (two of the reference's
// 4 x 8 groups side by side) that hold at least one GI pixel with a surface,
// listed each pass so that only they are launched. words: [0] heavy tiles (at
// least 48 surface pixels), [1] light tiles, [2] the temporal pass's listed
// candidate rays, the tile claim counters at kTileClaimWord + 32 c and the
// ray claim counters at kWalkClaimWord + 32 c for c = 0..7 (one 128-byte line
// each), and the list from kTileListWord: heavy tiles from its front, light
// ones from its back, so that the heavy ones are claimed first.
struct TileList {
    uint32_t* words;
    uint32_t tiles_x;   // tiles per row
    uint32_t tiles;     // tiles on the screen: the list's capacity
};
constexpr uint32_t kTileRayWord = 2u, kTileClaimWord = 32u, kWalkClaimWord = 32u + 8u * 32u,
                   kTileListWord = 32u + 16u * 32u;
constexpr int kHeavyTile = 48;
// Tiles per workgroup of the listing pass (one wave each). The listing runs
// beside the DI passes: a 1024-thread workgroup waits for a CU with four free
// wave slots on every SIMD, while smaller workgroups fit beside the DI waves;
// but each workgroup appends to the list's two counters, so smaller ones
// contend on them. 512 threads balance the two.
constexpr uint32_t kListTilesPerGroup = 8u;

// The temporal pass's candidate records (the listed form): for candidate k
// and GI pixel o, rec[k * pixels + o] holds the march's hit point (w =
// kCandMarched) or, for a ray the march leaves open, (slot bits, 0, 0,
// kCandOpen), the ray being listed at `slot` for the BVH when the ray mode
// traces such rays. Listed ray i is rays[2 i] (origin, max distance) and
// rays[2 i + 1] (direction); its closest hit goes to hit[i] and dist[i].
struct CandRays {
    float4* rec;
    float4* rays;
    int32_t* hit;
    float* dist;
    uint32_t pixels;
};
constexpr int kMaxCandRecords = 8;

// This lane's pixel of tile t: lanes [0, 32) the left 4 x 8 group, [32, 64)
// the right one, each row-major.
RGI_DEV void tile_pixel(const TileList& l, uint32_t t, int& x, int& y)
{
    uint32_t lane = threadIdx.x & 63u;
    x = (int)((t % l.tiles_x) * 8u + (lane >> 5) * 4u + (lane & 3u));
    y = (int)((t / l.tiles_x) * 8u + ((lane & 31u) >> 2));
}

// Listed item i (below the list's count): the heavy tiles, then the light.
RGI_DEV uint32_t listed_tile(const TileList& l, uint32_t i)
{
    uint32_t heavy = l.words[0];
    return i < heavy ? l.words[kTileListWord + i] : l.words[kTileListWord + l.tiles - 1u - (i - heavy)];
}

// The next listed tile for this wave (the result is wave-uniform). Workgroup
// b, which runs on XCD b mod 8, claims from counter c = b mod 8 the items
// c, c + 8, c + 16, ...: eight counters spread the claims of the chip's
// waves, and each XCD still takes every eighth tile of the heavy-first order.
RGI_DEV bool claim_tile(const TileList& l, uint32_t& t)
{
    const uint32_t c = blockIdx.x & 7u;
    uint32_t k = 0u;
    if ((threadIdx.x & 63u) == 0u) k = atomicAdd(l.words + kTileClaimWord + 32u * c, 1u);
    uint32_t item = (uint32_t)__shfl((int)k, 0) * 8u + c;
    if (item >= l.words[0] + l.words[1]) return false;
    t = listed_tile(l, item);
    return true;
}

} // namespace rgi
