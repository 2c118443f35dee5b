// Device-side building blocks of the gfx950 kernels: vector arithmetic with
// a fixed evaluation order, point-sampled plane fetches, the PCG stream and
// the per-launch constants. Compiled with FP contraction off, so every
// expression rounds exactly as written, identically on every lane and run.
//
// Definitions where GLSL leaves the result to the GPU (the reference's
// shaders run on whatever GPU bgfx picks):
//   dot = (x*x' + y*y') + z*z';  normalize(v) = v * (1 / sqrt(dot(v, v)));
//   inversesqrt(x) = 1 / sqrt(x);  min(a, b) = b < a ? b : a;
//   max(a, b) = a < b ? b : a;  mix(x, y, a) = x * (1 - a) + y * a;
//   sin/cos = rgi_sincos;  fp16 stores round to nearest even.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/restir_amd.h"
#include "rgi_math.h"

#define RGI_DEV __device__ __forceinline__

namespace rgi {

struct F2 { float x, y; };
struct F3 { float x, y, z; };
struct F4 { float x, y, z, w; };

// Workgroup -> screen tile.
// Order 0 is row-major in workgroup id.
// Order 1 hands each XCD a contiguous band of rows: the dispatcher places
// workgroup b on XCD b % 8, so numbering that XCD's workgroups b / 8 = 0, 1,
// ... through a band keeps the neighbourhood reads of concurrently running
// tiles in one XCD's L2 instead of spreading every region over all eight.
// Order 2 deals 4 x 4-tile blocks (64 x 64 pixels) to the XCDs in turn:
// block s (row-major over blocks) goes to XCD s % 8, and that XCD's
// workgroups walk its blocks tile by tile, so a block's tiles and most of
// their neighbourhood share one L2 while every region of the screen still
// spreads over all eight XCDs. Its grid is one-dimensional, 16 workgroups
// per block slot; slots past the screen edge have no tile.
// Every order visits every tile exactly once; false means this workgroup has
// no tile (order 2, or a batch slice's padding), and the whole workgroup
// leaves.
// A tile is 16 x 16 pixels, or 32 x 8 with `wide` (the same 256 threads).
struct TileMap {
    uint32_t order;
    uint32_t wide;
    uint32_t tiles_x, tiles_y;
    uint32_t blocks_x; // order 2: 4 x 4-tile blocks per row
};

// A workgroup's place in its view's grid: its linear index b and the grid's
// dimensions (the launch's own, or a slice of a batch launch).
struct BlockAt {
    uint32_t b, gx, gy;
};
RGI_DEV BlockAt this_block() { return BlockAt{blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y}; }

RGI_DEV bool block_tile(const TileMap& m, const BlockAt& at, uint32_t& tx, uint32_t& ty)
{
    uint32_t b = at.b;
    if (m.order == 2u) {
        uint32_t k = b >> 3, s = (b & 7u) + 8u * (k >> 4), j = k & 15u;
        tx = (s % m.blocks_x) * 4u + (j & 3u);
        ty = (s / m.blocks_x) * 4u + (j >> 2);
        return tx < m.tiles_x && ty < m.tiles_y;
    }
    // A batch pads each view's slice to whole rounds of 8 workgroups; the
    // padding has no tile.
    if (b >= at.gx * at.gy) return false;
    if (m.order == 1u) {
        uint32_t n = at.gx * at.gy, q = n >> 3, r = n & 7u, xcd = b & 7u, i = b >> 3;
        b = xcd < r ? xcd * (q + 1u) + i : r * (q + 1u) + (xcd - r) * q + i;
    }
    tx = b % at.gx;
    ty = b / at.gx;
    return true;
}
RGI_DEV bool block_tile(const TileMap& m, uint32_t& tx, uint32_t& ty) { return block_tile(m, this_block(), tx, ty); }

// Host side: the map and the grid of a width x height screen pass.
inline TileMap tile_map(int order, int wide, uint32_t width, uint32_t height)
{
    TileMap m;
    m.order = (uint32_t)order;
    m.wide = wide ? 1u : 0u;
    m.tiles_x = wide ? (width + 31u) / 32u : (width + 15u) / 16u;
    m.tiles_y = wide ? (height + 7u) / 8u : (height + 15u) / 16u;
    m.blocks_x = (m.tiles_x + 3u) / 4u;
    return m;
}

inline dim3 tile_grid(const TileMap& m)
{
    if (m.order != 2u) return dim3(m.tiles_x, m.tiles_y);
    uint32_t blocks = m.blocks_x * ((m.tiles_y + 3u) / 4u);
    return dim3(128u * ((blocks + 7u) / 8u));
}

// Lanes below this one whose bit is set in `mask` (a wave-64 ballot).
RGI_DEV uint32_t rank_in(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

RGI_DEV F3 f3(float x, float y, float z) { return F3{x, y, z}; }
RGI_DEV F3 splat3(float s) { return F3{s, s, s}; }
RGI_DEV F3 xyz(F4 v) { return F3{v.x, v.y, v.z}; }
RGI_DEV F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RGI_DEV F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RGI_DEV F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
RGI_DEV F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RGI_DEV F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
RGI_DEV F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
RGI_DEV F3 operator/(F3 a, F3 b) { return F3{a.x / b.x, a.y / b.y, a.z / b.z}; }
RGI_DEV float minf_(float a, float b) { return b < a ? b : a; }
RGI_DEV float maxf_(float a, float b) { return a < b ? b : a; }
RGI_DEV F3 min3(F3 a, F3 b) { return F3{minf_(a.x, b.x), minf_(a.y, b.y), minf_(a.z, b.z)}; }
RGI_DEV F3 max3(F3 a, F3 b) { return F3{maxf_(a.x, b.x), maxf_(a.y, b.y), maxf_(a.z, b.z)}; }
// Products of sums use explicit fused multiply-adds in a fixed order (x, y,
// z, then w); everything else is plain IEEE single precision, uncontracted.
RGI_DEV float dot(F3 a, F3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
RGI_DEV F3 cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RGI_DEV float length(F3 v) { return sqrtf(dot(v, v)); }
// Correctly rounded x / 255 and x / 3 in three instructions instead of the
// division's eleven: q = x * RN(1 / c), refined by one FMA step
// q + (x - c q) RN(1 / c). The same float as IEEE division for every float x
// and both constants (a sweep of all floats on the host; the oracle's
// orc_div_const_mismatches restates the check).
RGI_DEV float div255(float x)
{
    const float r = 1.0f / 255.0f;
    float q = x * r;
    return fmaf(fmaf(-255.0f, q, x), r, q);
}
RGI_DEV float div3(float x)
{
    const float r = 1.0f / 3.0f;
    float q = x * r;
    return fmaf(fmaf(-3.0f, q, x), r, q);
}
RGI_DEV float rsqrt_(float x) { return 1.0f / sqrtf(x); }
RGI_DEV F3 normalize(F3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
RGI_DEV float clampf_(float x, float lo, float hi) { return minf_(maxf_(x, lo), hi); }
RGI_DEV float mixf_(float x, float y, float a) { return x * (1.0f - a) + y * a; }
RGI_DEV F3 mix3(F3 x, F3 y, float a) { return x * (1.0f - a) + y * a; }
RGI_DEV float luminance(F3 c) { return dot(c, F3{0.2126f, 0.7152f, 0.0722f}); }

// int(x): truncation, saturating, NaN -> 0, written out so no out-of-range
// conversion is left undefined.
RGI_DEV int to_int(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (-2147483647 - 1);
    return (int)f;
}

constexpr float kPi = 3.14159265359f;
constexpr float kSkyDepth = 0.99999f;
constexpr float kSkySampleId = 30000000.0f;
constexpr uint32_t kGridCellCapacity = 32u;
constexpr uint32_t kGridMaxProbes = 8u;
constexpr uint32_t kGridInvalid = 0xffffffffu;

// mat4 * vec4, column-major.
RGI_DEV F4 mul(const float* m, F4 v)
{
    F4 r;
    r.x = fmaf(m[12], v.w, fmaf(m[8], v.z, fmaf(m[4], v.y, m[0] * v.x)));
    r.y = fmaf(m[13], v.w, fmaf(m[9], v.z, fmaf(m[5], v.y, m[1] * v.x)));
    r.z = fmaf(m[14], v.w, fmaf(m[10], v.z, fmaf(m[6], v.y, m[2] * v.x)));
    r.w = fmaf(m[15], v.w, fmaf(m[11], v.z, fmaf(m[7], v.y, m[3] * v.x)));
    return r;
}

// mat4 * (d, 0): the image of a direction.
RGI_DEV F4 mul_dir(const float* m, F3 d)
{
    return F4{fmaf(m[8], d.z, fmaf(m[4], d.y, m[0] * d.x)), fmaf(m[9], d.z, fmaf(m[5], d.y, m[1] * d.x)),
              fmaf(m[10], d.z, fmaf(m[6], d.y, m[2] * d.x)), fmaf(m[11], d.z, fmaf(m[7], d.y, m[3] * d.x))};
}

// Only the w row of mat4 * (p, 1): the view depth (fma(m15, 1, x) = x + m15).
RGI_DEV float mul_w(const float* m, F3 p) { return fmaf(m[11], p.z, fmaf(m[7], p.y, m[3] * p.x)) + m[15]; }

// ------------------------------------------------------------- plane fetch

// Point sampling with clamp: texel (clamp(floor(u * W)), clamp(floor(v * H))).
RGI_DEV int texel_coord(float u, int size)
{
    float f = floorf(u * (float)size);
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)(size - 1)) return size - 1;
    return (int)f;
}

RGI_DEV uint32_t texel_index(F2 uv, int width, int height)
{
    return (uint32_t)texel_coord(uv.y, height) * (uint32_t)width + (uint32_t)texel_coord(uv.x, width);
}

RGI_DEV F4 unorm8(uint32_t w)
{
    return F4{div255((float)(w & 255u)), div255((float)((w >> 8) & 255u)), div255((float)((w >> 16) & 255u)),
              div255((float)((w >> 24) & 255u))};
}

RGI_DEV F4 load_f4(const float* plane, uint32_t index)
{
    float4 v = reinterpret_cast<const float4*>(plane)[index];
    return F4{v.x, v.y, v.z, v.w};
}

RGI_DEV void store_f4(float* plane, uint32_t index, F4 v)
{
    reinterpret_cast<float4*>(plane)[index] = make_float4(v.x, v.y, v.z, v.w);
}

RGI_DEV float half_to_float(uint16_t bits)
{
    _Float16 h;
    __builtin_memcpy(&h, &bits, 2);
    return (float)h;
}

// float -> binary16 with integer operations (rgi_half_bits), not the
// hardware conversion: the backend fuses a multiply feeding a hardware f16
// conversion into v_fma_mix, which rounds once where the arithmetic asks for
// an f32 rounding first.
RGI_DEV uint16_t float_to_half(float v) { return rgi_half_bits(v); }

RGI_DEV F4 load_h4(const uint16_t* plane, uint32_t index)
{
    uint2 raw = reinterpret_cast<const uint2*>(plane)[index];
    return F4{half_to_float((uint16_t)(raw.x & 0xffffu)), half_to_float((uint16_t)(raw.x >> 16)),
              half_to_float((uint16_t)(raw.y & 0xffffu)), half_to_float((uint16_t)(raw.y >> 16))};
}

RGI_DEV void store_h4(uint16_t* plane, uint32_t index, F4 v)
{
    uint32_t lo = (uint32_t)float_to_half(v.x) | ((uint32_t)float_to_half(v.y) << 16);
    uint32_t hi = (uint32_t)float_to_half(v.z) | ((uint32_t)float_to_half(v.w) << 16);
    reinterpret_cast<uint2*>(plane)[index] = make_uint2(lo, hi);
}

// ------------------------------------------------------------------ random

RGI_DEV uint32_t pcg(uint32_t value)
{
    uint32_t state = value * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

RGI_DEV float next_random(uint32_t& state)
{
    state = pcg(state);
    return (float)(state >> 8u) * (1.0f / 16777216.0f);
}

RGI_DEV uint32_t checksum_hash(uint32_t v)
{
    v ^= v >> 15u;
    v *= 2246822519u;
    v ^= v >> 13u;
    v *= 3266489917u;
    v ^= v >> 16u;
    return v;
}

// ------------------------------------------------------------- projection

RGI_DEV float to_clip_depth(const rgi_frame& f, float d) { return f.homogeneous_depth ? d * 2.0f - 1.0f : d; }
RGI_DEV float to_clip_y(const rgi_frame& f, float y) { return f.origin_bottom_left ? y : -y; }

RGI_DEV F3 clip_to_world(const float* inv, F3 clip)
{
    F4 w = mul(inv, F4{clip.x, clip.y, clip.z, 1.0f});
    float r = 1.0f / w.w;
    return F3{w.x * r, w.y * r, w.z * r};
}

RGI_DEV F3 reconstruct_position(const rgi_frame& f, F2 uv, float depth)
{
    F3 clip = F3{uv.x * 2.0f - 1.0f, uv.y * 2.0f - 1.0f, to_clip_depth(f, depth)};
    clip.y = to_clip_y(f, clip.y);
    return clip_to_world(f.inv_view_proj, clip);
}

RGI_DEV F2 clip_to_uv(const rgi_frame& f, F4 clip)
{
    float sign = f.origin_bottom_left ? 1.0f : -1.0f;
    float r = 1.0f / clip.w;
    return F2{(clip.x * r) * 0.5f + 0.5f, ((clip.y * sign) * r) * 0.5f + 0.5f};
}

RGI_DEV bool on_screen(F2 uv) { return uv.x > 0.001f && uv.x < 0.999f && uv.y > 0.001f && uv.y < 0.999f; }

RGI_DEV bool project_to_screen(const rgi_frame& f, F3 p, F2& uv, float& view_depth)
{
    F4 clip = mul(f.view_proj, F4{p.x, p.y, p.z, 1.0f});
    view_depth = clip.w;
    uv = F2{0.0f, 0.0f};
    if (clip.w <= 0.0001f) return false;
    uv = clip_to_uv(f, clip);
    return on_screen(uv);
}

// Tangent frame shared by the grid jitter and the cosine lobe.
RGI_DEV void tangent_frame(F3 n, F3& t, F3& b)
{
    t = normalize(fabsf(n.y) < 0.99f ? cross(n, F3{0.0f, 1.0f, 0.0f}) : cross(n, F3{1.0f, 0.0f, 0.0f}));
    b = cross(n, t);
}

// ---------------------------------------------------------------- hash grid

// Cell size: the smallest size times 2^floor(log2(max(target / smallest, 1))),
// where floor(log2(m)) is read exactly from the binary exponent of m.
RGI_DEV float grid_cell_size(const rgi_frame& f, F3 camera, F3 p)
{
    float target = (length(p - camera) * f.grid_pixel_size) * f.grid_cell_pixels;
    float m = maxf_(target / f.grid_min_cell, 1.0f);
    int level = (int)((__float_as_uint(m) >> 23) & 0xffu) - 127;
    level = level < 0 ? 0 : (level > 127 ? 127 : level);
    return f.grid_min_cell * __uint_as_float((uint32_t)(level + 127) << 23);
}

RGI_DEV uint32_t grid_bucket(F3 n)
{
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    uint32_t bucket = n.z >= 0.0f ? 4u : 5u;
    if (ax >= ay && ax >= az) bucket = n.x >= 0.0f ? 0u : 1u;
    else if (ay >= az) bucket = n.y >= 0.0f ? 2u : 3u;
    return bucket;
}

RGI_DEV F3 grid_jitter(F3 p, F3 n, float cell, float r0, float r1)
{
    F3 t, b;
    tangent_frame(n, t, b);
    return p + (t * (r0 - 0.5f) + b * (r1 - 0.5f)) * cell;
}

RGI_DEV void grid_key(uint32_t mask, F3 p, F3 n, float cell, uint32_t& slot, uint32_t& checksum)
{
    uint32_t cx = (uint32_t)to_int(floorf(p.x / cell));
    uint32_t cy = (uint32_t)to_int(floorf(p.y / cell));
    uint32_t cz = (uint32_t)to_int(floorf(p.z / cell));
    uint32_t level = __float_as_uint(cell);
    uint32_t bucket = grid_bucket(n);
    slot = pcg(level + pcg(bucket + pcg(cx + pcg(cy + pcg(cz))))) & mask;
    uint32_t c = checksum_hash(level ^ checksum_hash(bucket ^ checksum_hash(cx ^ checksum_hash(cy ^ checksum_hash(cz + 0x9e3779b9u)))));
    checksum = c < 1u ? 1u : c;
}

RGI_DEV uint32_t grid_find(const uint32_t* grid, uint32_t mask, uint32_t slot0, uint32_t checksum)
{
    for (uint32_t probe = 0; probe < kGridMaxProbes; ++probe) {
        uint32_t s = (slot0 + probe) & mask;
        uint32_t stored = grid[s];
        if (stored == checksum) return s;
        if (stored == 0u) break;
    }
    return kGridInvalid;
}

RGI_DEV uint32_t grid_pick(const rgi_frame& f, const uint32_t* grid, F3 camera, F3 p, F3 n, float r0, float r1, float r2)
{
    uint32_t table = f.grid_table_size, mask = table - 1u;
    float cell = grid_cell_size(f, camera, p);
    F3 jittered = grid_jitter(p, n, cell, r0, r1);
    uint32_t slot, checksum;
    grid_key(mask, jittered, n, cell, slot, checksum);
    uint32_t found = grid_find(grid, mask, slot, checksum);
    if (found == kGridInvalid) return kGridInvalid;
    uint32_t count = grid[table + found];
    count = count > kGridCellCapacity ? kGridCellCapacity : count;
    if (count == 0u) return kGridInvalid;
    uint32_t index = (uint32_t)(r2 * (float)count);
    index = index > count - 1u ? count - 1u : index;
    return grid[2u * table + found * kGridCellCapacity + index];
}

// The jittered key of GI pixel (x, y) for the insert; false for the sky.
RGI_DEV bool insert_key(const rgi_frame& f, const rgi_gbuffer& g, F2 texel, int x, int y, uint32_t& slot,
                        uint32_t& checksum)
{
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    uint32_t t = texel_index(uv, g.width, g.height);
    float depth = g.depth[t];
    if (depth >= kSkyDepth) return false;
    F3 clip = F3{uv.x * 2.0f - 1.0f, uv.y * 2.0f - 1.0f, to_clip_depth(f, depth)};
    clip.y = to_clip_y(f, clip.y);
    F3 p = clip_to_world(f.inv_view_proj, clip);
    F4 nr = unorm8(g.normal[t]);
    F3 n = normalize(F3{nr.x * 2.0f - 1.0f, nr.y * 2.0f - 1.0f, nr.z * 2.0f - 1.0f});
    uint32_t rng = pcg((uint32_t)x + pcg((uint32_t)y + pcg(f.frame_index ^ 0x5bd1e995u)));
    float cell = grid_cell_size(f, F3{f.camera_position[0], f.camera_position[1], f.camera_position[2]}, p);
    float r0 = next_random(rng);
    float r1 = next_random(rng);
    grid_key(f.grid_table_size - 1u, grid_jitter(p, n, cell, r0, r1), n, cell, slot, checksum);
    return true;
}

} // namespace rgi
