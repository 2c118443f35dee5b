// Host oracle of the ReSTIR GI/DI hot path — TEST INFRASTRUCTURE ONLY (the
// checker of the HIP kernels and the benchmark's CPU baseline; nothing in
// the product links it). A scalar restatement of the reference's shaders,
// one function per reference function, each citing the lines it follows.
//
// Arithmetic is written in the evaluation order GLSL gives the reference's
// expressions (left to right, no contraction: build with -ffp-contract=off),
// with these definitions where GLSL leaves the result to the GPU:
//   dot(a, b)       = (a.x * b.x + a.y * b.y) + a.z * b.z
//   normalize(v)    = v * (1 / sqrt(dot(v, v)))
//   inversesqrt(x)  = 1 / sqrt(x);  length(v) = sqrt(dot(v, v))
//   min(a, b)       = b < a ? b : a;  max(a, b) = a < b ? b : a
//   mix(x, y, a)    = x * (1 - a) + y * a
//   pow(x, 5)       = ((x * x) * (x * x)) * x
//   floor(log2(m))  = the binary exponent of m (m >= 1), exactly
//   exp2(k)         = 2^k, exactly, for the integer k it is used with
//   sin, cos        = rgi_sincos (restir_amd/csrc/rgi_math.h)
//   int(x)          = truncation, saturating, NaN -> 0
//   RGBA16F stores  = round to nearest even (rgi_half_bits)
#include "oracle.h"

#include <math.h>
#include <string.h>

#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../restir_amd/csrc/rgi_math.h"

namespace {

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 v3s(float s) { return V3{s, s, s}; }
inline V3 xyz(const V4& v) { return V3{v.x, v.y, v.z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return a < b ? b : a; }
inline V3 vmin(V3 a, V3 b) { return V3{fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return V3{fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
// Products of sums are written with explicit fused multiply-adds in a fixed
// order (x, then y, then z, then w); all other arithmetic is plain IEEE
// single precision, uncontracted.
inline float dot(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
inline V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(V3 v) { return sqrtf(dot(v, v)); }
inline float inversesqrt(float x) { return 1.0f / sqrtf(x); }
inline V3 normalize(V3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
inline float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline V3 mix3(V3 x, V3 y, float a) { return x * (1.0f - a) + y * a; }
inline float pow5(float x) { return ((x * x) * (x * x)) * x; }
inline int to_int(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (-2147483647 - 1);
    return (int)f;
}

const float kPi = 3.14159265359f;
const float kSkyDepth = 0.99999f;

// mul(mat4, vec4) with a column-major matrix.
V4 mul(const float* m, V4 v)
{
    V4 r;
    r.x = fmaf(m[12], v.w, fmaf(m[8], v.z, fmaf(m[4], v.y, m[0] * v.x)));
    r.y = fmaf(m[13], v.w, fmaf(m[9], v.z, fmaf(m[5], v.y, m[1] * v.x)));
    r.z = fmaf(m[14], v.w, fmaf(m[10], v.z, fmaf(m[6], v.y, m[2] * v.x)));
    r.w = fmaf(m[15], v.w, fmaf(m[11], v.z, fmaf(m[7], v.y, m[3] * v.x)));
    return r;
}

// mul(mat4, vec4(d, 0)): the image of a direction.
V4 mul_dir(const float* m, V3 d)
{
    return V4{fmaf(m[8], d.z, fmaf(m[4], d.y, m[0] * d.x)), fmaf(m[9], d.z, fmaf(m[5], d.y, m[1] * d.x)),
              fmaf(m[10], d.z, fmaf(m[6], d.y, m[2] * d.x)), fmaf(m[11], d.z, fmaf(m[7], d.y, m[3] * d.x))};
}

// ---------------------------------------------------------------- textures

int texel_coord(float u, int size)
{
    float f = floorf(u * (float)size);
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)(size - 1)) return size - 1;
    return (int)f;
}

size_t texel(V2 uv, int width, int height)
{
    return (size_t)texel_coord(uv.y, height) * (size_t)width + (size_t)texel_coord(uv.x, width);
}

V4 unorm8(uint32_t w)
{
    return V4{(float)(w & 255u) / 255.0f, (float)((w >> 8) & 255u) / 255.0f,
              (float)((w >> 16) & 255u) / 255.0f, (float)((w >> 24) & 255u) / 255.0f};
}

V4 half4(const uint16_t* p)
{
    return V4{rgi_half_float(p[0]), rgi_half_float(p[1]), rgi_half_float(p[2]), rgi_half_float(p[3])};
}

V4 float4(const float* p) { return V4{p[0], p[1], p[2], p[3]}; }

void store_half4(uint16_t* p, V4 v)
{
    p[0] = rgi_half_bits(v.x);
    p[1] = rgi_half_bits(v.y);
    p[2] = rgi_half_bits(v.z);
    p[3] = rgi_half_bits(v.w);
}

void store_float4(float* p, V4 v)
{
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
    p[3] = v.w;
}

// ------------------------------------------------------------ random numbers

// pcg / gridHash (restir.sh:122-127, hashgrid.sh:61-67, cs_grid_insert.sc:27-32).
uint32_t pcg(uint32_t value)
{
    uint32_t state = value * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// nextRandom (restir.sh:131-135).
float next_random(uint32_t& state)
{
    state = pcg(state);
    return (float)(state >> 8u) * (1.0f / 16777216.0f);
}

// gridChecksumHash (hashgrid.sh:69-78).
uint32_t checksum_hash(uint32_t value)
{
    value ^= value >> 15u;
    value *= 2246822519u;
    value ^= value >> 13u;
    value *= 3266489917u;
    value ^= value >> 16u;
    return value;
}

// ------------------------------------------------------------- planted mistakes

// Test infrastructure for the oracle's own pins: orc_plant(k) makes the
// oracle commit one plausible mistake (0: none), so that a pin can show it
// fails on it. Never set outside those tests.
enum Plant {
    kPlantNone = 0,
    kPlantNoJacobian = 1,    // reconnectionJacobian taken as 1
    kPlantMisHere = 2,       // a neighbour's pairwise density evaluated at this receiver
    kPlantPdfAtOrigin = 3,   // candidate pdf left at the ray origin (no conversion to the receiver)
    kPlantThickness = 4,     // march thickness ten times the reference's
    kPlantBisection = 5,     // two bisection rounds instead of five
    kPlantDiTemporalMis = 6, // DI temporal: canonical sample without its MIS weight
    kPlantDiSpatialMis = 7,  // DI spatial: canonical takes every pair's whole share
    kPlantSkipShare = 8,     // a skipped GI tap does not hand its pair share to the canonical sample
};
static int g_plant = kPlantNone;

// ------------------------------------------------------------- pass context

struct Pass {
    const rgi_frame* f;
    const rgi_scene* s;
    const rgi_gbuffer* g;
    V2 texel_size;   // u_screenInfo.zw / u_screenSize.zw
    V2 size;         // u_screenInfo.xy
    int gi_w, gi_h;  // GI-size planes (reservoirs, history): u_screenInfo.xy
    V3 grid_camera;  // u_gridCamera of this pass
    const uint32_t* grid;
    const rgi_gi_inputs* in;
    uint32_t grid_mask;
};

float scene_scale(const Pass& p) { return p.f->scene_scale; }

// toClipSpaceDepth / toClipSpaceY (bgfx shaderlib), clipToWorld.
float to_clip_depth(const rgi_frame* f, float d) { return f->homogeneous_depth ? d * 2.0f - 1.0f : d; }
float to_clip_y(const rgi_frame* f, float y) { return f->origin_bottom_left ? y : -y; }

// The perspective divide is one reciprocal and a product per component.
V3 clip_to_world(const float* inv, V3 clip)
{
    V4 w = mul(inv, V4{clip.x, clip.y, clip.z, 1.0f});
    float r = 1.0f / w.w;
    return V3{w.x * r, w.y * r, w.z * r};
}

// reconstructWorldPosition (restir.sh:150-155, di.sh:71-76, cs_grid_insert.sc:58-60).
V3 reconstruct_world_position(const rgi_frame* f, V2 uv, float depth)
{
    V3 clip = V3{uv.x * 2.0f - 1.0f, uv.y * 2.0f - 1.0f, to_clip_depth(f, depth)};
    clip.y = to_clip_y(f, clip.y);
    return clip_to_world(f->inv_view_proj, clip);
}

// clipToUv (restir.sh:157-161).
V2 clip_to_uv(const rgi_frame* f, V4 clip)
{
    float origin_sign = f->origin_bottom_left ? 1.0f : -1.0f;
    float r = 1.0f / clip.w;
    return V2{(clip.x * r) * 0.5f + 0.5f, ((clip.y * origin_sign) * r) * 0.5f + 0.5f};
}

// isOnScreen (restir.sh:163-166).
bool is_on_screen(V2 uv) { return uv.x > 0.001f && uv.x < 0.999f && uv.y > 0.001f && uv.y < 0.999f; }

// projectToScreen (restir.sh:187-199).
bool project_to_screen(const rgi_frame* f, V3 position, V2& uv, float& view_depth)
{
    V4 clip = mul(f->view_proj, V4{position.x, position.y, position.z, 1.0f});
    view_depth = clip.w;
    uv = V2{0.0f, 0.0f};
    if (clip.w <= 0.0001f) return false;
    uv = clip_to_uv(f, clip);
    return is_on_screen(uv);
}

float fetch_depth(const Pass& p, V2 uv) { return p.g->depth[texel(uv, p.g->width, p.g->height)]; }
V4 fetch_normal_raw(const Pass& p, V2 uv) { return unorm8(p.g->normal[texel(uv, p.g->width, p.g->height)]); }

// decodeNormal (restir.sh:145-148).
V3 decode_normal(const Pass& p, V2 uv)
{
    V4 n = fetch_normal_raw(p, uv);
    return normalize(V3{n.x * 2.0f - 1.0f, n.y * 2.0f - 1.0f, n.z * 2.0f - 1.0f});
}

// diffuseAlbedo (restir.sh:138-143).
V3 diffuse_albedo(const Pass& p, V2 uv)
{
    V4 a = unorm8(p.g->albedo[texel(uv, p.g->width, p.g->height)]);
    float metallic = fetch_normal_raw(p, uv).w;
    return V3{a.x, a.y, a.z} * (1.0f - metallic);
}

// ---------------------------------------------------------------- hash grid

// gridCellSize (hashgrid.sh:82-87).
float grid_cell_size(const rgi_frame* f, V3 camera, V3 position)
{
    float target = (length(position - camera) * f->grid_pixel_size) * f->grid_cell_pixels;
    float m = fmax_(target / f->grid_min_cell, 1.0f);
    int level = (int)((rgi_float_bits(m) >> 23) & 0xffu) - 127;
    if (level < 0) level = 0;
    if (level > 127) level = 127;
    return f->grid_min_cell * rgi_bits_float((uint32_t)(level + 127) << 23);
}

// gridNormalBucket (hashgrid.sh:89-103).
uint32_t grid_normal_bucket(V3 n)
{
    V3 a = V3{fabsf(n.x), fabsf(n.y), fabsf(n.z)};
    uint32_t bucket = n.z >= 0.0f ? 4u : 5u;
    if (a.x >= a.y && a.x >= a.z) bucket = n.x >= 0.0f ? 0u : 1u;
    else if (a.y >= a.z) bucket = n.y >= 0.0f ? 2u : 3u;
    return bucket;
}

// The tangent frame of gridJitter and cosineHemisphere (hashgrid.sh:108-111, restir.sh:207-210).
void tangent_frame(V3 n, V3& tangent, V3& bitangent)
{
    tangent = normalize(fabsf(n.y) < 0.99f ? cross(n, v3(0.0f, 1.0f, 0.0f)) : cross(n, v3(1.0f, 0.0f, 0.0f)));
    bitangent = cross(n, tangent);
}

// gridJitter (hashgrid.sh:106-113).
V3 grid_jitter(V3 position, V3 n, float cell_size, float r0, float r1)
{
    V3 t, b;
    tangent_frame(n, t, b);
    return position + (t * (r0 - 0.5f) + b * (r1 - 0.5f)) * cell_size;
}

// gridKey (hashgrid.sh:116-127).
void grid_key(uint32_t mask, V3 position, V3 n, float cell_size, uint32_t& slot, uint32_t& checksum)
{
    int cx = to_int(floorf(position.x / cell_size));
    int cy = to_int(floorf(position.y / cell_size));
    int cz = to_int(floorf(position.z / cell_size));
    uint32_t level = rgi_float_bits(cell_size);
    uint32_t bucket = grid_normal_bucket(n);
    slot = pcg(level + pcg(bucket + pcg((uint32_t)cx + pcg((uint32_t)cy + pcg((uint32_t)cz))))) & mask;
    uint32_t c = checksum_hash(level ^ checksum_hash(bucket ^ checksum_hash((uint32_t)cx ^
                 checksum_hash((uint32_t)cy ^ checksum_hash((uint32_t)cz + 0x9e3779b9u)))));
    checksum = c < 1u ? 1u : c;
}

const uint32_t kGridCellCapacity = 32u;
const uint32_t kGridMaxProbes = 8u;
const uint32_t kGridInvalid = 0xffffffffu;

// gridFindCell (hashgrid.sh:132-150).
uint32_t grid_find_cell(const uint32_t* grid, uint32_t mask, uint32_t slot0, uint32_t checksum)
{
    for (uint32_t probe = 0; probe < kGridMaxProbes; ++probe) {
        uint32_t slot = (slot0 + probe) & mask;
        uint32_t stored = grid[slot];
        if (stored == checksum) return slot;
        if (stored == 0u) break;
    }
    return kGridInvalid;
}

// gridPickEntry (hashgrid.sh:155-177).
uint32_t grid_pick_entry(const rgi_frame* f, const uint32_t* grid, V3 camera, V3 position, V3 n, float r0,
                         float r1, float r2)
{
    uint32_t table = f->grid_table_size, mask = table - 1u;
    float cell_size = grid_cell_size(f, camera, position);
    V3 jittered = grid_jitter(position, n, cell_size, r0, r1);
    uint32_t slot, checksum;
    grid_key(mask, jittered, n, cell_size, slot, checksum);
    uint32_t cell = grid_find_cell(grid, mask, slot, checksum);
    if (cell == kGridInvalid) return kGridInvalid;
    uint32_t count = grid[table + cell];
    if (count > kGridCellCapacity) count = kGridCellCapacity;
    if (count == 0u) return kGridInvalid;
    uint32_t index = (uint32_t)(r2 * (float)count);
    if (index > count - 1u) index = count - 1u;
    return grid[2u * table + cell * kGridCellCapacity + index];
}

// Key of GI pixel (x, y) in cs_grid_insert (cs_grid_insert.sc:40-73); false for the sky.
bool insert_key(const rgi_frame* f, const rgi_gbuffer* g, int x, int y, uint32_t& slot, uint32_t& checksum,
                float& cell_size)
{
    V2 ts = V2{1.0f / (float)f->gi_width, 1.0f / (float)f->gi_height}; // u_screenInfo.zw
    V2 uv = V2{((float)x + 0.5f) * ts.x, ((float)y + 0.5f) * ts.y};
    float depth = g->depth[texel(uv, g->width, g->height)];
    if (depth >= kSkyDepth) return false;
    V3 clip = V3{uv.x * 2.0f - 1.0f, uv.y * 2.0f - 1.0f, to_clip_depth(f, depth)};
    clip.y = to_clip_y(f, clip.y);
    V3 position = clip_to_world(f->inv_view_proj, clip);
    V4 nr = unorm8(g->normal[texel(uv, g->width, g->height)]);
    V3 n = normalize(V3{nr.x * 2.0f - 1.0f, nr.y * 2.0f - 1.0f, nr.z * 2.0f - 1.0f});
    uint32_t rng = pcg((uint32_t)x + pcg((uint32_t)y + pcg(f->frame_index ^ 0x5bd1e995u)));
    cell_size = grid_cell_size(f, xyz(V4{f->camera_position[0], f->camera_position[1], f->camera_position[2], 0}),
                               position);
    float r0 = next_random(rng);
    float r1 = next_random(rng);
    V3 jittered = grid_jitter(position, n, cell_size, r0, r1);
    grid_key(f->grid_table_size - 1u, jittered, n, cell_size, slot, checksum);
    return true;
}

// -------------------------------------------------------------------- BVH

const int kBvhShortStack = 8;
const int kBvhMaxSteps = 2048;

V4 scene_vec4(const rgi_scene* s, size_t index) { return float4(s->bvh + index * 4); }
bool bvh_available(const rgi_scene* s) { return s->node_count > 0 && s->bvh != nullptr; }
V3 bvh_vertex(const rgi_scene* s, int tri, int corner)
{
    return xyz(scene_vec4(s, (size_t)s->triangle_offset + (size_t)tri * 4 + (size_t)corner));
}

// intersectBox (bvh.sh:32-41).
float intersect_box(V3 origin, V3 inv_dir, V3 mn, V3 mx, float max_distance)
{
    V3 t0 = (mn - origin) * inv_dir;
    V3 t1 = (mx - origin) * inv_dir;
    V3 n = vmin(t0, t1);
    V3 fa = vmax(t0, t1);
    float enter = fmax_(fmax_(n.x, n.y), fmax_(n.z, 0.0f));
    float leave = fmin_(fmin_(fa.x, fa.y), fmin_(fa.z, max_distance));
    return enter <= leave ? enter : -1.0f;
}

// The slab test made conservative for the watertight mode (Pharr, Jakob and
// Humphreys, PBR 3rd ed., 3.9.2): far distances scaled by 1 + 2 gamma(3);
// the twin of the kernels' box_enter_conservative, not a reference function.
float intersect_box_conservative(V3 origin, V3 inv_dir, V3 mn, V3 mx, float max_distance)
{
    const float grow = 1.0f + 2.0f * (3.0f * 0x1p-24f / (1.0f - 3.0f * 0x1p-24f));
    V3 t0 = (mn - origin) * inv_dir;
    V3 t1 = (mx - origin) * inv_dir;
    V3 n = vmin(t0, t1);
    V3 fa = vmax(t0, t1);
    float enter = fmax_(fmax_(n.x, n.y), fmax_(n.z, 0.0f));
    float leave = fmin_(fmin_(fa.x * grow, fa.y * grow), fmin_(fa.z * grow, max_distance));
    return enter <= leave ? enter : -1.0f;
}

float box_test(bool watertight, V3 origin, V3 inv_dir, V3 mn, V3 mx, float max_distance)
{
    return watertight ? intersect_box_conservative(origin, inv_dir, mn, mx, max_distance)
                      : intersect_box(origin, inv_dir, mn, mx, max_distance);
}

// intersectTriangle: Moeller-Trumbore, both sides (bvh.sh:44-71).
float intersect_triangle(V3 o, V3 d, V3 v0, V3 v1, V3 v2)
{
    V3 e1 = v1 - v0;
    V3 e2 = v2 - v0;
    V3 pv = cross(d, e2);
    float det = dot(e1, pv);
    if (fabsf(det) < 1e-12f) return -1.0f;
    float inv = 1.0f / det;
    V3 sv = o - v0;
    float u = dot(sv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    V3 q = cross(sv, e1);
    float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    return dot(e2, q) * inv;
}

// The watertight ray/triangle test of Woop, Benthin and Wald 2013 (JCGT
// 2(1), 65-82), both sides, in the operation order of the kernels'
// watertight_hit (restir_amd/csrc/bvh_device.h): not a reference function, the
// twin of rgi_trace's RGI_OPT_TRIANGLE_TEST = 1.
float axis_of(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

float intersect_triangle_watertight(V3 o, V3 d, V3 v0, V3 v1, V3 v2)
{
    float adx = fabsf(d.x), ady = fabsf(d.y), adz = fabsf(d.z);
    int kz = adx >= ady ? (adx >= adz ? 0 : 2) : (ady >= adz ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = axis_of(d, kz);
    if (dz < 0.0f) std::swap(kx, ky);
    float sx = axis_of(d, kx) / dz, sy = axis_of(d, ky) / dz, sz = 1.0f / dz;
    V3 a = v0 - o, b = v1 - o, c = v2 - o;
    float az = axis_of(a, kz), bz = axis_of(b, kz), cz = axis_of(c, kz);
    float ax = axis_of(a, kx) - sx * az, ay = axis_of(a, ky) - sy * az;
    float bx = axis_of(b, kx) - sx * bz, by = axis_of(b, ky) - sy * bz;
    float cx = axis_of(c, kx) - sx * cz, cy = axis_of(c, ky) - sy * cz;
    float u = cx * by - cy * bx;
    float v = ax * cy - ay * cx;
    float w = bx * ay - by * ax;
    if (u == 0.0f || v == 0.0f || w == 0.0f) {
        u = (float)((double)cx * (double)by - (double)cy * (double)bx);
        v = (float)((double)ax * (double)cy - (double)ay * (double)cx);
        w = (float)((double)bx * (double)ay - (double)by * (double)ax);
    }
    if ((u < 0.0f || v < 0.0f || w < 0.0f) && (u > 0.0f || v > 0.0f || w > 0.0f)) return -1.0f;
    float det = (u + v) + w;
    if (det == 0.0f) return -1.0f;
    return ((u * (sz * az) + v * (sz * bz)) + w * (sz * cz)) / det;
}

// bvhInverseDirection (bvh.sh:113-121).
float safe_component(float d) { return fabsf(d) > 1e-8f ? d : (d >= 0.0f ? 1e-8f : -1e-8f); }
V3 inverse_direction(V3 d)
{
    return V3{1.0f / safe_component(d.x), 1.0f / safe_component(d.y), 1.0f / safe_component(d.z)};
}

// bvhIntersectLeaf (bvh.sh:124-138).
int intersect_leaf(const rgi_scene* s, V3 o, V3 d, int first, int count, float& distance, int hit,
                   bool watertight = false)
{
    for (int tri = first; tri < first + count; ++tri) {
        float t = watertight ? intersect_triangle_watertight(o, d, bvh_vertex(s, tri, 0), bvh_vertex(s, tri, 1),
                                                             bvh_vertex(s, tri, 2))
                             : intersect_triangle(o, d, bvh_vertex(s, tri, 0), bvh_vertex(s, tri, 1),
                                                  bvh_vertex(s, tri, 2));
        if (t > 0.0f && t < distance) {
            distance = t;
            hit = tri;
        }
    }
    return hit;
}

// traceBvhStackless (bvh.sh:140-174).
int trace_stackless(const rgi_scene* s, V3 o, V3 d, bool any_hit, float& distance, bool watertight = false)
{
    V3 inv = inverse_direction(d);
    int node_count = (int)s->node_count;
    int hit = -1;
    int node = 0;
    for (int step = 0; step < kBvhMaxSteps; ++step) {
        if (node >= node_count) break;
        V4 n0 = scene_vec4(s, (size_t)node * 2);
        V4 n1 = scene_vec4(s, (size_t)node * 2 + 1);
        bool leaf = n1.w >= 0.0f;
        if (box_test(watertight, o, inv, xyz(n0), xyz(n1), distance) < 0.0f) {
            node = leaf ? node + 1 : to_int(-n1.w) - 1;
            continue;
        }
        if (leaf) {
            hit = intersect_leaf(s, o, d, to_int(n0.w), to_int(n1.w), distance, hit, watertight);
            if (any_hit && hit >= 0) return hit;
        }
        node = node + 1;
    }
    return hit;
}

// traceBvhSharedStack (bvh.sh:216-295), with this thread's 8-entry ring.
int trace_short_stack(const rgi_scene* s, V3 o, V3 d, float& distance, bool watertight = false)
{
    V3 inv = inverse_direction(d);
    if (box_test(watertight, o, inv, xyz(scene_vec4(s, 0)), xyz(scene_vec4(s, 1)), distance) < 0.0f) return -1;
    int stack[kBvhShortStack];
    int top = 0, held = 0;
    bool dropped = false;
    int hit = -1;
    int node = 0;
    for (int step = 0; step < kBvhMaxSteps; ++step) {
        V4 n0 = scene_vec4(s, (size_t)node * 2);
        V4 n1 = scene_vec4(s, (size_t)node * 2 + 1);
        if (n1.w >= 0.0f) {
            hit = intersect_leaf(s, o, d, to_int(n0.w), to_int(n1.w), distance, hit, watertight);
        } else {
            int left = node + 1;
            int right = to_int(n0.w);
            float ld = box_test(watertight, o, inv, xyz(scene_vec4(s, (size_t)left * 2)),
                                xyz(scene_vec4(s, (size_t)left * 2 + 1)), distance);
            float rd = box_test(watertight, o, inv, xyz(scene_vec4(s, (size_t)right * 2)),
                                xyz(scene_vec4(s, (size_t)right * 2 + 1)), distance);
            bool vl = ld >= 0.0f, vr = rd >= 0.0f;
            if (vl && vr) {
                bool left_first = ld <= rd;
                stack[top & (kBvhShortStack - 1)] = left_first ? right : left;
                ++top;
                if (held == kBvhShortStack) dropped = true;
                else ++held;
                node = left_first ? left : right;
                continue;
            }
            if (vl || vr) {
                node = vl ? left : right;
                continue;
            }
        }
        if (held == 0) break;
        --top;
        --held;
        node = stack[top & (kBvhShortStack - 1)];
    }
    if (dropped) {
        int again = trace_stackless(s, o, d, false, distance, watertight);
        if (again >= 0) hit = again;
    }
    return hit;
}

// traceBvh (bvh.sh:298-319): closest hit through the short stack (as the GI
// temporal pass compiles it), any hit stackless.
int trace_bvh(const rgi_scene* s, V3 o, V3 d, float max_distance, bool any_hit, float& distance,
              bool watertight = false)
{
    distance = max_distance;
    if (!bvh_available(s)) return -1;
    if (!any_hit) return trace_short_stack(s, o, d, distance, watertight);
    return trace_stackless(s, o, d, true, distance, watertight);
}

V3 bvh_triangle_albedo(const rgi_scene* s, int tri)
{
    return xyz(scene_vec4(s, (size_t)s->triangle_offset + (size_t)tri * 4 + 3));
}

V3 bvh_triangle_emission(const rgi_scene* s, int tri)
{
    size_t base = (size_t)s->triangle_offset + (size_t)tri * 4;
    return V3{scene_vec4(s, base + 1).w, scene_vec4(s, base + 2).w, scene_vec4(s, base + 3).w};
}

// ------------------------------------------------------------------ shadow

// shadowFace (shadow.sh:31-45).
int shadow_face(V3 d)
{
    V3 a = V3{fabsf(d.x), fabsf(d.y), fabsf(d.z)};
    if (a.x >= a.y && a.x >= a.z) return d.x >= 0.0f ? 0 : 1;
    if (a.y >= a.z) return d.y >= 0.0f ? 2 : 3;
    return d.z >= 0.0f ? 4 : 5;
}

// shadowAtlasUv (shadow.sh:48-64).
V2 shadow_atlas_uv(const rgi_frame* f, V3 position, int face)
{
    V4 clip = mul(f->shadow_view_proj + face * 16, V4{position.x, position.y, position.z, 1.0f});
    V2 ndc = V2{clip.x / clip.w, clip.y / clip.w};
    bool obl = f->shadow_info[3] > 0.5f;
    V2 local = V2{ndc.x * 0.5f + 0.5f, obl ? ndc.y * 0.5f + 0.5f : 0.5f - ndc.y * 0.5f};
    float row = face >= 3 ? 1.0f : 0.0f;
    float column = (float)face - 3.0f * row;
    if (obl) row = 1.0f - row;
    return V2{(column + local.x) * (1.0f / 3.0f), (row + local.y) * 0.5f};
}

// shadowLookup (shadow.sh:72-97): the atlas uv of a surface point and the
// biased, normalised distance compared there; false when shadows are off.
bool shadow_lookup(const rgi_frame* f, const float* atlas, V3 position, V3 n, float filter_radius, V2& uv,
                   float& reference)
{
    uv = V2{0.0f, 0.0f};
    reference = 0.0f;
    if (!f->shadows_enabled || atlas == nullptr) return false;
    V3 light = V3{f->light_position[0], f->light_position[1], f->light_position[2]};
    float light_distance = length(light - position);
    float texel_size = ((2.0f * light_distance) * f->shadow_info[1]) * f->shadow_info[2];
    V3 pos = position + n * (texel_size * 1.5f);
    V3 from_light = pos - light;
    uv = shadow_atlas_uv(f, pos, shadow_face(from_light));
    float cosine = clampf(dot(n, -normalize(from_light)), 0.0f, 1.0f);
    float slope = fmin_(sqrtf(1.0f - cosine * cosine) / fmax_(cosine, 0.0001f), 8.0f);
    float bias = texel_size * (1.0f + filter_radius * slope);
    reference = (length(from_light) - bias) / f->shadow_info[0];
    return true;
}

// lightVisibilityHard (shadow.sh:99-108).
float light_visibility_hard(const Pass& p, V3 position, V3 n)
{
    const rgi_frame* f = p.f;
    V2 uv;
    float reference;
    if (p.in == nullptr || !shadow_lookup(f, p.in->shadow_map, position, n, 0.5f, uv, reference)) return 1.0f;
    float stored = p.in->shadow_map[texel(uv, f->shadow_width, f->shadow_height)];
    return reference <= stored ? 1.0f : 0.0f;
}

// lightVisibilitySoft (shadow.sh:113-134): 4 x 4 compares one atlas texel apart.
float light_visibility_soft(const rgi_frame* f, const float* atlas, V3 position, V3 n)
{
    V2 uv;
    float reference;
    if (!shadow_lookup(f, atlas, position, n, 2.0f, uv, reference)) return 1.0f;
    V2 atlas_texel = V2{f->shadow_info[2] / 3.0f, f->shadow_info[2] * 0.5f};
    float lit = 0.0f;
    for (int y = 0; y < 4; ++y) {
        for (int x = 0; x < 4; ++x) {
            V2 offset = V2{((float)x - 1.5f) * atlas_texel.x, ((float)y - 1.5f) * atlas_texel.y};
            V2 tap = V2{uv.x + offset.x, uv.y + offset.y};
            lit += reference <= atlas[texel(tap, f->shadow_width, f->shadow_height)] ? 1.0f : 0.0f;
        }
    }
    return lit / 16.0f;
}

// --------------------------------------------------------------------- sky

bool sky_enabled(const rgi_frame* f) { return f->sky_zenith[3] > 0.5f; }

// skyRadiance (sky.sh:20-31).
V3 sky_radiance(const rgi_frame* f, V3 d)
{
    if (!sky_enabled(f)) return v3s(0.0f);
    V3 horizon = V3{f->sky_horizon[0], f->sky_horizon[1], f->sky_horizon[2]};
    float h = clampf(d.y, -1.0f, 1.0f);
    if (h >= 0.0f) return mix3(horizon, V3{f->sky_zenith[0], f->sky_zenith[1], f->sky_zenith[2]}, sqrtf(h));
    return mix3(horizon, V3{f->sky_ground[0], f->sky_ground[1], f->sky_ground[2]}, sqrtf(-h));
}

// ---------------------------------------------------------------- ReSTIR GI

float luminance(V3 c) { return dot(c, V3{0.2126f, 0.7152f, 0.0722f}); }

// cosineHemisphere (restir.sh:201-212).
V3 cosine_hemisphere(V3 n, float r0, float r1)
{
    float radius = sqrtf(r0);
    float angle = (2.0f * kPi) * r1;
    float sa, ca;
    rgi_sincos(angle, &sa, &ca);
    V3 local = V3{radius * ca, radius * sa, sqrtf(fmax_(1.0f - r0, 0.0f))};
    V3 t, b;
    tangent_frame(n, t, b);
    return normalize((t * local.x + b * local.y) + n * local.z);
}

V3 light_position(const rgi_frame* f) { return V3{f->light_position[0], f->light_position[1], f->light_position[2]}; }
V3 light_color(const rgi_frame* f) { return V3{f->light_color[0], f->light_color[1], f->light_color[2]}; }

// evaluateSecondaryDirect (restir.sh:216-229) and evaluateWorldDirect (232-244).
V3 world_direct(const Pass& p, V3 albedo, V3 position, V3 n)
{
    float sc = scene_scale(p);
    V3 to_light = light_position(p.f) - position;
    float d2 = fmax_(dot(to_light, to_light), (0.01f * sc) * sc);
    float nl = fmax_(dot(n, to_light * inversesqrt(d2)), 0.0f);
    V3 direct = v3s(0.0f);
    if (nl > 0.0f) {
        direct = ((albedo * (light_color(p.f) / d2)) * (nl / kPi)) * light_visibility_hard(p, position, n);
    }
    return direct;
}

const float kSkySampleId = 30000000.0f;
bool is_sky_sample(float id) { return id > kSkySampleId * 0.5f; }

// encodeSampleId (restir.sh:255-260).
float encode_sample_id(const rgi_scene* s, int tri, V3 dir)
{
    V3 v0 = bvh_vertex(s, tri, 0);
    V3 pn = cross(bvh_vertex(s, tri, 1) - v0, bvh_vertex(s, tri, 2) - v0);
    return dot(pn, dir) <= 0.0f ? (float)(tri + 1) : -(float)(tri + 1);
}

// Reservoir and history planes: GI size (the G-buffer is the window's size).
V4 fetch_f4(const float* plane, const Pass& p, V2 uv)
{
    return float4(plane + texel(uv, p.gi_w, p.gi_h) * 4);
}

V4 fetch_h4(const uint16_t* plane, const Pass& p, V2 uv)
{
    return half4(plane + texel(uv, p.gi_w, p.gi_h) * 4);
}

// s_surfaceLight (renderer.cpp:690-695): the DI diffuse light or the
// G-buffer emission, both window-size planes.
V3 fetch_surface_light(const Pass& p, V2 uv)
{
    return xyz(half4(p.in->surface_light + texel(uv, p.g->width, p.g->height) * 4));
}

// The same-surface test of previousIndirect / historyIrradiance / reprojectHistory
// (restir.sh:288-295, 312-319; cs_restir_spatial.sc:67-76).
bool history_matches(const Pass& p, V2 uv, V3 position, V3 n)
{
    float sc = scene_scale(p);
    V3 rp = xyz(fetch_f4(p.in->history_receiver, p, uv));
    V3 rn = xyz(fetch_f4(p.in->history_normal, p, uv));
    V3 offset = rp - position;
    float bounce = 0.1f * sc;
    if (dot(rn, rn) < 0.5f || dot(offset, offset) > bounce * bounce || fabsf(dot(offset, n)) > 0.03f * sc ||
        dot(normalize(rn), n) < 0.9f)
        return false;
    return true;
}

// previousIndirect (restir.sh:265-298).
V3 previous_indirect(const Pass& p, V3 position, V3 n, V3 albedo)
{
    if (!p.f->history_valid || !p.f->multiple_bounces) return v3s(0.0f);
    V4 clip = mul(p.f->prev_view_proj, V4{position.x, position.y, position.z, 1.0f});
    if (clip.w <= 0.0001f) return v3s(0.0f);
    V2 uv = clip_to_uv(p.f, clip);
    if (!is_on_screen(uv)) return v3s(0.0f);
    if (!history_matches(p, uv, position, n)) return v3s(0.0f);
    return xyz(fetch_h4(p.in->history_color, p, uv)) * albedo;
}

// historyIrradiance (restir.sh:302-323).
bool history_irradiance(const Pass& p, V2 uv, V3 position, V3 n, V3& irradiance)
{
    irradiance = v3s(0.0f);
    if (!is_on_screen(uv)) return false;
    if (!history_matches(p, uv, position, n)) return false;
    irradiance = xyz(fetch_h4(p.in->history_color, p, uv));
    return true;
}

// A marched ray and its clip-space image: the point at distance t is
// origin + direction * t, whose clip position is clip_origin + t * clip_direction
// (projectToScreen of that point, restir.sh:187-199, with the projection
// distributed over the ray).
struct ClipRay {
    V3 origin, direction;
    V4 clip_origin, clip_direction;
};

ClipRay clip_ray(const rgi_frame* f, V3 origin, V3 direction)
{
    return ClipRay{origin, direction, mul(f->view_proj, V4{origin.x, origin.y, origin.z, 1.0f}),
                   mul_dir(f->view_proj, direction)};
}

// depthBehindSurface (restir.sh:327-358) at distance t along the ray.
float depth_behind_surface(const Pass& p, const ClipRay& ray, float t, V2& uv, V3& surface_position, bool& on_screen)
{
    surface_position = v3s(0.0f);
    V4 clip = V4{fmaf(ray.clip_direction.x, t, ray.clip_origin.x), fmaf(ray.clip_direction.y, t, ray.clip_origin.y),
                 0.0f, fmaf(ray.clip_direction.w, t, ray.clip_origin.w)};
    float ray_depth = clip.w;
    uv = V2{0.0f, 0.0f};
    on_screen = false;
    if (clip.w <= 0.0001f) return -1.0f;
    uv = clip_to_uv(p.f, clip);
    on_screen = is_on_screen(uv);
    if (!on_screen) return -1.0f;
    float depth = fetch_depth(p, uv);
    if (depth >= kSkyDepth) return -1.0f;
    // The surface's view depth, (view_proj * reconstructWorldPosition(uv,
    // depth)).w, is 1 / (c z + d) for any uv, with (0, 0, c, d) the w row of
    // the inverse view-projection of a perspective camera. The position is
    // needed only when the ray is behind the surface.
    const float* inv = p.f->inv_view_proj;
    float behind = ray_depth - 1.0f / fmaf(inv[11], to_clip_depth(p.f, depth), inv[15]);
    if (behind > 0.0f) {
        surface_position = reconstruct_world_position(p.f, uv, depth);
        V3 sn = decode_normal(p, uv);
        V3 ray_position = ray.origin + ray.direction * t;
        if (dot(ray_position - surface_position, sn) > -(0.01f * scene_scale(p))) return -1.0f;
    }
    return behind;
}

// traceScreenSpaceRay (restir.sh:364-436).
bool trace_screen_space_ray(const Pass& p, V3 origin, V3 direction, float max_distance, V3& hit_position)
{
    float sc = scene_scale(p);
    hit_position = v3s(0.0f);
    float previous_distance = 0.0f;
    bool previous_in_front = true;
    int steps = p.f->trace_steps;
    float step_scale = 32.0f / (float)steps;
    ClipRay ray = clip_ray(p.f, origin, direction);
    for (int i = 1; i <= steps; ++i) {
        if (previous_distance >= max_distance) break;
        float sp = (float)i * step_scale;
        float ray_distance = fmin_((0.02f + (sp * sp) * 0.0035f) * sc, max_distance);
        V2 uv;
        V3 surface;
        bool on_screen;
        float behind = depth_behind_surface(p, ray, ray_distance, uv, surface, on_screen);
        if (!on_screen) return false;
        bool in_front = behind <= 0.0f;
        if (!in_front && previous_in_front) {
            float near_d = previous_distance;
            float far_d = ray_distance;
            for (int r = 0; r < (g_plant == kPlantBisection ? 2 : 5); ++r) {
                float middle = 0.5f * (near_d + far_d);
                V2 muv;
                V3 msurf;
                bool mon;
                float mb = depth_behind_surface(p, ray, middle, muv, msurf, mon);
                if (mon && mb > 0.0f) {
                    far_d = middle;
                    behind = mb;
                    surface = msurf;
                } else {
                    near_d = middle;
                }
            }
            float thickness = 0.03f * sc + 0.03f * far_d;
            if (g_plant == kPlantThickness) thickness = 0.3f * sc + 0.3f * far_d;
            if (behind < thickness) {
                hit_position = surface;
                return true;
            }
        }
        previous_in_front = in_front;
        previous_distance = ray_distance;
    }
    return false;
}

// evaluateSample (restir.sh:445-530).
bool evaluate_sample(const Pass& p, V3 rpos, V3 rn, V3 ralbedo, V3 spos, float sid, V3& contribution, V3& snormal)
{
    const rgi_frame* f = p.f;
    float sc = scene_scale(p);
    contribution = v3s(0.0f);
    snormal = v3(0.0f, 0.0f, 1.0f);
    if (is_sky_sample(sid)) {
        snormal = -spos;
        contribution = (ralbedo * sky_radiance(f, spos)) * (fmax_(dot(rn, spos), 0.0f) / kPi);
        return sky_enabled(f);
    }
    bool visible_surface = false;
    V2 suv;
    float sview;
    if (project_to_screen(f, spos, suv, sview)) {
        float sdepth = fetch_depth(p, suv);
        visible_surface = sdepth < kSkyDepth && length(reconstruct_world_position(f, suv, sdepth) - spos) <= 0.1f * sc;
    }
    int tri = to_int(fabsf(sid) + 0.5f) - 1;
    if (!visible_surface && (tri < 0 || !bvh_available(p.s))) return false;
    if (visible_surface) {
        snormal = decode_normal(p, suv);
    } else {
        V3 v0 = bvh_vertex(p.s, tri, 0);
        snormal = normalize(cross(bvh_vertex(p.s, tri, 1) - v0, bvh_vertex(p.s, tri, 2) - v0)) * (sid < 0.0f ? -1.0f : 1.0f);
    }
    V3 to_receiver = rpos - spos;
    float d2 = dot(to_receiver, to_receiver);
    if (d2 < ((0.0001f * sc) * sc) * 0.01f) return true;
    V3 dir = to_receiver * inversesqrt(d2);
    float scos = dot(snormal, dir);
    float rcos = -dot(rn, dir);
    if (scos <= 0.0f || rcos <= 0.0f) return true;
    V3 radiance;
    if (visible_surface) {
        V3 albedo = diffuse_albedo(p, suv);
        radiance = (world_direct(p, albedo, spos, snormal) + previous_indirect(p, spos, snormal, albedo)) +
                   fetch_surface_light(p, suv) * f->lighting_surface;
    } else {
        V3 albedo = bvh_triangle_albedo(p.s, tri);
        radiance = (world_direct(p, albedo, spos, snormal) + previous_indirect(p, spos, snormal, albedo)) +
                   bvh_triangle_emission(p.s, tri) * f->lighting_emission;
    }
    contribution = (ralbedo * radiance) * (rcos / kPi);
    return true;
}

// geometryTerm (restir.sh:538-558).
float geometry_term(const Pass& p, V3 rpos, V3 rn, V3 spos, V3 sn, float sid)
{
    if (is_sky_sample(sid)) return fmax_(dot(rn, spos), 0.0f);
    float sc = scene_scale(p);
    V3 offset = spos - rpos;
    float d2 = dot(offset, offset);
    if (d2 < ((0.0001f * sc) * sc) * 0.01f) return 0.0f;
    V3 dir = offset * inversesqrt(d2);
    float rc = dot(rn, dir);
    float scos = -dot(sn, dir);
    return rc > 0.0f && scos > 0.0f ? (rc * scos) / d2 : 0.0f;
}

// isVisible (restir.sh:564-601).
bool is_visible(const Pass& p, V3 rpos, V3 rn, V3 spos, float sid)
{
    float sc = scene_scale(p);
    float bias = 0.02f * sc;
    float max_bvh = 4.0f * sc;
    V3 origin = rpos + rn * bias;
    bool sky = is_sky_sample(sid);
    V3 to_sample = sky ? spos * max_bvh : spos - origin;
    float sd = length(to_sample);
    V3 dir = to_sample / fmax_(sd, 0.0001f);
    bool visible = true;
    if (sky) {
        float occ = 0.0f;
        visible = false;
        if (bvh_available(p.s)) visible = trace_bvh(p.s, origin, dir, max_bvh, true, occ) < 0;
    } else if (p.f->ray_mode == RGI_RAY_BVH && bvh_available(p.s)) {
        float occ = 0.0f;
        visible = trace_bvh(p.s, origin, dir, sd - bias, true, occ) < 0;
    } else {
        V3 occ_pos;
        visible = !trace_screen_space_ray(p, origin, dir, sd * 0.95f - 0.05f * sc, occ_pos);
    }
    return visible;
}

// reconnectionJacobian (restir.sh:623-649).
float reconnection_jacobian(const Pass& p, V3 here, V3 there, V3 spos, V3 sn, float sid)
{
    if (is_sky_sample(sid)) return 1.0f;
    float sc = scene_scale(p);
    V3 to_here = here - spos;
    V3 to_there = there - spos;
    float hs = dot(to_here, to_here);
    float ts = dot(to_there, to_there);
    float lim = ((0.0001f * sc) * sc) * 0.01f;
    if (hs < lim || ts < lim) return 0.0f;
    float ch = dot(sn, to_here) * inversesqrt(hs);
    float ct = dot(sn, to_there) * inversesqrt(ts);
    if (ct <= 0.001f) return 0.0f;
    float j = fmax_((ch / ct) * (ts / hs), 0.0f);
    if (g_plant == kPlantNoJacobian) j = 1.0f;
    float clamp_value = p.f->jacobian_clamp;
    return clamp_value > 0.0f ? fmin_(j, clamp_value) : j;
}

// Reservoir (restir.sh:651-673).
struct Reservoir {
    V3 contribution = V3{0, 0, 0};
    V3 position = V3{0, 0, 0};
    V3 normal = V3{0, 0, 1};
    float id = 0.0f;
    float target = 0.0f;
    float weight_sum = 0.0f;
    float count = 0.0f;
};

// updateReservoir (restir.sh:676-699).
bool update_reservoir(Reservoir& r, V3 contribution, V3 position, V3 n, float id, float target, float weight,
                      float random)
{
    r.weight_sum += weight;
    if (random * r.weight_sum < weight) {
        r.contribution = contribution;
        r.position = position;
        r.normal = n;
        r.id = id;
        r.target = target;
        return true;
    }
    return false;
}

struct Neighbour {
    V3 sample_position;
    float sample_id;
    V3 receiver_position;
    V3 receiver_normal;
    float count;
    float weight;
};

// acceptNeighbour (restir.sh:715-741).
bool accept_neighbour(const Pass& p, V4 sample, V4 receiver, V4 rnormal, V3 position, V3 n, float radius,
                      Neighbour& nb)
{
    nb.sample_position = xyz(sample);
    nb.sample_id = rnormal.w;
    nb.receiver_position = xyz(receiver);
    nb.receiver_normal = v3(0.0f, 0.0f, 1.0f);
    nb.count = fmin_(sample.w, 20.0f);
    nb.weight = receiver.w;
    V3 rn = xyz(rnormal);
    if (sample.w <= 0.0f || dot(rn, rn) < 0.5f) return false;
    nb.receiver_normal = normalize(rn);
    V3 off = xyz(receiver) - position;
    return dot(off, off) <= radius * radius && fabsf(dot(off, n)) <= 0.03f * scene_scale(p) &&
           dot(nb.receiver_normal, n) >= 0.9f;
}

// reuseRadius (restir.sh:746-751).
float reuse_radius(const Pass& p, V3 position)
{
    float r = 0.4f * scene_scale(p);
    return p.f->world_space_reuse ? fmax_(r, 2.0f * grid_cell_size(p.f, p.grid_camera, position)) : r;
}

// PairwiseMis (restir.sh:789-931).
struct Mis {
    Reservoir combined;
    V3 canonical_position, canonical_normal;
    float canonical_id;
    float pair_canonical, canonical_geometry, canonical_share, total_count;
    bool has_canonical, neighbour_selected;
};

Mis begin_mis(const Pass& p, V3 position, V3 n, const Reservoir& c, float ccount, int taps)
{
    Mis m;
    m.canonical_position = c.position;
    m.canonical_normal = c.normal;
    m.canonical_id = c.id;
    m.pair_canonical = ccount / (float)(taps > 1 ? taps : 1);
    m.total_count = ccount;
    m.has_canonical = c.target > 0.000001f;
    m.canonical_geometry = m.has_canonical ? geometry_term(p, position, n, c.position, c.normal, c.id) : 0.0f;
    m.canonical_share = taps > 0 ? 0.0f : ccount;
    m.neighbour_selected = false;
    return m;
}

void skip_neighbour(Mis& m)
{
    if (g_plant != kPlantSkipShare) m.canonical_share += m.pair_canonical;
}

// A neighbour's pairwise MIS weight (restir.sh:886-892), times M_total, of
// its sample with normal `sn` shifted to this receiver.
float neighbour_mis(const Pass& p, const Mis& m, V3 position, V3 n, const Neighbour& nb, V3 sn, float pair_weight)
{
    float hg = geometry_term(p, position, n, nb.sample_position, sn, nb.sample_id);
    float tg = geometry_term(p, nb.receiver_position, nb.receiver_normal, nb.sample_position, sn, nb.sample_id);
    float nd = nb.count * (g_plant == kPlantMisHere ? hg : tg);
    float den = nd + m.pair_canonical * hg;
    return den > 0.0f ? (pair_weight * nd) / den : 0.0f;
}

void add_neighbour(const Pass& p, Mis& m, V3 position, V3 n, V3 albedo, const Neighbour& nb, float random)
{
    float ncount = nb.count;
    float pair_weight = ncount + m.pair_canonical;
    m.total_count += ncount;
    if (m.has_canonical) {
        float ng = geometry_term(p, nb.receiver_position, nb.receiver_normal, m.canonical_position, m.canonical_normal,
                                 m.canonical_id);
        if (ng > 0.0f) {
            if (!is_visible(p, nb.receiver_position, nb.receiver_normal, m.canonical_position, m.canonical_id)) ng = 0.0f;
        }
        float cd = m.pair_canonical * m.canonical_geometry;
        float den = ncount * ng + cd;
        m.canonical_share += den > 0.0f ? (pair_weight * cd) / den : pair_weight;
    }
    V3 contribution = v3s(0.0f);
    V3 sn = v3(0.0f, 0.0f, 1.0f);
    bool shifted = false;
    if (nb.weight > 0.0f) {
        shifted = evaluate_sample(p, position, n, albedo, nb.sample_position, nb.sample_id, contribution, sn);
    }
    if (shifted) {
        float target = luminance(contribution);
        float j = reconnection_jacobian(p, position, nb.receiver_position, nb.sample_position, sn, nb.sample_id);
        float mis = neighbour_mis(p, m, position, n, nb, sn, pair_weight);
        float w = ((mis * target) * nb.weight) * j;
        if (w > 0.0f && update_reservoir(m.combined, contribution, nb.sample_position, sn, nb.sample_id, target, w, random))
            m.neighbour_selected = true;
    }
}

float finish_mis(const Pass& p, Mis& m, V3 position, V3 n, const Reservoir& c, float cweight, float random)
{
    if (m.has_canonical) {
        float w = (m.canonical_share * c.target) * cweight;
        if (update_reservoir(m.combined, c.contribution, c.position, c.normal, c.id, c.target, w, random))
            m.neighbour_selected = false;
    }
    float cw = m.combined.target > 0.000001f ? m.combined.weight_sum / (m.total_count * m.combined.target) : 0.0f;
    if (cw > 0.0f && m.neighbour_selected) {
        if (!is_visible(p, position, n, m.combined.position, m.combined.id)) cw = 0.0f;
    }
    return cw;
}

// giPixelUv (restir.sh:180-184).
V2 gi_pixel_uv(const Pass& p, uint32_t index)
{
    uint32_t w = (uint32_t)p.gi_w;
    return V2{((float)(index % w) + 0.5f) * p.texel_size.x, ((float)(index / w) + 0.5f) * p.texel_size.y};
}

Pass make_pass(const rgi_frame* f, const rgi_scene* s, const rgi_gbuffer* g, const rgi_gi_inputs* in, bool previous_camera)
{
    Pass p;
    p.f = f;
    p.s = s;
    p.g = g;
    p.in = in;
    p.gi_w = f->gi_width;
    p.gi_h = f->gi_height;
    p.size = V2{(float)f->gi_width, (float)f->gi_height};
    p.texel_size = V2{1.0f / (float)f->gi_width, 1.0f / (float)f->gi_height};
    const float* cam = previous_camera ? f->prev_camera_position : f->camera_position;
    p.grid_camera = V3{cam[0], cam[1], cam[2]};
    p.grid = in ? in->grid : nullptr;
    p.grid_mask = f->grid_table_size - 1u;
    return p;
}

// The DI passes run at the window size (u_screenSize, cs_di_temporal.sc:28-30).
Pass make_di_pass(const rgi_frame* f, const rgi_scene* s, const rgi_gbuffer* g)
{
    Pass p = make_pass(f, s, g, nullptr, false);
    p.gi_w = g->width;
    p.gi_h = g->height;
    p.size = V2{(float)g->width, (float)g->height};
    p.texel_size = V2{1.0f / (float)g->width, 1.0f / (float)g->height};
    return p;
}

// ----------------------------------------------------------- specular GI

// sampleVisibleNormal (fs_specular.sc:51-68): Heitz 2018's sampling of the
// GGX normals visible from `view` (tangent frame, z up).
V3 sample_visible_normal(V3 view, float alpha, float r0, float r1)
{
    V3 stretched = normalize(V3{alpha * view.x, alpha * view.y, view.z});
    float length_squared = stretched.x * stretched.x + stretched.y * stretched.y;
    V3 t1 = length_squared > 0.0f ? V3{-stretched.y, stretched.x, 0.0f} * inversesqrt(length_squared)
                                  : V3{1.0f, 0.0f, 0.0f};
    V3 t2 = cross(stretched, t1);
    float radius = sqrtf(r0);
    float angle = (2.0f * kPi) * r1;
    float sa, ca;
    rgi_sincos(angle, &sa, &ca);
    float p1 = radius * ca;
    float p2 = radius * sa;
    float s = 0.5f * (1.0f + stretched.z);
    p2 = (1.0f - s) * sqrtf(fmax_(1.0f - p1 * p1, 0.0f)) + s * p2;
    V3 normal = (t1 * p1 + t2 * p2) + stretched * sqrtf(fmax_((1.0f - p1 * p1) - p2 * p2, 0.0f));
    return normalize(V3{alpha * normal.x, alpha * normal.y, fmax_(normal.z, 0.0f)});
}

// smithG1 (fs_specular.sc:71-75).
float smith_g1_ggx(float c, float alpha)
{
    float a2 = alpha * alpha;
    return (2.0f * c) / (c + sqrtf(a2 + ((1.0f - a2) * c) * c));
}

V3 reflect3(V3 i, V3 n) { return i - n * (2.0f * dot(n, i)); }

// The reflected direction of fs_specular.sc:160-167 at a surface.
V3 specular_direction(V3 n, V3 view, float alpha, float r0, float r1, V3& halfway)
{
    V3 tangent, bitangent;
    tangent_frame(n, tangent, bitangent);
    V3 local_view = V3{dot(view, tangent), dot(view, bitangent), fmax_(dot(n, view), 0.001f)};
    V3 lh = sample_visible_normal(local_view, alpha, r0, r1);
    halfway = (tangent * lh.x + bitangent * lh.y) + n * lh.z;
    return reflect3(-view, halfway);
}

// hitRadiance (fs_specular.sc:78-124).
V3 specular_hit_radiance(const Pass& p, const uint16_t* previous_hdr, V3 hit, float id, V3 dir)
{
    const rgi_frame* f = p.f;
    if (is_sky_sample(id)) return sky_radiance(f, dir);
    V2 uv;
    float view_depth;
    bool visible_surface = false;
    if (project_to_screen(f, hit, uv, view_depth)) {
        float depth = fetch_depth(p, uv);
        visible_surface = depth < kSkyDepth && length(reconstruct_world_position(f, uv, depth) - hit) <= 0.1f * scene_scale(p);
    }
    if (visible_surface) {
        V4 clip = mul(f->prev_view_proj, V4{hit.x, hit.y, hit.z, 1.0f});
        if (f->history_valid && clip.w > 0.0001f && previous_hdr != nullptr) {
            V2 previous_uv = clip_to_uv(f, clip);
            V4 previous = half4(previous_hdr + texel(previous_uv, p.g->width, p.g->height) * 4);
            if (is_on_screen(previous_uv) && previous.w > 0.5f) return xyz(previous);
        }
        V3 n = decode_normal(p, uv);
        V3 albedo = diffuse_albedo(p, uv);
        return (world_direct(p, albedo, hit, n) + previous_indirect(p, hit, n, albedo)) +
               fetch_surface_light(p, uv) * f->lighting_surface;
    }
    int tri = to_int(fabsf(id) + 0.5f) - 1;
    if (tri < 0 || !bvh_available(p.s)) return v3s(0.0f);
    V3 v0 = bvh_vertex(p.s, tri, 0);
    V3 n = normalize(cross(bvh_vertex(p.s, tri, 1) - v0, bvh_vertex(p.s, tri, 2) - v0));
    if (dot(n, -dir) < 0.0f) n = -n;
    V3 albedo = bvh_triangle_albedo(p.s, tri);
    return (world_direct(p, albedo, hit, n) + previous_indirect(p, hit, n, albedo)) +
           bvh_triangle_emission(p.s, tri) * f->lighting_emission;
}

// One thread of cs_restir_temporal up to the group sum (cs_restir_temporal.sc:57-269).
struct TemporalThread {
    bool inside, working;
    V3 position, normal, albedo;
    Reservoir canonical;
    float canonical_weight;
    Mis mis;
    float combined_weight;
    V2 previous_uv;
    float estimate;
};

void temporal_thread(const Pass& p, int x, int y, TemporalThread& t)
{
    const rgi_frame* f = p.f;
    float sc = scene_scale(p);
    t.inside = x < (int)p.size.x && y < (int)p.size.y;
    V2 uv = V2{((float)x + 0.5f) * p.texel_size.x, ((float)y + 0.5f) * p.texel_size.y};
    float depth = t.inside ? fetch_depth(p, uv) : 1.0f;
    t.working = t.inside && depth < kSkyDepth;
    t.position = reconstruct_world_position(f, uv, depth);
    t.normal = decode_normal(p, uv);
    t.albedo = diffuse_albedo(p, uv);
    V3 ray_origin = t.position + t.normal * (0.02f * sc);
    float gx = floorf(uv.x * p.size.x), gy = floorf(uv.y * p.size.y);
    uint32_t rng = pcg((uint32_t)to_int(gx) + pcg((uint32_t)to_int(gy) + pcg(f->frame_index)));

    Reservoir canonical;
    for (int k = 0; k < f->initial_candidates; ++k) {
        float r0 = next_random(rng);
        float r1 = next_random(rng);
        float rsel = next_random(rng);
        V3 dir = cosine_hemisphere(t.normal, r0, r1);
        canonical.count += 1.0f;
        V3 hit_position = v3s(0.0f);
        float hit_id = 0.0f;
        bool hit = false;
        if (t.working && f->ray_mode != RGI_RAY_BVH) hit = trace_screen_space_ray(p, ray_origin, dir, 3.6f * sc, hit_position);
        bool need_bvh = t.working && !hit && f->ray_mode != RGI_RAY_SCREEN;
        if (need_bvh) {
            float dist;
            int tri = trace_bvh(p.s, ray_origin, dir, 4.0f * sc, false, dist);
            if (tri >= 0) {
                hit = true;
                hit_position = ray_origin + dir * dist;
                hit_id = encode_sample_id(p.s, tri, dir);
            } else if (sky_enabled(f)) {
                hit = true;
                hit_position = dir;
                hit_id = kSkySampleId;
            }
        }
        V3 contribution = v3s(0.0f), hit_normal = v3s(0.0f);
        bool evaluated = false;
        if (hit) evaluated = evaluate_sample(p, t.position, t.normal, t.albedo, hit_position, hit_id, contribution, hit_normal);
        if (evaluated) {
            float target = luminance(contribution);
            float source_pdf = dot(t.normal, dir) / kPi;
            if (!is_sky_sample(hit_id)) {
                V3 from_receiver = hit_position - t.position;
                V3 from_origin = hit_position - ray_origin;
                float lim = ((0.0001f * sc) * sc) * 0.01f;
                float rs = fmax_(dot(from_receiver, from_receiver), lim);
                float os = fmax_(dot(from_origin, from_origin), lim);
                float hco = fmax_(-dot(hit_normal, from_origin), 0.0f) * inversesqrt(os);
                float hcr = -dot(hit_normal, from_receiver) * inversesqrt(rs);
                source_pdf = hcr > 0.0001f ? (source_pdf * (hco / os)) * (rs / hcr) : 0.0f;
                if (g_plant == kPlantPdfAtOrigin) source_pdf = dot(t.normal, dir) / kPi;
            }
            if (target > 0.0f && source_pdf > 0.0001f)
                update_reservoir(canonical, contribution, hit_position, hit_normal, hit_id, target, target / source_pdf, rsel);
        }
    }
    t.canonical = canonical;
    t.canonical_weight = canonical.target > 0.000001f ? canonical.weight_sum / (canonical.count * canonical.target) : 0.0f;

    bool history_valid = f->history_valid != 0;
    t.previous_uv = V2{0.0f, 0.0f};
    V4 clip_prev = mul(f->prev_view_proj, V4{t.position.x, t.position.y, t.position.z, 1.0f});
    if (clip_prev.w > 0.0001f) t.previous_uv = clip_to_uv(f, clip_prev);

    float radius = reuse_radius(p, t.position);
    t.mis = begin_mis(p, t.position, t.normal, canonical, canonical.count, f->temporal_taps);
    t.combined_weight = 0.0f;
    if (t.working) {
        for (int tap = 0; tap < f->temporal_taps; ++tap) {
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            float r2 = next_random(rng);
            float rsel = next_random(rng);
            V2 huv = V2{-1.0f, -1.0f};
            if (history_valid) {
                if (f->world_space_reuse) {
                    uint32_t entry = grid_pick_entry(f, p.grid, p.grid_camera, t.position, t.normal, r0, r1, r2);
                    if (entry != kGridInvalid) huv = gi_pixel_uv(p, entry);
                } else if (tap == 0 && clip_prev.w > 0.0001f) {
                    huv = t.previous_uv;
                }
            }
            Neighbour nb;
            if (!is_on_screen(huv) ||
                !accept_neighbour(p, fetch_f4(p.in->history_sample, p, huv), fetch_f4(p.in->history_receiver, p, huv),
                                  fetch_f4(p.in->history_normal, p, huv), t.position, t.normal, radius, nb)) {
                skip_neighbour(t.mis);
                continue;
            }
            add_neighbour(p, t.mis, t.position, t.normal, t.albedo, nb, rsel);
        }
        float rfinal = next_random(rng);
        t.combined_weight = finish_mis(p, t.mis, t.position, t.normal, canonical, t.canonical_weight, rfinal);
    }
    t.estimate = t.working ? t.mis.combined.target * t.combined_weight : 0.0f;
}

// -------------------------------------------------------------- ReSTIR DI

struct Surface {
    V3 position, normal, surface_normal, albedo;
    float roughness, metallic;
    V3 eye;
    bool valid;
};

// loadSurface (di.sh:90-105).
Surface load_surface(const Pass& p, V2 uv)
{
    Surface s;
    size_t t = texel(uv, p.g->width, p.g->height);
    float depth = p.g->depth[t];
    V4 ar = unorm8(p.g->albedo[t]);
    V4 nm = unorm8(p.g->normal[t]);
    V4 sn = unorm8(p.g->shading_normal[t]);
    s.valid = depth < kSkyDepth;
    s.position = reconstruct_world_position(p.f, uv, fmin_(depth, kSkyDepth));
    s.surface_normal = normalize(V3{nm.x * 2.0f - 1.0f, nm.y * 2.0f - 1.0f, nm.z * 2.0f - 1.0f});
    s.normal = normalize(V3{sn.x * 2.0f - 1.0f, sn.y * 2.0f - 1.0f, sn.z * 2.0f - 1.0f});
    s.albedo = V3{ar.x, ar.y, ar.z};
    s.roughness = ar.w;
    s.metallic = nm.w;
    s.eye = V3{p.f->camera_position[0], p.f->camera_position[1], p.f->camera_position[2]};
    return s;
}

V4 light_data(const rgi_scene* s, int light, int row) { return float4(s->lights + ((size_t)light * 4 + (size_t)row) * 4); }
bool is_point_light(const rgi_scene* s, int light) { return light_data(s, light, 0).w > 0.5f; }

// selectLight (di.sh:123-142).
int select_light(const Pass& p, float random, float& pmf)
{
    int low = 0, high = p.f->di_light_count - 1;
    for (int step = 0; step < 24 && low < high; ++step) {
        int middle = (low + high) >> 1;
        if (light_data(p.s, middle, 2).w < random) low = middle + 1;
        else high = middle;
    }
    pmf = light_data(p.s, low, 3).w;
    return low;
}

// triangleArea (di.sh:144-148).
float triangle_area(const rgi_scene* s, int light)
{
    V3 v0 = xyz(light_data(s, light, 0));
    return 0.5f * length(cross(xyz(light_data(s, light, 1)) - v0, xyz(light_data(s, light, 2)) - v0));
}

// sampleLightPoint (di.sh:153-167).
V3 sample_light_point(const rgi_scene* s, int light, float r0, float r1, float& density)
{
    V3 v0 = xyz(light_data(s, light, 0));
    if (is_point_light(s, light)) {
        density = 1.0f;
        return v0;
    }
    float root = sqrtf(r0);
    float u = 1.0f - root;
    float v = r1 * root;
    density = 1.0f / fmax_(triangle_area(s, light), 1e-12f);
    return (v0 + (xyz(light_data(s, light, 1)) - v0) * u) + (xyz(light_data(s, light, 2)) - v0) * v;
}

// distributionGGX, geometrySchlickGGX, fresnelSchlick, evaluateBrdf (brdf.sh:14-62).
float distribution_ggx(V3 n, V3 h, float roughness)
{
    float alpha = roughness * roughness;
    float a2 = alpha * alpha;
    float nh = fmax_(dot(n, h), 0.0f);
    float den = (nh * nh) * (a2 - 1.0f) + 1.0f;
    return a2 / fmax_((kPi * den) * den, 0.0001f);
}

float geometry_schlick_ggx(float nd, float roughness)
{
    float k = ((roughness + 1.0f) * (roughness + 1.0f)) / 8.0f;
    return nd / fmax_(nd * (1.0f - k) + k, 0.0001f);
}

V3 fresnel_schlick(V3 h, V3 v, V3 f0)
{
    float c = clampf(dot(h, v), 0.0f, 1.0f);
    float w = pow5(1.0f - c);
    return f0 + (v3s(1.0f) - f0) * w;
}

void evaluate_brdf(V3 n, V3 v, V3 l, V3 albedo, float roughness_in, float metallic, V3& diffuse, V3& specular)
{
    float nl = fmax_(dot(n, l), 0.0f);
    float nv = fmax_(dot(n, v), 0.0f);
    V3 h = normalize(v + l);
    V3 f0 = mix3(v3s(0.04f), albedo, metallic);
    V3 fr = fresnel_schlick(h, v, f0);
    float roughness = fmax_(roughness_in, 0.05f);
    float d = distribution_ggx(n, h, roughness);
    float g = geometry_schlick_ggx(nl, roughness) * geometry_schlick_ggx(nv, roughness);
    specular = ((fr * (d * g)) / fmax_((4.0f * nl) * nv, 0.001f)) * nl;
    diffuse = (((v3s(1.0f) - fr) * (1.0f - metallic)) * albedo / kPi) * nl;
}

// specularAlbedo (brdf.sh:68-76): Karis' fit of the split-sum specular albedo.
V3 specular_albedo(V3 f0, float roughness, float nv)
{
    float rx = -1.0f * roughness + 1.0f;
    float ry = -0.0275f * roughness + 0.0425f;
    float rz = -0.572f * roughness + 1.04f;
    float rw = 0.022f * roughness + -0.04f;
    float a004 = fmin_(rx * rx, rgi_exp2f(-9.28f * nv)) * rx + ry;
    float sx = -1.04f * a004 + rz;
    float sy = 1.04f * a004 + rw;
    return f0 * sx + v3s(sy);
}

// evaluateLightSample (di.sh:180-212).
void evaluate_light_sample(const Pass& p, const Surface& s, int light, V3 point, V3& diffuse, V3& specular)
{
    diffuse = v3s(0.0f);
    specular = v3s(0.0f);
    if (light < 0 || light >= p.f->di_light_count) return;
    float sc = scene_scale(p);
    V3 to_light = point - s.position;
    float d2 = fmax_(dot(to_light, to_light), (0.0001f * sc) * sc);
    V3 dir = to_light * inversesqrt(d2);
    if (dot(s.surface_normal, dir) <= 0.0f) return;
    V3 radiance = xyz(light_data(p.s, light, 3)) / d2;
    if (!is_point_light(p.s, light)) {
        V3 v0 = xyz(light_data(p.s, light, 0));
        V3 ln = normalize(cross(xyz(light_data(p.s, light, 1)) - v0, xyz(light_data(p.s, light, 2)) - v0));
        radiance = radiance * fabsf(dot(ln, dir));
    }
    V3 d, sp;
    evaluate_brdf(s.normal, normalize(s.eye - s.position), dir, s.albedo, s.roughness, s.metallic, d, sp);
    diffuse = d * radiance;
    specular = sp * radiance;
}

float light_target(const Pass& p, const Surface& s, int light, V3 point)
{
    V3 d, sp;
    evaluate_light_sample(p, s, light, point, d, sp);
    return luminance(d + sp);
}

// lightVisible (di.sh:224-232).
bool light_visible(const Pass& p, const Surface& s, V3 point)
{
    float bias = 0.01f * scene_scale(p);
    V3 origin = s.position + s.surface_normal * bias;
    V3 to_point = point - origin;
    float pd = length(to_point);
    float occ = 0.0f;
    return trace_bvh(p.s, origin, to_point / fmax_(pd, 0.0001f), pd * 0.999f - bias, true, occ) < 0;
}

// sameSurface (di.sh:235-242).
bool same_surface(const Pass& p, V3 position, V3 n, V3 other_position, V3 other_normal, float radius)
{
    V3 off = other_position - position;
    return dot(other_normal, other_normal) > 0.5f && dot(off, off) <= radius * radius &&
           fabsf(dot(off, n)) <= 0.03f * scene_scale(p) && dot(normalize(other_normal), n) >= 0.9f;
}

// ------------------------------------------------------------- denoiser

// geometryWeight (svgf.sh:27-37): exp(-plane distance / (0.02 scale)) times
// the normal cosine to the 64th power; 0 for a neighbour without a normal.
float svgf_geometry_weight(float scale, V3 p, V3 n, V3 np, V3 nn)
{
    if (dot(nn, nn) < 0.5f) return 0.0f;
    float plane = fabsf(dot(n, np - p)) / (0.02f * scale);
    float normal_weight = rgi_pow64(fmax_(dot(n, normalize(nn)), 0.0f));
    return rgi_expf(-plane) * normal_weight;
}

// isInside (svgf.sh:39-42).
bool svgf_inside(V2 uv) { return uv.x >= 0.0f && uv.x <= 1.0f && uv.y >= 0.0f && uv.y <= 1.0f; }

// Plane fetches at a uv, point sampled with clamp.
struct Planes {
    int w, h;
    V4 f4(const float* p, V2 uv) const { return float4(p + texel(uv, w, h) * 4); }
    V4 h4(const uint16_t* p, V2 uv) const { return half4(p + texel(uv, w, h) * 4); }
};

// splineWeight (fs_svgf_atrous.sc:29-33): the B3 spline 3/8, 1/4, 1/16.
float spline_weight(int offset)
{
    int d = offset < 0 ? -offset : offset;
    return d == 0 ? 0.375f : (d == 1 ? 0.25f : 0.0625f);
}

int thread_count = 0;

template <typename F>
void parallel_for(int64_t n, F&& fn)
{
#ifdef _OPENMP
    int threads = thread_count > 0 ? thread_count : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int64_t i = 0; i < n; ++i) fn(i);
#else
    for (int64_t i = 0; i < n; ++i) fn(i);
#endif
}

} // namespace

// ===================================================================== C API

extern "C" {

void orc_set_threads(int threads) { thread_count = threads; }
void orc_plant(int mistake) { g_plant = mistake; }

// traceScreenSpaceRay (restir.sh:364-436) of one ray over a G-buffer, for
// the march's pins: returns whether it hit, and the hit surface point.
int orc_screen_march(const rgi_frame* f, const rgi_gbuffer* g, const float* origin3, const float* dir3,
                     float max_distance, float* hit3)
{
    Pass p = make_pass(f, nullptr, g, nullptr, false);
    V3 hit;
    bool h = trace_screen_space_ray(p, V3{origin3[0], origin3[1], origin3[2]}, V3{dir3[0], dir3[1], dir3[2]},
                                    max_distance, hit);
    hit3[0] = hit.x;
    hit3[1] = hit.y;
    hit3[2] = hit.z;
    return h ? 1 : 0;
}

// reconnectionJacobian (restir.sh:623-649) of a sample point (position,
// normal, not the sky) shifted from receiver `there` to receiver `here`.
float orc_reconnection_jacobian(const rgi_frame* f, const float* here3, const float* there3, const float* sample3,
                                const float* normal3)
{
    Pass p = make_pass(f, nullptr, nullptr, nullptr, false);
    return reconnection_jacobian(p, V3{here3[0], here3[1], here3[2]}, V3{there3[0], there3[1], there3[2]},
                                 V3{sample3[0], sample3[1], sample3[2]}, V3{normal3[0], normal3[1], normal3[2]}, 1.0f);
}

// The pairwise MIS of restir.sh:789-931 over one sample point x (position,
// normal) that the canonical reservoir and `taps` neighbour reservoirs all
// hold: begin_mis, then per neighbour add_neighbour (a negative count marks
// a tap that found no reservoir: skip_neighbour), with every visibility test
// passing (`scene` must see nothing between the points). Writes the
// canonical share m_c(x) M_total, each neighbour's m_i(x) M_total (0 for a
// skipped tap) and M_total. Neighbours: `taps` rows of 7 floats, receiver
// position, receiver normal, count.
void orc_gi_pairwise_weights(const rgi_frame* f, const rgi_scene* s, const float* receiver6, float canonical_count,
                             const float* sample6, int taps, const float* neighbours, float* out_canonical_share,
                             float* out_neighbour_mis, float* out_total)
{
    Pass p = make_pass(f, s, nullptr, nullptr, false);
    V3 position = V3{receiver6[0], receiver6[1], receiver6[2]}, n = V3{receiver6[3], receiver6[4], receiver6[5]};
    Reservoir c;
    c.position = V3{sample6[0], sample6[1], sample6[2]};
    c.normal = V3{sample6[3], sample6[4], sample6[5]};
    c.target = 1.0f;
    c.id = 1.0f;
    Mis m = begin_mis(p, position, n, c, canonical_count, taps);
    for (int i = 0; i < taps; ++i) {
        const float* r = neighbours + (size_t)i * 7;
        out_neighbour_mis[i] = 0.0f;
        if (r[6] < 0.0f) {
            skip_neighbour(m);
            continue;
        }
        // The neighbour's MIS weight of x as add_neighbour forms it (its
        // shift to this receiver is replaced by x itself).
        Neighbour nb;
        nb.sample_position = c.position;
        nb.sample_id = c.id;
        nb.receiver_position = V3{r[0], r[1], r[2]};
        nb.receiver_normal = V3{r[3], r[4], r[5]};
        nb.count = r[6];
        nb.weight = 0.0f; // add_neighbour then only updates the canonical share and M_total
        add_neighbour(p, m, position, n, v3s(0.0f), nb, 0.5f);
        out_neighbour_mis[i] = neighbour_mis(p, m, position, n, nb, c.normal, nb.count + m.pair_canonical);
    }
    *out_canonical_share = m.canonical_share;
    *out_total = m.total_count;
}
int orc_get_threads(void)
{
#ifdef _OPENMP
    return thread_count > 0 ? thread_count : omp_get_max_threads();
#else
    return 1;
#endif
}

uint32_t orc_pcg(uint32_t value) { return pcg(value); }
uint32_t orc_checksum_hash(uint32_t value) { return checksum_hash(value); }
float orc_next_random(uint32_t* state) { return next_random(*state); }
float orc_grid_cell_size(const rgi_frame* frame, const float* position3, const float* camera3)
{
    return grid_cell_size(frame, V3{camera3[0], camera3[1], camera3[2]}, V3{position3[0], position3[1], position3[2]});
}
uint32_t orc_grid_normal_bucket(const float* n) { return grid_normal_bucket(V3{n[0], n[1], n[2]}); }
void orc_grid_key(const rgi_frame* frame, const float* position3, const float* normal3, float cell_size,
                  uint32_t* slot, uint32_t* checksum)
{
    grid_key(frame->grid_table_size - 1u, V3{position3[0], position3[1], position3[2]},
             V3{normal3[0], normal3[1], normal3[2]}, cell_size, *slot, *checksum);
}
void orc_cosine_hemisphere(const float* n, float r0, float r1, float* out)
{
    V3 d = cosine_hemisphere(V3{n[0], n[1], n[2]}, r0, r1);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}
void orc_sincos(float x, float* s, float* c) { rgi_sincos(x, s, c); }
float orc_expf(float x) { return rgi_expf(x); }
float orc_pow64(float x) { return rgi_pow64(x); }
float orc_exp2f(float x) { return rgi_exp2f(x); }
void orc_specular_albedo(const float* f0, float roughness, float nv, float* out)
{
    V3 r = specular_albedo(V3{f0[0], f0[1], f0[2]}, roughness, nv);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

// fs_composite.sc:82-143 over gbuffer->width x height pixels; GI and indirect
// specular planes at frame->gi_width x gi_height (u_screenInfo).
void orc_composite(const rgi_frame* f, const rgi_gbuffer* g, const rgi_composite_inputs* in, uint16_t* out)
{
    int w = g->width, h = g->height;
    V2 ts = V2{1.0f / (float)w, 1.0f / (float)h};
    V2 gi_size = V2{(float)f->gi_width, (float)f->gi_height};
    V2 gi_texel = V2{1.0f / (float)f->gi_width, 1.0f / (float)f->gi_height};
    V3 camera = V3{f->camera_position[0], f->camera_position[1], f->camera_position[2]};
    auto h4 = [](const uint16_t* plane, size_t index) {
        return plane ? half4(plane + index * 4) : V4{0.0f, 0.0f, 0.0f, 0.0f};
    };
    auto gnormal = [](uint32_t word) {
        V4 n = unorm8(word);
        return normalize(V3{n.x * 2.0f - 1.0f, n.y * 2.0f - 1.0f, n.z * 2.0f - 1.0f});
    };
    parallel_for((int64_t)w * h, [&](int64_t i) {
        int x = (int)(i % w), y = (int)(i / w);
        V2 uv = V2{((float)x + 0.5f) * ts.x, ((float)y + 0.5f) * ts.y};
        size_t t = texel(uv, w, h);
        float depth = g->depth[t];
        if (depth >= kSkyDepth) {
            V3 far_position = reconstruct_world_position(f, uv, 1.0f);
            V3 direction = normalize(far_position - camera);
            V3 sky = sky_radiance(f, direction);
            store_half4(out + i * 4, sky_enabled(f) ? V4{sky.x, sky.y, sky.z, 1.0f} : V4{0.0f, 0.0f, 0.0f, 0.0f});
            return;
        }
        V4 albedo_roughness = unorm8(g->albedo[t]);
        V4 normal_metallic = unorm8(g->normal[t]);
        V3 albedo = xyz(albedo_roughness);
        float roughness = fmax_(albedo_roughness.w, 0.05f);
        float metallic = normal_metallic.w;
        V3 surface_normal = normalize(V3{normal_metallic.x * 2.0f - 1.0f, normal_metallic.y * 2.0f - 1.0f,
                                         normal_metallic.z * 2.0f - 1.0f});
        V3 n = gnormal(g->shading_normal[t]);
        V4 eo = half4(g->emission + t * 4);
        V3 position = reconstruct_world_position(f, uv, depth);
        V3 view_direction = normalize(camera - position);
        V3 to_light = V3{f->light_position[0], f->light_position[1], f->light_position[2]} - position;
        float distance_squared = fmax_(dot(to_light, to_light), (0.01f * f->scene_scale) * f->scene_scale);
        V3 light_direction = to_light * inversesqrt(distance_squared);
        V3 diffuse, specular;
        evaluate_brdf(n, view_direction, light_direction, albedo, roughness, metallic, diffuse, specular);
        float visibility = 0.0f;
        if (dot(n, light_direction) > 0.0f)
            visibility = light_visibility_soft(f, in->shadow_map, position, surface_normal);
        V3 light_color = V3{f->light_color[0], f->light_color[1], f->light_color[2]};
        // A resolved diffuse plane (the cache's prediction; no counterpart in
        // the reference) stands for the DI diffuse light and the GI together.
        V3 diffuse_light = in->resolved_diffuse ? xyz(h4(in->resolved_diffuse, t)) : xyz(h4(in->di_diffuse, t));
        V3 direct = (((diffuse + specular) * (light_color / distance_squared)) * visibility + diffuse_light) +
                    xyz(h4(in->di_specular, t));
        // upsampleIndirect (fs_composite.sc:48-80).
        float tcx = uv.x * gi_size.x - 0.5f, tcy = uv.y * gi_size.y - 0.5f;
        float bx = floorf(tcx), by = floorf(tcy);
        float blx = tcx - bx, bly = tcy - by;
        V3 sum = v3s(0.0f), specular_sum = v3s(0.0f);
        float weight_sum = 0.0f;
        for (int k = 0; k < 4; ++k) {
            float ox = (float)(k & 1), oy = (float)(k >> 1);
            V2 tuv = V2{((bx + ox) + 0.5f) * gi_texel.x, ((by + oy) + 0.5f) * gi_texel.y};
            size_t gt = texel(tuv, w, h);
            float tap_depth = g->depth[gt];
            if (tap_depth >= kSkyDepth) continue;
            V3 tap_position = reconstruct_world_position(f, tuv, tap_depth);
            V3 tap_normal = gnormal(g->normal[gt]);
            float wx = mixf(1.0f - blx, blx, ox), wy = mixf(1.0f - bly, bly, oy);
            float plane_step = 0.03f * f->scene_scale >= fabsf(dot(tap_position - position, surface_normal)) ? 1.0f : 0.0f;
            float normal_step = dot(tap_normal, surface_normal) >= 0.9f ? 1.0f : 0.0f;
            float weight = ((wx * wy + 0.01f) * plane_step) * normal_step;
            size_t it = texel(tuv, f->gi_width, f->gi_height);
            sum = sum + xyz(h4(in->gi, it)) * weight;
            specular_sum = specular_sum + xyz(h4(in->specular, it)) * weight;
            weight_sum += weight;
        }
        size_t own = texel(uv, f->gi_width, f->gi_height);
        V3 irradiance = weight_sum > 0.0f ? sum / weight_sum : xyz(h4(in->gi, own));
        V3 specular_radiance = weight_sum > 0.0f ? specular_sum / weight_sum : xyz(h4(in->specular, own));
        V3 modulate = in->options[2] > 0.5f ? (albedo * (1.0f - metallic)) * eo.w : v3s(1.0f);
        V3 indirect = in->resolved_diffuse ? v3s(0.0f) : irradiance * modulate;
        V3 f0 = mix3(v3s(0.04f), albedo, metallic);
        indirect = indirect + ((specular_radiance * specular_albedo(f0, roughness, fmax_(dot(n, view_direction), 0.001f))) *
                               eo.w) *
                                  in->specular_weight;
        V3 color = (direct * in->options[0] + indirect * in->options[1]) + xyz(eo) * in->options[3];
        store_half4(out + i * 4, V4{color.x, color.y, color.z, 1.0f});
    });
}

// fs_svgf_variance.sc:24-73: per-pixel luminance variance, from the pixel's
// own moments after 4 frames of history, else from its 7 x 7 same-surface
// neighbourhood, boosted by 4 / history length.
// The reflected direction and its weight G1(l) for roughness `roughness`
// (alpha = roughness^2) from two randoms: fs_specular.sc:139-170.
void orc_specular_sample(const float* n3, const float* v3, float roughness, float r0, float r1, float* out_dir,
                         float* out_g1)
{
    V3 n = V3{n3[0], n3[1], n3[2]}, view = V3{v3[0], v3[1], v3[2]};
    float alpha = roughness * roughness;
    V3 halfway;
    V3 d = specular_direction(n, view, alpha, r0, r1, halfway);
    out_dir[0] = d.x;
    out_dir[1] = d.y;
    out_dir[2] = d.z;
    float nl = dot(n, d);
    *out_g1 = nl > 0.0f ? smith_g1_ggx(nl, alpha) : 0.0f;
}

// fs_specular.sc:127-226: one ray per pixel, sequentially per pixel.
void orc_specular_gi(const rgi_frame* f, const rgi_scene* s, const rgi_gbuffer* g, const rgi_gi_inputs* in,
                     const uint16_t* previous_hdr, const uint16_t* history, uint16_t* out)
{
    Pass p = make_pass(f, s, g, in, false);
    float sc = f->scene_scale;
    int w = g->width, h = g->height;        // G-buffer
    int gw = f->gi_width, gh = f->gi_height; // this pass's pixels and planes
    V3 camera = V3{f->camera_position[0], f->camera_position[1], f->camera_position[2]};
    parallel_for((int64_t)gw * gh, [&](int64_t i) {
        int x = (int)(i % gw), y = (int)(i / gw);
        V2 uv = V2{((float)x + 0.5f) * p.texel_size.x, ((float)y + 0.5f) * p.texel_size.y};
        float depth = fetch_depth(p, uv);
        if (depth >= kSkyDepth) {
            store_half4(out + i * 4, V4{0.0f, 0.0f, 0.0f, 0.0f});
            return;
        }
        V3 position = reconstruct_world_position(f, uv, depth);
        V3 n = decode_normal(p, uv);
        size_t t = texel(uv, w, h);
        V4 albedo_roughness = unorm8(g->albedo[t]);
        float metallic = unorm8(g->normal[t]).w;
        float roughness = fmax_(albedo_roughness.w, 0.02f);
        float alpha = roughness * roughness;
        V3 view = normalize(camera - position);
        float normal_view = fmax_(dot(n, view), 0.001f);
        V3 f0 = mix3(v3s(0.04f), xyz(albedo_roughness), metallic);
        float gx = floorf(uv.x * p.size.x), gy = floorf(uv.y * p.size.y);
        uint32_t rng = pcg((uint32_t)to_int(gx) + pcg((uint32_t)to_int(gy) + pcg(f->frame_index ^ 0x2545f491u)));
        float r0 = next_random(rng);
        float r1 = next_random(rng);
        V3 halfway;
        V3 direction = specular_direction(n, view, alpha, r0, r1, halfway);
        float normal_light = dot(n, direction);
        V3 estimate = v3s(0.0f);
        if (normal_light > 0.0f) {
            V3 origin = position + n * (0.02f * sc);
            V3 hit_position = v3s(0.0f);
            float hit_id = 0.0f;
            bool hit = false;
            if (f->ray_mode != RGI_RAY_BVH) hit = trace_screen_space_ray(p, origin, direction, 3.6f * sc, hit_position);
            if (!hit && f->ray_mode != RGI_RAY_SCREEN) {
                float dist = 0.0f;
                int tri = trace_bvh(s, origin, direction, 4.0f * sc, false, dist);
                if (tri >= 0) {
                    hit = true;
                    hit_position = origin + direction * dist;
                    hit_id = encode_sample_id(s, tri, direction);
                } else if (sky_enabled(f)) {
                    hit = true;
                    hit_id = kSkySampleId;
                }
            }
            if (hit) {
                V3 fresnel = fresnel_schlick(halfway, view, f0);
                estimate = (specular_hit_radiance(p, previous_hdr, hit_position, hit_id, direction) * fresnel) *
                           smith_g1_ggx(normal_light, alpha);
            }
        }
        V3 demodulated = estimate / vmax(specular_albedo(f0, roughness, normal_view), v3s(0.001f));
        if (f->outlier_clamp > 0.0f) demodulated = vmin(demodulated, v3s(f->outlier_clamp));
        V3 previous = v3s(0.0f);
        float history_length = 0.0f;
        V4 clip_previous = mul(f->prev_view_proj, V4{position.x, position.y, position.z, 1.0f});
        if (f->history_valid && clip_previous.w > 0.0001f && history != nullptr) {
            V2 cuv = clip_to_uv(f, clip_previous);
            V2 puv = V2{(floorf(cuv.x * p.size.x) + 0.5f) * p.texel_size.x, (floorf(cuv.y * p.size.y) + 0.5f) * p.texel_size.y};
            if (is_on_screen(puv) && history_matches(p, puv, position, n)) {
                V4 hv = fetch_h4(history, p, puv);
                previous = xyz(hv);
                history_length = hv.w;
            }
        }
        float max_history = mixf(4.0f, 32.0f, clampf(roughness * 1.5f, 0.0f, 1.0f));
        history_length = fmin_(history_length + 1.0f, max_history);
        V3 color = mix3(previous, demodulated, 1.0f / history_length);
        store_half4(out + i * 4, V4{color.x, color.y, color.z, history_length});
    });
}

void orc_svgf_variance(const rgi_frame* f, const uint16_t* color, const uint16_t* moments, const float* receiver,
                       const float* normal, uint16_t* out)
{
    int w = f->gi_width, h = f->gi_height;
    Planes t{w, h};
    V2 ts = V2{1.0f / (float)w, 1.0f / (float)h};
    float scale = f->scene_scale;
    parallel_for((int64_t)w * h, [&](int64_t i) {
        int x = (int)(i % w), y = (int)(i / w);
        V2 uv0 = V2{((float)x + 0.5f) * ts.x, ((float)y + 0.5f) * ts.y};
        V3 n = xyz(t.f4(normal, uv0));
        if (dot(n, n) < 0.5f) {
            store_half4(out + i * 4, V4{0.0f, 0.0f, 0.0f, 0.0f});
            return;
        }
        n = normalize(n);
        V3 c = xyz(t.h4(color, uv0));
        V3 m = xyz(t.h4(moments, uv0));
        float history = m.z;
        if (history >= 4.0f) {
            store_half4(out + i * 4, V4{c.x, c.y, c.z, fmax_(m.y - m.x * m.x, 0.0f)});
            return;
        }
        V3 p = xyz(t.f4(receiver, uv0));
        float sx = 0.0f, sy = 0.0f, wsum = 0.0f;
        for (int dy = -3; dy <= 3; ++dy) {
            for (int dx = -3; dx <= 3; ++dx) {
                V2 uv = V2{uv0.x + (float)dx * ts.x, uv0.y + (float)dy * ts.y};
                if (!svgf_inside(uv)) continue;
                float wt = svgf_geometry_weight(scale, p, n, xyz(t.f4(receiver, uv)), xyz(t.f4(normal, uv)));
                V4 mm = t.h4(moments, uv);
                sx = sx + mm.x * wt;
                sy = sy + mm.y * wt;
                wsum = wsum + wt;
            }
        }
        float d = fmax_(wsum, 0.0001f);
        float mx = sx / d, my = sy / d;
        float variance = fmax_(my - mx * mx, 0.0f) * (4.0f / fmax_(history, 1.0f));
        store_half4(out + i * 4, V4{c.x, c.y, c.z, variance});
    });
}

// fs_svgf_atrous.sc:35-101: one a-trous iteration with taps `step` texels
// apart; rgb filtered with the weights, variance with their squares.
void orc_svgf_atrous(const rgi_frame* f, const uint16_t* input, const float* receiver, const float* normal,
                     int step, uint16_t* out)
{
    int w = f->gi_width, h = f->gi_height;
    Planes t{w, h};
    V2 ts = V2{1.0f / (float)w, 1.0f / (float)h};
    float scale = f->scene_scale;
    V2 spacing = V2{(float)step * ts.x, (float)step * ts.y};
    parallel_for((int64_t)w * h, [&](int64_t i) {
        int x = (int)(i % w), y = (int)(i / w);
        V2 uv0 = V2{((float)x + 0.5f) * ts.x, ((float)y + 0.5f) * ts.y};
        V3 n = xyz(t.f4(normal, uv0));
        if (dot(n, n) < 0.5f) {
            store_half4(out + i * 4, V4{0.0f, 0.0f, 0.0f, 0.0f});
            return;
        }
        n = normalize(n);
        V3 p = xyz(t.f4(receiver, uv0));
        V4 center = t.h4(input, uv0);
        float center_lum = luminance(xyz(center));
        float variance = 0.0f;
        for (int by = -1; by <= 1; ++by) {
            for (int bx = -1; bx <= 1; ++bx) {
                float wt = (bx == 0 ? 0.5f : 0.25f) * (by == 0 ? 0.5f : 0.25f);
                V2 uv = V2{uv0.x + (float)bx * ts.x, uv0.y + (float)by * ts.y};
                variance = variance + t.h4(input, uv).w * wt;
            }
        }
        float lum_scale = 1.0f / (4.0f * sqrtf(fmax_(variance, 0.0f)) + 0.0001f);
        V3 csum = v3s(0.0f);
        float vsum = 0.0f, wsum = 0.0f;
        for (int dy = -2; dy <= 2; ++dy) {
            for (int dx = -2; dx <= 2; ++dx) {
                V2 uv = V2{uv0.x + (float)dx * spacing.x, uv0.y + (float)dy * spacing.y};
                if (!svgf_inside(uv)) continue;
                V4 tap = t.h4(input, uv);
                float wt = spline_weight(dx) * spline_weight(dy);
                if (dx != 0 || dy != 0) {
                    wt = wt * svgf_geometry_weight(scale, p, n, xyz(t.f4(receiver, uv)), xyz(t.f4(normal, uv)));
                    wt = wt * rgi_expf(-fabsf(luminance(xyz(tap)) - center_lum) * lum_scale);
                }
                csum = csum + xyz(tap) * wt;
                vsum = vsum + (tap.w * wt) * wt;
                wsum = wsum + wt;
            }
        }
        V3 c = csum / wsum;
        store_half4(out + i * 4, V4{c.x, c.y, c.z, vsum / (wsum * wsum)});
    });
}
void orc_brdf(const float* n, const float* v, const float* l, const float* a, float roughness, float metallic,
              float* diffuse, float* specular)
{
    V3 d, sp;
    evaluate_brdf(V3{n[0], n[1], n[2]}, V3{v[0], v[1], v[2]}, V3{l[0], l[1], l[2]}, V3{a[0], a[1], a[2]}, roughness,
                  metallic, d, sp);
    diffuse[0] = d.x; diffuse[1] = d.y; diffuse[2] = d.z;
    specular[0] = sp.x; specular[1] = sp.y; specular[2] = sp.z;
}
uint16_t orc_half_bits(float value) { return rgi_half_bits(value); }

void orc_grid_keys(const rgi_frame* frame, const rgi_gbuffer* g, uint32_t* out_slot, uint32_t* out_checksum,
                   float* out_cell_size)
{
    // One key per GI pixel (frame->gi_width x gi_height) from the window-size G-buffer.
    int64_t n = (int64_t)frame->gi_width * frame->gi_height;
    parallel_for(n, [&](int64_t i) {
        int x = (int)(i % frame->gi_width), y = (int)(i / frame->gi_width);
        uint32_t slot = 0, checksum = 0;
        float cell = 0.0f;
        if (!insert_key(frame, g, x, y, slot, checksum, cell)) {
            slot = 0;
            checksum = 0;
            cell = 0.0f;
        }
        out_slot[i] = slot;
        out_checksum[i] = checksum;
        if (out_cell_size) out_cell_size[i] = cell;
    });
}

// cs_grid_clear (cs_grid_clear.sc:12-21).
void orc_grid_clear(uint32_t* grid, uint32_t table_size)
{
    memset(grid, 0, sizeof(uint32_t) * 2u * table_size);
}

// cs_grid_insert's gridInsert (hashgrid.sh:182-200), in the deterministic
// linearization: CAS rounds by probe index then pixel index, then the adds in
// pixel order.
void orc_grid_insert(const rgi_frame* frame, const rgi_gbuffer* g, uint32_t* grid)
{
    uint32_t table = frame->grid_table_size, mask = table - 1u;
    int64_t n = (int64_t)frame->gi_width * frame->gi_height;
    std::vector<uint32_t> slot0((size_t)n), checksum((size_t)n);
    orc_grid_keys(frame, g, slot0.data(), checksum.data(), nullptr);
    std::vector<uint32_t> resolved((size_t)n, kGridInvalid);
    std::vector<uint32_t> pending;
    for (int64_t i = 0; i < n; ++i)
        if (checksum[(size_t)i] != 0u) pending.push_back((uint32_t)i);
    for (uint32_t probe = 0; probe < kGridMaxProbes && !pending.empty(); ++probe) {
        std::vector<uint32_t> next;
        for (uint32_t i : pending) {
            uint32_t s = (slot0[i] + probe) & mask;
            if (grid[s] == 0u) {
                grid[s] = checksum[i];
                resolved[i] = s;
            } else if (grid[s] == checksum[i]) {
                resolved[i] = s;
            } else {
                next.push_back(i);
            }
        }
        pending.swap(next);
    }
    for (int64_t i = 0; i < n; ++i) {
        uint32_t s = resolved[(size_t)i];
        if (s == kGridInvalid) continue;
        uint32_t index = grid[table + s]++;
        if (index < kGridCellCapacity) grid[2u * table + s * kGridCellCapacity + index] = (uint32_t)i;
    }
}

uint32_t orc_grid_pick_entry(const rgi_frame* frame, const uint32_t* grid, const float* camera3, const float* position3,
                             const float* normal3, float r0, float r1, float r2)
{
    return grid_pick_entry(frame, grid, V3{camera3[0], camera3[1], camera3[2]},
                           V3{position3[0], position3[1], position3[2]}, V3{normal3[0], normal3[1], normal3[2]}, r0,
                           r1, r2);
}

// rgi_trace with RGI_OPT_TRIANGLE_TEST = 1: the same traversal with the
// watertight triangle test.
void orc_trace_watertight(const rgi_scene* scene, const float* rays, uint32_t count, int32_t* hit_triangle,
                          float* hit_distance, int any_hit)
{
    parallel_for((int64_t)count, [&](int64_t i) {
        const float* r = rays + (size_t)i * 8;
        float dist;
        int tri = trace_bvh(scene, V3{r[0], r[1], r[2]}, V3{r[4], r[5], r[6]}, r[3], any_hit != 0, dist, true);
        hit_triangle[i] = tri;
        hit_distance[i] = dist;
    });
}

void orc_trace(const rgi_scene* scene, const float* rays, uint32_t count, int32_t* hit_triangle, float* hit_distance,
               int any_hit)
{
    parallel_for((int64_t)count, [&](int64_t i) {
        const float* r = rays + (size_t)i * 8;
        float dist;
        int tri = trace_bvh(scene, V3{r[0], r[1], r[2]}, V3{r[4], r[5], r[6]}, r[3], any_hit != 0, dist);
        hit_triangle[i] = tri;
        hit_distance[i] = dist;
    });
}

// cs_restir_temporal (cs_restir_temporal.sc:57-297), one 4 x 8 group at a time.
void orc_gi_temporal(const rgi_frame* f, const rgi_scene* s, const rgi_gbuffer* g, const rgi_gi_inputs* in,
                     const rgi_gi_reservoirs* out)
{
    Pass p = make_pass(f, s, g, in, true);
    int groups_x = (p.gi_w + 3) / 4, groups_y = (p.gi_h + 7) / 8;
    parallel_for((int64_t)groups_x * groups_y, [&](int64_t gi) {
        int gx = (int)(gi % groups_x), gy = (int)(gi / groups_x);
        TemporalThread th[32];
        for (int t = 0; t < 32; ++t) temporal_thread(p, gx * 4 + (t & 3), gy * 8 + (t >> 2), th[t]);
        // groupSum (compute.sh:376-386): the tree in shared memory.
        float sx[32], sy[32];
        for (int t = 0; t < 32; ++t) {
            sx[t] = th[t].estimate;
            sy[t] = th[t].working ? 1.0f : 0.0f;
        }
        for (int stride = 16; stride > 0; stride >>= 1)
            for (int t = 0; t < stride; ++t) {
                sx[t] = sx[t] + sx[t + stride];
                sy[t] = sy[t] + sy[t + stride];
            }
        float tile_mean = sx[0] / fmax_(sy[0], 1.0f);
        for (int t = 0; t < 32; ++t) {
            TemporalThread& h = th[t];
            float cw = h.combined_weight;
            // Boiling filter (cs_restir_temporal.sc:272-289).
            if (h.working && f->boiling_filter > 0.0f && cw > 0.0f) {
                float reference = h.canonical.target * h.canonical_weight;
                V3 hc;
                if (f->history_valid && history_irradiance(p, h.previous_uv, h.position, h.normal, hc))
                    reference = fmax_(reference, luminance(hc * h.albedo));
                else
                    reference = fmax_(reference, tile_mean);
                float limit = f->boiling_filter * reference;
                if (h.estimate > limit) cw = cw * (limit / h.estimate);
            }
            if (!h.inside) continue;
            int x = gx * 4 + (t & 3), y = gy * 8 + (t >> 2);
            size_t o = ((size_t)y * p.gi_w + x) * 4;
            if (h.working) {
                store_float4(out->sample + o, V4{h.mis.combined.position.x, h.mis.combined.position.y,
                                                 h.mis.combined.position.z, h.mis.total_count});
                store_float4(out->receiver + o, V4{h.position.x, h.position.y, h.position.z, cw});
                store_float4(out->normal + o, V4{h.normal.x, h.normal.y, h.normal.z, h.mis.combined.id});
            } else {
                store_float4(out->sample + o, V4{0, 0, 0, 0});
                store_float4(out->receiver + o, V4{0, 0, 0, 0});
                store_float4(out->normal + o, V4{0, 0, 0, 0});
            }
        }
    });
}

// cs_restir_spatial (cs_restir_spatial.sc:49-246).
void orc_gi_spatial(const rgi_frame* f, const rgi_scene* s, const rgi_gbuffer* g, const rgi_gi_inputs* in,
                    const rgi_gi_reservoirs* res, uint16_t* out_color, uint16_t* out_moments)
{
    Pass p = make_pass(f, s, g, in, false);
    int64_t n = (int64_t)p.gi_w * p.gi_h;
    parallel_for(n, [&](int64_t i) {
        int x = (int)(i % p.gi_w), y = (int)(i / p.gi_w);
        V2 uv = V2{((float)x + 0.5f) * p.texel_size.x, ((float)y + 0.5f) * p.texel_size.y};
        uint16_t* oc = out_color + (size_t)i * 4;
        uint16_t* om = out_moments + (size_t)i * 4;
        float depth = fetch_depth(p, uv);
        if (depth >= kSkyDepth) {
            store_half4(oc, V4{0, 0, 0, 0});
            store_half4(om, V4{0, 0, 0, 0});
            return;
        }
        V3 position = reconstruct_world_position(f, uv, depth);
        V3 n3 = decode_normal(p, uv);
        V3 albedo = diffuse_albedo(p, uv);
        float px = floorf(uv.x * p.size.x), py = floorf(uv.y * p.size.y);
        uint32_t pixel_index = (uint32_t)to_int(py) * (uint32_t)p.gi_w + (uint32_t)to_int(px);
        uint32_t rng = pcg((uint32_t)to_int(px) + pcg((uint32_t)to_int(py) + pcg(f->frame_index ^ 0x68e31da4u)));

        V4 own_sample = fetch_f4(res->sample, p, uv);
        float cweight = fetch_f4(res->receiver, p, uv).w;
        float ccount = fmax_(fmin_(own_sample.w, 20.0f), 1.0f);
        Reservoir canonical;
        V3 ccontrib = v3s(0.0f), cnormal = v3(0.0f, 0.0f, 1.0f);
        bool cvalid = false;
        float own_id = fetch_f4(res->normal, p, uv).w;
        if (cweight > 0.0f) cvalid = evaluate_sample(p, position, n3, albedo, xyz(own_sample), own_id, ccontrib, cnormal);
        if (cvalid) {
            canonical.contribution = ccontrib;
            canonical.position = xyz(own_sample);
            canonical.normal = cnormal;
            canonical.id = own_id;
            canonical.target = luminance(ccontrib);
        }
        float radius = reuse_radius(p, position);
        Mis mis = begin_mis(p, position, n3, canonical, ccount, f->spatial_taps);
        for (int tap = 0; tap < f->spatial_taps; ++tap) {
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            float r2 = next_random(rng);
            float rsel = next_random(rng);
            V2 nuv = V2{-1.0f, -1.0f};
            if (f->world_space_reuse) {
                uint32_t entry = grid_pick_entry(f, p.grid, p.grid_camera, position, n3, r0, r1, r2);
                if (entry != kGridInvalid && entry != pixel_index) nuv = gi_pixel_uv(p, entry);
            } else {
                float disk = 12.0f * sqrtf(r0);
                float angle = (2.0f * kPi) * r1;
                float sa, ca;
                rgi_sincos(angle, &sa, &ca);
                V2 np = V2{floorf(px + ca * disk), floorf(py + sa * disk)};
                if (np.x != px || np.y != py) nuv = V2{(np.x + 0.5f) * p.texel_size.x, (np.y + 0.5f) * p.texel_size.y};
            }
            Neighbour nb;
            if (!is_on_screen(nuv) || !accept_neighbour(p, fetch_f4(res->sample, p, nuv), fetch_f4(res->receiver, p, nuv),
                                                        fetch_f4(res->normal, p, nuv), position, n3, radius, nb)) {
                skip_neighbour(mis);
                continue;
            }
            add_neighbour(p, mis, position, n3, albedo, nb, rsel);
        }
        float rfinal = next_random(rng);
        float cw = finish_mis(p, mis, position, n3, canonical, cweight, rfinal);
        V3 estimate = mis.combined.contribution * cw;
        if (f->outlier_clamp > 0.0f) estimate = vmin(estimate, v3s(f->outlier_clamp));
        V3 irradiance = estimate / vmax(albedo, v3s(0.001f));
        float lum = luminance(irradiance);

        V3 prev_color = v3s(0.0f), prev_moments = v3s(0.0f);
        V4 clip_prev = mul(f->prev_view_proj, V4{position.x, position.y, position.z, 1.0f});
        if (f->history_valid && clip_prev.w > 0.0001f) {
            // reprojectHistory (cs_restir_spatial.sc:49-93).
            V2 puv = clip_to_uv(f, clip_prev);
            V2 tc = V2{puv.x * p.size.x - 0.5f, puv.y * p.size.y - 0.5f};
            V2 base = V2{floorf(tc.x), floorf(tc.y)};
            V2 blend = V2{tc.x - base.x, tc.y - base.y};
            float wsum = 0.0f;
            V3 col = v3s(0.0f), mom = v3s(0.0f);
            for (int k = 0; k < 4; ++k) {
                V2 off = V2{(float)(k & 1), (float)(k >> 1)};
                V2 tuv = V2{((base.x + off.x) + 0.5f) * p.texel_size.x, ((base.y + off.y) + 0.5f) * p.texel_size.y};
                if (!is_on_screen(tuv)) continue;
                if (!history_matches(p, tuv, position, n3)) continue;
                V2 bil = V2{mixf(1.0f - blend.x, blend.x, off.x), mixf(1.0f - blend.y, blend.y, off.y)};
                float w = bil.x * bil.y;
                col = col + xyz(fetch_h4(in->history_color, p, tuv)) * w;
                mom = mom + xyz(fetch_h4(in->history_moments, p, tuv)) * w;
                wsum += w;
            }
            if (wsum < 0.01f) {
                prev_color = v3s(0.0f);
                prev_moments = v3s(0.0f);
            } else {
                prev_color = col / wsum;
                prev_moments = mom / wsum;
            }
        }
        float history_length = fmin_(prev_moments.z + 1.0f, 16.0f);
        float bw = 1.0f / history_length;
        V3 color = mix3(prev_color, irradiance, bw);
        float m0 = mixf(prev_moments.x, lum, bw);
        float m1 = mixf(prev_moments.y, lum * lum, bw);
        store_half4(oc, V4{color.x, color.y, color.z, history_length});
        store_half4(om, V4{m0, m1, history_length, 0.0f});
    });
}

// cs_di_temporal (cs_di_temporal.sc:32-180).
void orc_di_temporal(const rgi_frame* f, const rgi_scene* s, const rgi_gbuffer* g, const rgi_di_reservoirs* prev,
                     const rgi_di_reservoirs* out)
{
    Pass p = make_di_pass(f, s, g);
    int64_t n = (int64_t)g->width * g->height;
    parallel_for(n, [&](int64_t i) {
        int x = (int)(i % g->width), y = (int)(i / g->width);
        V2 uv = V2{((float)x + 0.5f) * p.texel_size.x, ((float)y + 0.5f) * p.texel_size.y};
        float* os = out->sample + (size_t)i * 4;
        float* orc = out->receiver + (size_t)i * 4;
        uint16_t* on = out->normal + (size_t)i * 4;
        Surface surface = load_surface(p, uv);
        if (!surface.valid || f->di_light_count == 0) {
            store_float4(os, V4{0, 0, 0, -1.0f});
            store_float4(orc, V4{0, 0, 0, 0});
            store_half4(on, V4{0, 0, 0, 0});
            return;
        }
        uint32_t rng = pcg((uint32_t)x + pcg((uint32_t)y + pcg(f->frame_index ^ 0x3c6ef372u)));
        int candidates = f->di_candidates;
        float wsum = 0.0f;
        V3 chosen_point = v3s(0.0f);
        int chosen_light = -1;
        float chosen_target = 0.0f;
        for (int c = 0; c < candidates; ++c) {
            float pmf;
            float rl = next_random(rng);
            int light = select_light(p, rl, pmf);
            float density;
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            V3 lp = sample_light_point(s, light, r0, r1, density);
            float target = light_target(p, surface, light, lp);
            float source_pdf = pmf * density;
            float w = source_pdf > 0.0f ? target / source_pdf : 0.0f;
            wsum += w;
            float rsel = next_random(rng);
            if (rsel * wsum < w) {
                chosen_point = lp;
                chosen_light = light;
                chosen_target = target;
            }
        }
        float ccount = (float)candidates;
        float cweight = chosen_target > 0.0f ? wsum / (ccount * chosen_target) : 0.0f;
        V3 cpoint = chosen_point;
        int clight = chosen_light;
        float ctarget = chosen_target;

        float tcount = 0.0f, tweight = 0.0f;
        V3 tpoint = v3s(0.0f);
        int tlight = -1;
        Surface previous = surface;
        V4 clip = mul(f->prev_view_proj, V4{surface.position.x, surface.position.y, surface.position.z, 1.0f});
        if (f->di_history_valid && clip.w > 0.0001f) {
            V2 puv = V2{(clip.x / clip.w) * 0.5f + 0.5f,
                        ((f->origin_bottom_left ? clip.y : -clip.y) / clip.w) * 0.5f + 0.5f};
            if (puv.x > 0.0f && puv.x < 1.0f && puv.y > 0.0f && puv.y < 1.0f) {
                size_t t = texel(puv, g->width, g->height);
                V4 sd = float4(prev->sample + t * 4);
                V4 rc = float4(prev->receiver + t * 4);
                V4 nm = half4(prev->normal + t * 4);
                int light = to_int(sd.w + 0.5f);
                if (sd.w >= 0.0f && light < f->di_light_count &&
                    same_surface(p, surface.position, surface.normal, xyz(rc), xyz(nm), 0.1f * scene_scale(p))) {
                    tcount = fmin_(nm.w, 20.0f);
                    tweight = rc.w;
                    tpoint = xyz(sd);
                    tlight = light;
                    previous.position = xyz(rc);
                    previous.normal = normalize(xyz(nm));
                    previous.eye = V3{f->prev_camera_position[0], f->prev_camera_position[1], f->prev_camera_position[2]};
                }
            }
        }
        wsum = 0.0f;
        chosen_light = -1;
        chosen_target = 0.0f;
        if (ctarget > 0.0f) {
            float ptarget = 0.0f;
            if (tcount > 0.0f) {
                if (light_visible(p, previous, cpoint)) ptarget = light_target(p, previous, clight, cpoint);
            }
            float den = ccount * ctarget + tcount * ptarget;
            float w = (((ccount * ctarget) / den) * ctarget) * cweight;
            if (g_plant == kPlantDiTemporalMis) w = ctarget * cweight;
            wsum += w;
            chosen_point = cpoint;
            chosen_light = clight;
            chosen_target = ctarget;
        }
        if (tcount > 0.0f && tweight > 0.0f) {
            float target = light_target(p, surface, tlight, tpoint);
            float ptarget = light_target(p, previous, tlight, tpoint);
            float den = tcount * ptarget + ccount * target;
            float w = den > 0.0f ? (((tcount * ptarget) / den) * target) * tweight : 0.0f;
            wsum += w;
            if (w > 0.0f) {
                float rsel = next_random(rng);
                if (rsel * wsum < w) {
                    chosen_point = tpoint;
                    chosen_light = tlight;
                    chosen_target = target;
                }
            }
        }
        float cw = chosen_target > 0.0f ? wsum / chosen_target : 0.0f;
        if (cw > 0.0f) {
            if (!light_visible(p, surface, chosen_point)) cw = 0.0f;
        }
        store_float4(os, V4{chosen_point.x, chosen_point.y, chosen_point.z, (float)chosen_light});
        store_float4(orc, V4{surface.position.x, surface.position.y, surface.position.z, cw});
        store_half4(on, V4{surface.normal.x, surface.normal.y, surface.normal.z, ccount + tcount});
    });
}

// cs_di_spatial (cs_di_spatial.sc:30-169).
void orc_di_spatial(const rgi_frame* f, const rgi_scene* s, const rgi_gbuffer* g, const rgi_di_reservoirs* res,
                    uint16_t* out_diffuse, uint16_t* out_specular)
{
    Pass p = make_di_pass(f, s, g);
    int64_t n = (int64_t)g->width * g->height;
    parallel_for(n, [&](int64_t i) {
        int x = (int)(i % g->width), y = (int)(i / g->width);
        V2 uv = V2{((float)x + 0.5f) * p.texel_size.x, ((float)y + 0.5f) * p.texel_size.y};
        uint16_t* od = out_diffuse + (size_t)i * 4;
        uint16_t* osp = out_specular + (size_t)i * 4;
        Surface surface = load_surface(p, uv);
        size_t t0 = texel(uv, g->width, g->height);
        V4 own_sample = float4(res->sample + t0 * 4);
        V4 own_receiver = float4(res->receiver + t0 * 4);
        V4 own_normal = half4(res->normal + t0 * 4);
        if (!surface.valid || f->di_light_count == 0) {
            store_half4(od, V4{0, 0, 0, 0});
            store_half4(osp, V4{0, 0, 0, 0});
            return;
        }
        uint32_t rng = pcg((uint32_t)x + pcg((uint32_t)y + pcg(f->frame_index ^ 0x9e3779b9u)));
        int taps = f->di_spatial_taps;
        float ccount = fmax_(fmin_(own_normal.w, 20.0f), 1.0f);
        int clight = own_sample.w >= 0.0f ? to_int(own_sample.w + 0.5f) : -1;
        V3 cpoint = xyz(own_sample);
        float ctarget = clight >= 0 && own_receiver.w > 0.0f ? light_target(p, surface, clight, cpoint) : 0.0f;
        float pair_canonical = ccount / (float)(taps > 1 ? taps : 1);
        float total = ccount;
        float share = taps > 0 ? 0.0f : ccount;
        float wsum = 0.0f;
        V3 chosen_point = cpoint;
        int chosen_light = -1;
        float chosen_target = 0.0f;
        bool neighbour_selected = false;
        float sc = scene_scale(p);
        for (int tap = 0; tap < taps; ++tap) {
            float ra = next_random(rng);
            float radius = 16.0f * sqrtf(ra);
            float rb = next_random(rng);
            float angle = (2.0f * kPi) * rb;
            float rsel = next_random(rng);
            float sa, ca;
            rgi_sincos(angle, &sa, &ca);
            V2 nuv = V2{uv.x + (ca * radius) * p.texel_size.x, uv.y + (sa * radius) * p.texel_size.y};
            Surface nb = load_surface(p, nuv);
            size_t tn = texel(nuv, g->width, g->height);
            V4 sd = float4(res->sample + tn * 4);
            V4 rc = float4(res->receiver + tn * 4);
            V4 nm = half4(res->normal + tn * 4);
            int light = sd.w >= 0.0f ? to_int(sd.w + 0.5f) : -1;
            if (nuv.x <= 0.0f || nuv.x >= 1.0f || nuv.y <= 0.0f || nuv.y >= 1.0f || !nb.valid ||
                !same_surface(p, surface.position, surface.normal, nb.position, nb.normal, 4.0f * (0.1f * sc))) {
                share += pair_canonical;
                continue;
            }
            float ncount = fmin_(nm.w, 20.0f);
            float pair_weight = ncount + pair_canonical;
            total += ncount;
            if (ctarget > 0.0f) {
                float ntarget = 0.0f;
                if (light_visible(p, nb, cpoint)) ntarget = light_target(p, nb, clight, cpoint);
                float cd = pair_canonical * ctarget;
                float den = ncount * ntarget + cd;
                share += g_plant == kPlantDiSpatialMis ? pair_weight : (pair_weight * cd) / den;
            }
            if (light >= 0 && light < f->di_light_count && rc.w > 0.0f) {
                float target = light_target(p, surface, light, xyz(sd));
                float ntarget = light_target(p, nb, light, xyz(sd));
                float nd = ncount * ntarget;
                float den = nd + pair_canonical * target;
                float w = den > 0.0f ? (((pair_weight * nd) / den) * target) * rc.w : 0.0f;
                wsum += w;
                if (w > 0.0f && rsel * wsum < w) {
                    chosen_point = xyz(sd);
                    chosen_light = light;
                    chosen_target = target;
                    neighbour_selected = true;
                }
            }
        }
        if (ctarget > 0.0f) {
            float w = (share * ctarget) * own_receiver.w;
            wsum += w;
            float rsel = next_random(rng);
            if (rsel * wsum < w) {
                chosen_point = cpoint;
                chosen_light = clight;
                chosen_target = ctarget;
                neighbour_selected = false;
            }
        }
        float cw = chosen_target > 0.0f ? wsum / (total * chosen_target) : 0.0f;
        if (cw > 0.0f && neighbour_selected) {
            if (!light_visible(p, surface, chosen_point)) cw = 0.0f;
        }
        V3 d, sp;
        evaluate_light_sample(p, surface, chosen_light, chosen_point, d, sp);
        d = d * cw;
        sp = sp * cw;
        store_half4(od, V4{d.x, d.y, d.z, 1.0f});
        store_half4(osp, V4{sp.x, sp.y, sp.z, 1.0f});
    });
}

// The displayed image of a pixel (GI re-modulated by the diffuse albedo, plus
// DI), false over the sky (fs_composite.sc's light, cs_error.sc's mask).
// The compared image at pixel i: the composite itself (`composited`), or the
// GI re-modulated by the albedo plus the DI planes; geometry pixels only.
static bool reference_image(const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds,
                            bool composited, size_t i, V3& L)
{
    if (g->depth[i] >= kSkyDepth) return false;
    V4 c = half4(color + i * 4);
    if (composited) {
        L = xyz(c);
        return true;
    }
    V4 a = unorm8(g->albedo[i]);
    float metallic = unorm8(g->normal[i]).w;
    L = V3{c.x, c.y, c.z} * (V3{a.x, a.y, a.z} * (1.0f - metallic));
    if (dd) L = L + xyz(half4(dd + i * 4));
    if (ds) L = L + xyz(half4(ds + i * 4));
    return true;
}

// fs_reference.sc:18-25: the running mean, here as an FP64 running sum.
static void reference_accumulate(const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds,
                                 bool composited, double* accum)
{
    size_t n = (size_t)g->width * g->height;
    for (size_t i = 0; i < n; ++i) {
        V3 L;
        if (!reference_image(g, color, dd, ds, composited, i, L)) continue;
        if (L.x != L.x || L.y != L.y || L.z != L.z) continue;
        accum[3 * i + 0] += (double)L.x;
        accum[3 * i + 1] += (double)L.y;
        accum[3 * i + 2] += (double)L.z;
    }
}

static double lum64(double r, double g, double b) { return (0.2126 * r + 0.7152 * g) + 0.0722 * b; }

// cs_error.sc:35-92 in FP64: 256-pixel blocks reduced by a halving tree, block
// partials summed with stride 256 per lane, then the same tree.
static void reference_error(const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds,
                            bool composited, const double* accum, uint32_t frames, double* out)
{
    const int B = 256, K = 5;
    size_t n = (size_t)g->width * g->height;
    size_t blocks = (n + B - 1) / B;
    double inv = 1.0 / (double)frames;
    std::vector<double> partial(blocks * K);
    auto tree = [&](std::vector<double>& s) {
        for (int stride = B / 2; stride > 0; stride >>= 1)
            for (int t = 0; t < stride; ++t)
                for (int k = 0; k < K; ++k) s[k * B + t] = s[k * B + t] + s[k * B + t + stride];
    };
    for (size_t b = 0; b < blocks; ++b) {
        std::vector<double> s(K * B, 0.0);
        for (int t = 0; t < B; ++t) {
            size_t i = b * B + t;
            V3 L;
            if (i >= n || !reference_image(g, color, dd, ds, composited, i, L)) continue;
            double rr = accum[3 * i] * inv, rg = accum[3 * i + 1] * inv, rb = accum[3 * i + 2] * inv;
            double lr = L.x, lg = L.y, lb = L.z;
            double li = lum64(lr, lg, lb), lref = lum64(rr, rg, rb);
            double d = li - lref;
            s[0 * B + t] = (d * d) / (lref * lref + 0.01);
            s[1 * B + t] = ((lr - rr) * (lr - rr) + (lg - rg) * (lg - rg)) + (lb - rb) * (lb - rb);
            s[2 * B + t] = li;
            s[3 * B + t] = lref;
            s[4 * B + t] = 1.0;
        }
        tree(s);
        for (int k = 0; k < K; ++k) partial[b * K + k] = s[k * B];
    }
    std::vector<double> s(K * B, 0.0);
    for (int t = 0; t < B; ++t)
        for (size_t b = (size_t)t; b < blocks; b += B)
            for (int k = 0; k < K; ++k) s[k * B + t] = s[k * B + t] + partial[b * K + k];
    tree(s);
    double count = s[4 * B];
    double nn = count > 1.0 ? count : 1.0;
    out[0] = s[0] / nn;
    out[1] = s[1 * B] / (3.0 * nn);
    out[2] = s[2 * B] / nn;
    out[3] = s[3 * B] / nn;
    out[4] = count;
}

void orc_reference_accumulate(const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds,
                              double* accum)
{
    reference_accumulate(g, color, dd, ds, false, accum);
}

void orc_reference_error(const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds,
                         const double* accum, uint32_t frames, double* out)
{
    reference_error(g, color, dd, ds, false, accum, frames, out);
}

void orc_reference_accumulate_composite(const rgi_gbuffer* g, const uint16_t* hdr, double* accum)
{
    reference_accumulate(g, hdr, nullptr, nullptr, true, accum);
}

void orc_reference_error_composite(const rgi_gbuffer* g, const uint16_t* hdr, const double* accum, uint32_t frames,
                                   double* out)
{
    reference_error(g, hdr, nullptr, nullptr, true, accum, frames, out);
}

// The G-buffer producer: one closest-hit primary ray per pixel centre.
// The shadow pass (renderer.cpp:566-596, fs_shadow.sc): the distance from
// the light to the nearest surface through each atlas texel centre, over the
// far plane, both faces of every triangle, clipped to [near, far]; 1 (the
// clear value) where nothing is hit. Texel (X, Y) is on face column + 3 row,
// column = X / F, row = Y / F from the top (from the bottom when the origin
// is bottom-left): shadowAtlasUv (shadow.sh:52-69) inverted at texel centres.
void orc_shadow_atlas(const rgi_shadow* sh, const rgi_scene* s, float* atlas)
{
    int F = sh->face_size, W = 3 * F, H = 2 * F;
    bool obl = sh->info[3] > 0.5f;
    V3 light = V3{sh->light_position[0], sh->light_position[1], sh->light_position[2]};
    parallel_for((int64_t)W * H, [&](int64_t i) {
        int X = (int)(i % W), Y = (int)(i / W);
        int column = X / F, texture_row = Y / F;
        int face = column + 3 * (obl ? 1 - texture_row : texture_row);
        float lx = ((float)(X - column * F) + 0.5f) / (float)F;
        float ly = ((float)(Y - texture_row * F) + 0.5f) / (float)F;
        V3 ndc = V3{lx * 2.0f - 1.0f, obl ? ly * 2.0f - 1.0f : 1.0f - ly * 2.0f, 1.0f};
        V3 far_point = clip_to_world(sh->inv_view_proj + face * 16, ndc);
        V3 to_far = far_point - light;
        float far_distance = length(to_far);
        V3 dir = to_far * (1.0f / far_distance);
        float near_distance = far_distance * (sh->near_plane / sh->info[0]);
        V3 origin = light + dir * near_distance;
        float dist;
        int tri = trace_bvh(s, origin, dir, far_distance - near_distance, false, dist);
        atlas[i] = tri >= 0 ? length((origin + dir * dist) - light) / sh->info[0] : 1.0f;
    });
}

void orc_gbuffer_primary(const rgi_frame* f, const rgi_scene* s, const float* materials, uint32_t material_count,
                         float* depth, uint32_t* albedo, uint32_t* normal, uint32_t* shading_normal, uint16_t* emission)
{
    int w = f->width, h = f->height;
    V2 ts = V2{1.0f / (float)w, 1.0f / (float)h};
    V3 eye = V3{f->camera_position[0], f->camera_position[1], f->camera_position[2]};
    parallel_for((int64_t)w * h, [&](int64_t i) {
        int x = (int)(i % w), y = (int)(i / w);
        V2 uv = V2{((float)x + 0.5f) * ts.x, ((float)y + 0.5f) * ts.y};
        V3 clip = V3{uv.x * 2.0f - 1.0f, to_clip_y(f, uv.y * 2.0f - 1.0f), 1.0f};
        V3 far_point = clip_to_world(f->inv_view_proj, clip);
        V3 dir = normalize(far_point - eye);
        float dist;
        int tri = trace_bvh(s, eye, dir, 1e30f, false, dist);
        float d = 1.0f;
        if (tri >= 0) {
            V3 hit = eye + dir * dist;
            V4 c = mul(f->view_proj, V4{hit.x, hit.y, hit.z, 1.0f});
            float z = c.z / c.w;
            if (f->homogeneous_depth) z = z * 0.5f + 0.5f;
            d = z;
            if (!(z >= 0.0f && z < kSkyDepth)) tri = -1;
        }
        if (tri < 0) {
            depth[i] = 1.0f;
            albedo[i] = 0u;
            normal[i] = 0u;
            shading_normal[i] = 0u;
            for (int c = 0; c < 4; ++c) emission[(size_t)i * 4 + c] = 0;
            return;
        }
        depth[i] = d;
        V3 v0 = bvh_vertex(s, tri, 0);
        V3 nn = normalize(cross(bvh_vertex(s, tri, 1) - v0, bvh_vertex(s, tri, 2) - v0));
        if (dot(nn, dir) > 0.0f) nn = -nn;
        size_t rec = (size_t)s->triangle_offset + (size_t)tri * 4;
        uint32_t material = (uint32_t)to_int(scene_vec4(s, rec).w);
        float roughness = 1.0f, metallic = 0.0f;
        if (material < material_count) {
            roughness = materials[material * 4 + 0];
            metallic = materials[material * 4 + 1];
        }
        auto byte = [](float v) -> uint32_t {
            float q = floorf(clampf(v, 0.0f, 1.0f) * 255.0f + 0.5f);
            return (uint32_t)to_int(q);
        };
        V3 a = bvh_triangle_albedo(s, tri);
        albedo[i] = byte(a.x) | (byte(a.y) << 8) | (byte(a.z) << 16) | (byte(roughness) << 24);
        uint32_t nw = byte(nn.x * 0.5f + 0.5f) | (byte(nn.y * 0.5f + 0.5f) << 8) | (byte(nn.z * 0.5f + 0.5f) << 16);
        normal[i] = nw | (byte(metallic) << 24);
        shading_normal[i] = nw;
        V3 e = bvh_triangle_emission(s, tri);
        // a = ambient occlusion: 1 without an occlusion map (fs_gbuffer.sc:81-93).
        store_half4(emission + (size_t)i * 4, V4{e.x, e.y, e.z, 1.0f});
    });
}

} // extern "C"

// The kernels' division by a constant c (device.h div255, div3): q = x *
// RN(1 / c) refined by one FMA step. Counts the non-negative floats with bit
// patterns lo, lo + stride, ... below hi where it differs from IEEE x / c.
extern "C" uint64_t orc_div_const_mismatches(float c, uint32_t lo, uint32_t hi, uint32_t stride)
{
    volatile float one = 1.0f;
    const float r = one / c;
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i += stride) {
        uint32_t u = (uint32_t)i;
        float x;
        memcpy(&x, &u, 4);
        float q = x * r;
        float refined = fmaf(fmaf(-c, q, x), r, q);
        float exact = x / c;
        if (memcmp(&refined, &exact, 4) != 0) ++bad;
    }
    return bad;
}
