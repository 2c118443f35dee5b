// ReSTIR DI passes on gfx950 (di.sh, brdf.sh, cs_di_temporal.sc,
// cs_di_spatial.sc) at full resolution: fresh light candidates by power with
// the generalized balance heuristic against last frame's reservoir, then
// pairwise-MIS spatial reuse and shading. Shadow rays are any-hit walks of
// the BVH. One pixel per lane, 16 x 16 pixels per 256-thread block.
#include "bvh_device.h"
#include "shading.h"
#include "batch.h"
#include "internal.h"

namespace rgi {
namespace {

struct DiPass {
    const rgi_frame* f;
    SceneView v;
    rgi_gbuffer g;
    const float4* lights;
    F2 texel;
    float sc;
    const uint32_t* guide;    // the CDF's guide table (k_light_guide), or NULL
    uint32_t guide_buckets;   // its K, a power of two <= 2^16
};

struct Surf {
    F3 position, normal, surface_normal, albedo;
    float roughness, metallic;
    F3 eye;
    bool valid;
};

RGI_DEV Surf load_surface(const DiPass& p, F2 uv)
{
    Surf s;
    uint32_t t = texel_index(uv, p.g.width, p.g.height);
    float depth = p.g.depth[t];
    F4 ar = unorm8(p.g.albedo[t]);
    F4 nm = unorm8(p.g.normal[t]);
    F4 sn = unorm8(p.g.shading_normal[t]);
    s.valid = depth < kSkyDepth;
    s.position = reconstruct_position(*p.f, uv, minf_(depth, kSkyDepth));
    s.surface_normal = normalize(F3{nm.x * 2.0f - 1.0f, nm.y * 2.0f - 1.0f, nm.z * 2.0f - 1.0f});
    s.normal = normalize(F3{sn.x * 2.0f - 1.0f, sn.y * 2.0f - 1.0f, sn.z * 2.0f - 1.0f});
    s.albedo = F3{ar.x, ar.y, ar.z};
    s.roughness = ar.w;
    s.metallic = nm.w;
    s.eye = F3{p.f->camera_position[0], p.f->camera_position[1], p.f->camera_position[2]};
    return s;
}

RGI_DEV F4 light_row(const DiPass& p, int light, int row)
{
    float4 q = p.lights[light * 4 + row];
    return F4{q.x, q.y, q.z, q.w};
}

RGI_DEV bool point_light(const DiPass& p, int light) { return light_row(p, light, 0).w > 0.5f; }

// Power-proportional choice: binary search of the CDF (di.sh:123-142), for
// the first light whose CDF value reaches `random`. With a guide table the
// search starts from the bracket [guide[k], guide[k + 1]], k = floor(random
// K), which holds that light (k_light_guide), and so ends on the same light
// in fewer steps. `random` is a nextRandom value, a multiple of 2^-24, so
// random K is exact.
RGI_DEV int pick_light(const DiPass& p, float random, float& pmf)
{
    int low = 0, high = p.f->di_light_count - 1;
    if (p.guide) {
        uint32_t k = (uint32_t)(random * (float)p.guide_buckets);
        low = (int)p.guide[k];
        high = (int)p.guide[k + 1u];
    }
    for (int step = 0; step < 24 && low < high; ++step) {
        int middle = (low + high) >> 1;
        if (light_row(p, middle, 2).w < random) low = middle + 1;
        else high = middle;
    }
    pmf = light_row(p, low, 3).w;
    return low;
}

RGI_DEV float light_area(const DiPass& p, int light)
{
    F3 v0 = xyz(light_row(p, light, 0));
    return 0.5f * length(cross(xyz(light_row(p, light, 1)) - v0, xyz(light_row(p, light, 2)) - v0));
}

RGI_DEV F3 light_point(const DiPass& p, int light, float r0, float r1, float& density)
{
    F3 v0 = xyz(light_row(p, light, 0));
    if (point_light(p, light)) {
        density = 1.0f;
        return v0;
    }
    float root = sqrtf(r0);
    float u = 1.0f - root;
    float v = r1 * root;
    density = 1.0f / maxf_(light_area(p, light), 1e-12f);
    return (v0 + (xyz(light_row(p, light, 1)) - v0) * u) + (xyz(light_row(p, light, 2)) - v0) * v;
}

RGI_DEV void light_sample(const DiPass& p, const Surf& s, int light, F3 point, F3& diffuse, F3& specular)
{
    diffuse = splat3(0.0f);
    specular = splat3(0.0f);
    if (light < 0 || light >= p.f->di_light_count) return;
    F3 to_light = point - s.position;
    float d2 = maxf_(dot(to_light, to_light), (0.0001f * p.sc) * p.sc);
    F3 dir = to_light * rsqrt_(d2);
    if (dot(s.surface_normal, dir) <= 0.0f) return;
    F3 radiance = xyz(light_row(p, light, 3)) / d2;
    if (!point_light(p, light)) {
        F3 v0 = xyz(light_row(p, light, 0));
        F3 ln = normalize(cross(xyz(light_row(p, light, 1)) - v0, xyz(light_row(p, light, 2)) - v0));
        radiance = radiance * fabsf(dot(ln, dir));
    }
    F3 d, sp;
    brdf(s.normal, normalize(s.eye - s.position), dir, s.albedo, s.roughness, s.metallic, d, sp);
    diffuse = d * radiance;
    specular = sp * radiance;
}

RGI_DEV float light_target(const DiPass& p, const Surf& s, int light, F3 point)
{
    F3 d, sp;
    light_sample(p, s, light, point, d, sp);
    return luminance(d + sp);
}

RGI_DEV bool light_visible(const DiPass& p, const Surf& s, F3 point)
{
    float bias = 0.01f * p.sc;
    F3 origin = s.position + s.surface_normal * bias;
    F3 to_point = point - origin;
    float pd = length(to_point);
    return trace_any(p.v, origin, to_point / maxf_(pd, 0.0001f), pd * 0.999f - bias) < 0;
}

RGI_DEV bool same_surface(const DiPass& p, F3 position, F3 n, F3 other_position, F3 other_normal, float radius)
{
    F3 off = other_position - position;
    return dot(other_normal, other_normal) > 0.5f && dot(off, off) <= radius * radius &&
           fabsf(dot(off, n)) <= 0.03f * p.sc && dot(normalize(other_normal), n) >= 0.9f;
}

RGI_DEV bool block_pixel(const TileMap& tiles, const BlockAt& at, int& x, int& y)
{
    uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, tx, ty;
    if (!block_tile(tiles, at, tx, ty)) return false;
    if (tiles.wide) {
        x = (int)(tx * 32u + wave * 8u + (lane & 7u));
        y = (int)(ty * 8u + (lane >> 3));
    } else {
        x = (int)(tx * 16u + (wave & 1u) * 8u + (lane & 7u));
        y = (int)(ty * 16u + (wave >> 1) * 8u + (lane >> 3));
    }
    return true;
}

RGI_DEV DiPass make_pass(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, F2 texel)
{
    DiPass p;
    p.f = &f;
    p.v = scene_view(s);
    p.g = g;
    p.lights = reinterpret_cast<const float4*>(s.lights);
    p.texel = texel;
    p.sc = f.scene_scale;
    p.guide = nullptr;
    p.guide_buckets = 0u;
    return p;
}

// The CDF's guide table (the cutpoint method of Chen and Asau, 1974): entry k
// of K + 1 is the first light whose CDF value reaches k / K (the last light
// when none does), found by the same search as pick_light. A random r in
// [k / K, (k + 1) / K) then picks a light between entries k and k + 1: every
// light before entry k has a CDF value below k / K <= r, and entry k + 1's
// reaches (k + 1) / K > r (or is the last light). K is a power of two, so k /
// K is exact.
__global__ void __launch_bounds__(256) k_light_guide(const float4* lights, int n, uint32_t buckets, uint32_t* guide)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > buckets) return;
    float t = (float)k / (float)buckets;
    int low = 0, high = n - 1;
    for (int step = 0; step < 24 && low < high; ++step) {
        int middle = (low + high) >> 1;
        if (lights[middle * 4 + 2].w < t) low = middle + 1;
        else high = middle;
    }
    guide[k] = (uint32_t)low;
}

// Light records staged in LDS per workgroup when the scene's fit: the light
// choice's binary search over the CDF and every light evaluation then read
// LDS instead of a chain of dependent global loads. The records are the
// same words, so results do not change.
constexpr uint32_t kLdsLights = 256; // 16 KB

RGI_DEV void stage_lights(DiPass& p, const rgi_frame& f, const rgi_scene& s, float4* lds)
{
    const uint32_t n = s.light_count;
    if (f.di_light_count == 0 || n == 0 || n > kLdsLights) return; // the same for the whole workgroup
    const float4* src = reinterpret_cast<const float4*>(s.lights);
    for (uint32_t i = threadIdx.x; i < 4u * n; i += blockDim.x) lds[i] = src[i];
    __syncthreads();
    p.lights = lds;
}

// One workgroup of DI temporal: the tile at `at` of the view's grid.
RGI_DEV void di_temporal_body(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_di_reservoirs& prev,
                              const rgi_di_reservoirs& out, F2 texel, const TileMap& tiles, const BlockAt& at,
                              const uint32_t* guide, uint32_t guide_buckets, float4* lights)
{
    DiPass p = make_pass(f, s, g, texel);
    p.guide = guide;
    p.guide_buckets = guide_buckets;
    int x, y;
    if (!block_pixel(tiles, at, x, y)) return; // the whole workgroup, before the barrier
    stage_lights(p, f, s, lights);
    if (x >= g.width || y >= g.height) return;
    uint32_t o = (uint32_t)y * (uint32_t)g.width + (uint32_t)x;
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    Surf surface = load_surface(p, uv);
    if (!surface.valid || f.di_light_count == 0) {
        store_f4(out.sample, o, F4{0.0f, 0.0f, 0.0f, -1.0f});
        store_f4(out.receiver, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        store_h4(out.normal, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    uint32_t rng = pcg((uint32_t)x + pcg((uint32_t)y + pcg(f.frame_index ^ 0x3c6ef372u)));
    int candidates = f.di_candidates;
    float wsum = 0.0f;
    F3 chosen_point = splat3(0.0f);
    int chosen_light = -1;
    float chosen_target = 0.0f;
    for (int c = 0; c < candidates; ++c) {
        float pmf;
        float rl = next_random(rng);
        int light = pick_light(p, rl, pmf);
        float density;
        float r0 = next_random(rng);
        float r1 = next_random(rng);
        F3 lp = light_point(p, light, r0, r1, density);
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
    F3 cpoint = chosen_point;
    int clight = chosen_light;
    float ctarget = chosen_target;

    float tcount = 0.0f, tweight = 0.0f;
    F3 tpoint = splat3(0.0f);
    int tlight = -1;
    Surf previous = surface;
    F4 clip = mul(f.prev_view_proj, F4{surface.position.x, surface.position.y, surface.position.z, 1.0f});
    if (f.di_history_valid && clip.w > 0.0001f) {
        F2 puv = F2{(clip.x / clip.w) * 0.5f + 0.5f, ((f.origin_bottom_left ? clip.y : -clip.y) / clip.w) * 0.5f + 0.5f};
        if (puv.x > 0.0f && puv.x < 1.0f && puv.y > 0.0f && puv.y < 1.0f) {
            uint32_t t = texel_index(puv, g.width, g.height);
            F4 sd = load_f4(prev.sample, t);
            F4 rc = load_f4(prev.receiver, t);
            F4 nm = load_h4(prev.normal, t);
            int light = to_int(sd.w + 0.5f);
            if (sd.w >= 0.0f && light < f.di_light_count &&
                same_surface(p, surface.position, surface.normal, xyz(rc), xyz(nm), 0.1f * p.sc)) {
                tcount = minf_(nm.w, 20.0f);
                tweight = rc.w;
                tpoint = xyz(sd);
                tlight = light;
                previous.position = xyz(rc);
                previous.normal = normalize(xyz(nm));
                previous.eye = F3{f.prev_camera_position[0], f.prev_camera_position[1], f.prev_camera_position[2]};
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
    store_f4(out.sample, o, F4{chosen_point.x, chosen_point.y, chosen_point.z, (float)chosen_light});
    store_f4(out.receiver, o, F4{surface.position.x, surface.position.y, surface.position.z, cw});
    store_h4(out.normal, o, F4{surface.normal.x, surface.normal.y, surface.normal.z, ccount + tcount});
}

// One workgroup of DI spatial: the tile at `at` of the view's grid.
RGI_DEV void di_spatial_body(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_di_reservoirs& res,
                             uint16_t* out_diffuse, uint16_t* out_specular, F2 texel, const TileMap& tiles,
                             const BlockAt& at, float4* lights)
{
    DiPass p = make_pass(f, s, g, texel);
    int x, y;
    if (!block_pixel(tiles, at, x, y)) return; // the whole workgroup, before the barrier
    stage_lights(p, f, s, lights);
    if (x >= g.width || y >= g.height) return;
    uint32_t o = (uint32_t)y * (uint32_t)g.width + (uint32_t)x;
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    Surf surface = load_surface(p, uv);
    uint32_t t0 = texel_index(uv, g.width, g.height);
    F4 own_sample = load_f4(res.sample, t0);
    F4 own_receiver = load_f4(res.receiver, t0);
    F4 own_normal = load_h4(res.normal, t0);
    if (!surface.valid || f.di_light_count == 0) {
        store_h4(out_diffuse, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        store_h4(out_specular, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    uint32_t rng = pcg((uint32_t)x + pcg((uint32_t)y + pcg(f.frame_index ^ 0x9e3779b9u)));
    int taps = f.di_spatial_taps;
    float ccount = maxf_(minf_(own_normal.w, 20.0f), 1.0f);
    int clight = own_sample.w >= 0.0f ? to_int(own_sample.w + 0.5f) : -1;
    F3 cpoint = xyz(own_sample);
    float ctarget = clight >= 0 && own_receiver.w > 0.0f ? light_target(p, surface, clight, cpoint) : 0.0f;
    float pair_c = ccount / (float)(taps > 1 ? taps : 1);
    float total = ccount;
    float share = taps > 0 ? 0.0f : ccount;
    float wsum = 0.0f;
    F3 chosen_point = cpoint;
    int chosen_light = -1;
    float chosen_target = 0.0f;
    bool neighbour_selected = false;
    for (int tap = 0; tap < taps; ++tap) {
        float ra = next_random(rng);
        float radius = 16.0f * sqrtf(ra);
        float rb = next_random(rng);
        float angle = (2.0f * kPi) * rb;
        float rsel = next_random(rng);
        float sa, ca;
        rgi_sincos(angle, &sa, &ca);
        F2 nuv = F2{uv.x + (ca * radius) * texel.x, uv.y + (sa * radius) * texel.y};
        Surf nb = load_surface(p, nuv);
        uint32_t tn = texel_index(nuv, g.width, g.height);
        F4 sd = load_f4(res.sample, tn);
        F4 rc = load_f4(res.receiver, tn);
        F4 nm = load_h4(res.normal, tn);
        int light = sd.w >= 0.0f ? to_int(sd.w + 0.5f) : -1;
        if (nuv.x <= 0.0f || nuv.x >= 1.0f || nuv.y <= 0.0f || nuv.y >= 1.0f || !nb.valid ||
            !same_surface(p, surface.position, surface.normal, nb.position, nb.normal, 4.0f * (0.1f * p.sc))) {
            share += pair_c;
            continue;
        }
        float ncount = minf_(nm.w, 20.0f);
        float pair_weight = ncount + pair_c;
        total += ncount;
        if (ctarget > 0.0f) {
            float ntarget = 0.0f;
            if (light_visible(p, nb, cpoint)) ntarget = light_target(p, nb, clight, cpoint);
            float cd = pair_c * ctarget;
            float den = ncount * ntarget + cd;
            share += (pair_weight * cd) / den;
        }
        if (light >= 0 && light < f.di_light_count && rc.w > 0.0f) {
            float target = light_target(p, surface, light, xyz(sd));
            float ntarget = light_target(p, nb, light, xyz(sd));
            float nd = ncount * ntarget;
            float den = nd + pair_c * target;
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
    F3 d, sp;
    light_sample(p, surface, chosen_light, chosen_point, d, sp);
    d = d * cw;
    sp = sp * cw;
    store_h4(out_diffuse, o, F4{d.x, d.y, d.z, 1.0f});
    store_h4(out_specular, o, F4{sp.x, sp.y, sp.z, 1.0f});
}

__global__ void __launch_bounds__(256) k_di_temporal(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_di_reservoirs prev,
                                                     rgi_di_reservoirs out, F2 texel, TileMap tiles,
                                                     const uint32_t* guide, uint32_t guide_buckets)
{
    __shared__ float4 lights[4 * kLdsLights];
    di_temporal_body(f, s, g, prev, out, texel, tiles, this_block(), guide, guide_buckets, lights);
}

__global__ void __launch_bounds__(256) k_di_spatial(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_di_reservoirs res,
                                                    uint16_t* out_diffuse, uint16_t* out_specular, F2 texel,
                                                    TileMap tiles)
{
    __shared__ float4 lights[4 * kLdsLights];
    di_spatial_body(f, s, g, res, out_diffuse, out_specular, texel, tiles, this_block(), lights);
}

// The batched forms: workgroups [first[v], first[v + 1]) run view v's grid.
__global__ void __launch_bounds__(256) k_di_temporal_batch(const BatchView* views, const uint32_t* first, uint32_t count)
{
    __shared__ float4 lights[4 * kLdsLights];
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    di_temporal_body(v.f, v.s, v.g, v.di_prev, v.di_next, v.texel, v.di_tiles, BlockAt{b, v.di_gx, v.di_gy}, v.guide,
                     v.guide_buckets, lights);
}

// At most 128 VGPRs, as the compiler gives the single-view kernel: 4 waves
// per SIMD.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) k_di_spatial_batch(const BatchView* views, const uint32_t* first, uint32_t count)
{
    __shared__ float4 lights[4 * kLdsLights];
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    di_spatial_body(v.f, v.s, v.g, v.di_next, v.di_diffuse, v.di_specular, v.texel, v.di_tiles,
                    BlockAt{b, v.di_gx, v.di_gy}, lights);
}

} // namespace

int di_temporal(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                const rgi_di_reservoirs& prev, const rgi_di_reservoirs& out, const Tuning& tuning, hipStream_t stream,
                bool fixed_lights)
{
    TileMap tiles = tile_map(tuning.tile_order, tuning.tile_shape, g.width, g.height);
    F2 texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
    // More lights than the workgroups stage in LDS: the light choice searches
    // the CDF in memory, so it starts from a guide table built here (with a
    // context to hold it).
    uint32_t* guide = nullptr;
    uint32_t buckets = 0u;
    const int e = di_light_guide(ctx, f, s, &guide, &buckets, stream, fixed_lights);
    if (e != RGI_OK) return e;
    hipLaunchKernelGGL(k_di_temporal, tile_grid(tiles), dim3(256), 0, stream, f, s, g, prev, out,
                       texel, tiles, guide, buckets);
    return check_launch("di temporal");
}

int di_spatial(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_di_reservoirs& res,
               uint16_t* diffuse, uint16_t* specular, const Tuning& tuning, hipStream_t stream)
{
    TileMap tiles = tile_map(tuning.tile_order, tuning.tile_shape, g.width, g.height);
    F2 texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
    hipLaunchKernelGGL(k_di_spatial, tile_grid(tiles), dim3(256), 0, stream, f, s, g, res,
                       diffuse, specular, texel, tiles);
    return check_launch("di spatial");
}

int di_light_guide(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, uint32_t** guide, uint32_t* buckets,
                   hipStream_t stream, bool fixed_lights)
{
    *guide = nullptr;
    *buckets = 0u;
    const int n = f.di_light_count;
    if (!ctx || n <= (int)kLdsLights || n > (1 << 24)) return RGI_OK;
    uint32_t k = 1024u;
    while (k < (uint32_t)n && k < (1u << 16)) k <<= 1;
    *guide = static_cast<uint32_t*>(context_buffer(ctx, kBufLightGuide, (k + 1u) * sizeof(uint32_t)));
    if (!*guide) return RGI_E_NOMEM;
    *buckets = k;
    // The bucket count follows from n, so the same table and n find the
    // same guide in the same buffer.
    if (fixed_lights && ctx->guide_lights == s.lights && ctx->guide_count == n && ctx->guide_buffer == *guide)
        return RGI_OK;
    hipLaunchKernelGGL(k_light_guide, dim3((k + 256u) / 256u), dim3(256), 0, stream,
                       reinterpret_cast<const float4*>(s.lights), n, k, *guide);
    ctx->guide_lights = fixed_lights ? s.lights : nullptr;
    ctx->guide_count = n;
    ctx->guide_buffer = *guide;
    return check_launch("light guide");
}

int di_temporal_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, hipStream_t stream)
{
    if (blocks == 0u) return RGI_OK;
    hipLaunchKernelGGL(k_di_temporal_batch, dim3(blocks), dim3(256), 0, stream, views, first, count);
    return check_launch("di temporal (batch)");
}

int di_spatial_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, hipStream_t stream)
{
    if (blocks == 0u) return RGI_OK;
    hipLaunchKernelGGL(k_di_spatial_batch, dim3(blocks), dim3(256), 0, stream, views, first, count);
    return check_launch("di spatial (batch)");
}

} // namespace rgi
