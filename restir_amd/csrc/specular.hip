// Indirect specular GI on gfx950 (fs_specular.sc): one GGX visible-normal
// ray per GI pixel (Heitz 2018), traced like the GI rays (depth-buffer march,
// then the BVH, then the sky), its radiance weighted by F * G1(l),
// demodulated by the split-sum specular albedo and accumulated over 4 to 32
// frames by roughness.
//
// Launch shape as the GI temporal pass: 16 x 16 pixels per block, rays the
// depth buffer cannot answer compacted per wave into an LDS queue and traced
// by the first lanes of the wave.
#include "gi_device.h"
#include "batch.h"
#include "internal.h"

namespace rgi {
namespace {

struct SpecularLds {
    int ring[kStackRing * 256];
    float4 q_origin[256];
    float4 q_dir[256];
    int q_tri[256];
    float q_dist[256];
};

// sampleVisibleNormal (fs_specular.sc:51-68): a microfacet normal from the
// GGX distribution of normals visible from `view` (tangent frame, z up).
RGI_DEV F3 sample_visible_normal(F3 view, float alpha, float r0, float r1)
{
    F3 stretched = normalize(F3{alpha * view.x, alpha * view.y, view.z});
    float length_squared = stretched.x * stretched.x + stretched.y * stretched.y;
    F3 t1 = length_squared > 0.0f ? F3{-stretched.y, stretched.x, 0.0f} * rsqrt_(length_squared)
                                  : F3{1.0f, 0.0f, 0.0f};
    F3 t2 = cross(stretched, t1);
    float radius = sqrtf(r0);
    float angle = (2.0f * kPi) * r1;
    float sa, ca;
    rgi_sincos(angle, &sa, &ca);
    float p1 = radius * ca;
    float p2 = radius * sa;
    float s = 0.5f * (1.0f + stretched.z);
    p2 = (1.0f - s) * sqrtf(maxf_(1.0f - p1 * p1, 0.0f)) + s * p2;
    F3 normal = (t1 * p1 + t2 * p2) + stretched * sqrtf(maxf_((1.0f - p1 * p1) - p2 * p2, 0.0f));
    return normalize(F3{alpha * normal.x, alpha * normal.y, maxf_(normal.z, 0.0f)});
}

// smithG1 (fs_specular.sc:71-75).
RGI_DEV float smith_g1_ggx(float c, float alpha)
{
    float a2 = alpha * alpha;
    return (2.0f * c) / (c + sqrtf(a2 + ((1.0f - a2) * c) * c));
}

// reflect(i, n) = i - 2 dot(n, i) n.
RGI_DEV F3 reflect3(F3 i, F3 n) { return i - n * (2.0f * dot(n, i)); }

// hitRadiance (fs_specular.sc:78-124): radiance leaving a hit back along the
// ray: last frame's composite where the hit is a visible surface that was on
// screen, else its direct, bounce and surface light; the sky for sky hits.
RGI_DEV F3 hit_radiance(const GiPass& p, const uint16_t* previous_hdr, F3 hit, float id, F3 dir)
{
    const rgi_frame& f = *p.f;
    if (is_sky(id)) return sky_radiance(f, dir);
    F2 uv;
    float view_depth;
    bool visible_surface = false;
    if (project_to_screen(f, hit, uv, view_depth)) {
        float depth = fetch_depth(p, uv);
        visible_surface = depth < kSkyDepth && length(reconstruct_position(f, uv, depth) - hit) <= 0.1f * p.sc;
    }
    if (visible_surface) {
        F4 clip = mul(f.prev_view_proj, F4{hit.x, hit.y, hit.z, 1.0f});
        if (f.history_valid && clip.w > 0.0001f && previous_hdr != nullptr) {
            F2 previous_uv = clip_to_uv(f, clip);
            F4 previous = load_h4(previous_hdr, texel_index(previous_uv, p.g.width, p.g.height));
            if (on_screen(previous_uv) && previous.w > 0.5f) return xyz(previous);
        }
        F3 n = decode_normal(p, uv);
        F3 albedo = diffuse_albedo(p, uv);
        return (point_light_direct(p, albedo, hit, n) + previous_indirect(p, hit, n, albedo)) +
               fetch_light(p, uv) * f.lighting_surface;
    }
    int tri = to_int(fabsf(id) + 0.5f) - 1;
    if (tri < 0 || !bvh_ready(p.v)) return splat3(0.0f);
    F3 v0 = vertex(p.v, tri, 0);
    F3 n = normalize(cross(vertex(p.v, tri, 1) - v0, vertex(p.v, tri, 2) - v0));
    if (dot(n, -dir) < 0.0f) n = -n;
    F3 albedo = tri_albedo(p.v, tri);
    return (point_light_direct(p, albedo, hit, n) + previous_indirect(p, hit, n, albedo)) +
           tri_emission(p.v, tri) * f.lighting_emission;
}

// fs_specular.sc:127-226 (kSameSize: a G-buffer of the GI size, as in the GI
// passes).
template <bool kSameSize>
RGI_DEV void specular_body(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                           const uint16_t* previous_hdr, const uint16_t* history, uint16_t* out, F2 texel,
                           const TileMap& tiles, const BlockAt& at, SpecularLds& lds)
{
    GiPass p;
    p.f = &f;
    p.v = scene_view(s);
    p.g = g;
    p.in = in;
    set_gi_size<kSameSize>(p, f, g, texel);
    p.grid_camera = F3{f.camera_position[0], f.camera_position[1], f.camera_position[2]};
    p.sc = f.scene_scale;

    int x, y;
    if (!group_pixel(tiles, at, x, y)) return;
    uint32_t lane = threadIdx.x & 63u, qbase = threadIdx.x & ~63u;
    bool inside = x < p.hw && y < p.hh;
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    float depth = inside ? fetch_depth(p, uv) : 1.0f;
    bool working = inside && depth < kSkyDepth;
    // A wave with no surface pixel writes zeros and leaves (the rest of the
    // pass is wave-local).
    if (__ballot(working) == 0ull) {
        if (inside) store_h4(out, (uint32_t)y * (uint32_t)p.hw + (uint32_t)x, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    F3 position = reconstruct_position(f, uv, depth);
    F3 n = decode_normal(p, uv);
    uint32_t t = tix(p, uv);
    F4 albedo_roughness = unorm8(g.albedo[t]);
    float metallic = unorm8(g.normal[t]).w;
    float roughness = maxf_(albedo_roughness.w, 0.02f);
    float alpha = roughness * roughness;
    F3 camera = F3{f.camera_position[0], f.camera_position[1], f.camera_position[2]};
    F3 view = normalize(camera - position);
    float normal_view = maxf_(dot(n, view), 0.001f);
    F3 f0 = mix3(splat3(0.04f), xyz(albedo_roughness), metallic);
    float gx = floorf(uv.x * p.size.x), gy = floorf(uv.y * p.size.y);
    uint32_t rng = pcg((uint32_t)to_int(gx) + pcg((uint32_t)to_int(gy) + pcg(f.frame_index ^ 0x2545f491u)));

    F3 tangent, bitangent;
    tangent_frame(n, tangent, bitangent);
    F3 local_view = F3{dot(view, tangent), dot(view, bitangent), normal_view};
    float r0 = next_random(rng);
    float r1 = next_random(rng);
    F3 lh = sample_visible_normal(local_view, alpha, r0, r1);
    F3 halfway = (tangent * lh.x + bitangent * lh.y) + n * lh.z;
    F3 direction = reflect3(-view, halfway);
    float normal_light = dot(n, direction);
    bool active = working && normal_light > 0.0f;

    F3 origin = position + n * (0.02f * p.sc);
    F3 hit_position = splat3(0.0f);
    float hit_id = 0.0f;
    bool hit = false;
    if (active && f.ray_mode != RGI_RAY_BVH) hit = march(p, origin, direction, 3.6f * p.sc, hit_position);
    bool need = active && !hit && f.ray_mode != RGI_RAY_SCREEN;
    uint64_t queue = __ballot(need);
    if (queue != 0ull) {
        uint32_t slot = rank_in(queue);
        if (need) {
            lds.q_origin[qbase + slot] = make_float4(origin.x, origin.y, origin.z, 0.0f);
            lds.q_dir[qbase + slot] = make_float4(direction.x, direction.y, direction.z, 0.0f);
        }
        wave_sync();
        if (lane < (uint32_t)__popcll(queue)) {
            float4 qo = lds.q_origin[qbase + lane], qd = lds.q_dir[qbase + lane];
            float dist;
            int tri = trace_closest(p.v, F3{qo.x, qo.y, qo.z}, F3{qd.x, qd.y, qd.z}, 4.0f * p.sc, dist,
                                    lds.ring + threadIdx.x, 256);
            lds.q_tri[qbase + lane] = tri;
            lds.q_dist[qbase + lane] = dist;
        }
        wave_sync();
        if (need) {
            int tri = lds.q_tri[qbase + slot];
            float dist = lds.q_dist[qbase + slot];
            if (tri >= 0) {
                hit = true;
                hit_position = origin + direction * dist;
                hit_id = sample_id_of(p.v, tri, direction);
            } else if (sky_on(f)) {
                hit = true;
                hit_id = kSkySampleId;
            }
        }
    }
    F3 estimate = splat3(0.0f);
    if (active && hit) {
        F3 fresnel = fresnel_schlick(halfway, view, f0);
        estimate = (hit_radiance(p, previous_hdr, hit_position, hit_id, direction) * fresnel) *
                   smith_g1_ggx(normal_light, alpha);
    }
    if (!inside) return;
    uint32_t o = (uint32_t)y * (uint32_t)p.hw + (uint32_t)x;
    if (!working) {
        store_h4(out, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    F3 demodulated = estimate / max3(specular_albedo(f0, roughness, normal_view), splat3(0.001f));
    if (f.outlier_clamp > 0.0f) demodulated = min3(demodulated, splat3(f.outlier_clamp));

    // Last frame's result at this surface: the nearest same-surface texel.
    F3 previous = splat3(0.0f);
    float history_length = 0.0f;
    F4 clip_previous = mul(f.prev_view_proj, F4{position.x, position.y, position.z, 1.0f});
    if (f.history_valid && clip_previous.w > 0.0001f && history != nullptr) {
        F2 cuv = clip_to_uv(f, clip_previous);
        F2 puv = F2{(floorf(cuv.x * p.size.x) + 0.5f) * texel.x, (floorf(cuv.y * p.size.y) + 0.5f) * texel.y};
        if (on_screen(puv) && same_history_surface(p, puv, position, n)) {
            F4 h = fetch_h4(history, p, puv);
            previous = xyz(h);
            history_length = h.w;
        }
    }
    float max_history = mixf_(4.0f, 32.0f, clampf_(roughness * 1.5f, 0.0f, 1.0f));
    history_length = minf_(history_length + 1.0f, max_history);
    F3 color = mix3(previous, demodulated, 1.0f / history_length);
    store_h4(out, o, F4{color.x, color.y, color.z, history_length});
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) k_specular(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in,
                                                  const uint16_t* previous_hdr, const uint16_t* history,
                                                  uint16_t* out, F2 texel, TileMap tiles)
{
    __shared__ SpecularLds lds;
    specular_body<kSameSize>(f, s, g, in, previous_hdr, history, out, texel, tiles, this_block(), lds);
}

// The batched form: workgroups [first[v], first[v + 1]) run view v's tiles
// (its GI spatial tile map), for the views whose G-buffer size matches
// kSameSize.
template <bool kSameSize>
__global__ void __launch_bounds__(256) k_specular_batch(const BatchView* __restrict__ views,
                                                        const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ SpecularLds lds;
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    if (v.same != (kSameSize ? 1u : 0u)) return; // the whole workgroup, before the barriers
    specular_body<kSameSize>(v.fg, v.s, v.g, v.sin, v.spec_previous, v.spec_history, v.spec_out, v.gi_texel,
                             v.gi_tiles, BlockAt{b, v.gi_gx, v.gi_gy}, lds);
}

} // namespace

int specular_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, bool same,
                   bool larger, hipStream_t stream)
{
    if (blocks == 0u) return RGI_OK;
    if (same) hipLaunchKernelGGL(k_specular_batch<true>, dim3(blocks), dim3(256), 0, stream, views, first, count);
    if (larger) hipLaunchKernelGGL(k_specular_batch<false>, dim3(blocks), dim3(256), 0, stream, views, first, count);
    return check_launch("specular gi (batch)");
}

int specular_gi(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                const uint16_t* previous_hdr, const uint16_t* history, uint16_t* out, const Tuning& tuning,
                hipStream_t stream)
{
    TileMap tiles = tile_map(tuning.tile_order, tuning.tile_shape, f.gi_width, f.gi_height);
    F2 texel = F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height};
    const bool same = g.width == f.gi_width && g.height == f.gi_height;
    hipLaunchKernelGGL(same ? k_specular<true> : k_specular<false>, tile_grid(tiles), dim3(256), 0, stream, f, s, g, in,
                       previous_hdr, history, out, texel, tiles);
    return check_launch("specular gi");
}

} // namespace rgi
