// ReSTIR GI passes on gfx950: temporal (fresh candidates + reuse from last
// frame's grid + boiling filter) and spatial (reuse from this frame's grid,
// shading, temporal accumulation).
//
// Launch shape: 256 threads cover 16 x 16 pixels; each wave64 covers 8 x 8,
// i.e. two of the reference's 4 x 8 groups side by side (compute.sh:33-35),
// in the half-waves [0, 32) and [32, 64). The boiling filter's group mean
// (groupSum, compute.sh:376-386) is the same tree of additions done with
// 32-wide shuffles in each half-wave, so its rounding matches.
//
// Rays the depth buffer cannot answer are compacted per wave (ballot + mbcnt)
// into an LDS queue and traced by the first lanes of the wave, as the
// reference's group queue does (cs_restir_temporal.sc:108-149); which lane
// traces a ray does not change its result.
#include "gi_device.h"
#include "internal.h"

namespace rgi {
namespace {

// Per-block LDS of the temporal pass: the closest-hit stack rings and the
// per-wave ray queue.
struct TemporalLds {
    int ring[kStackRing * 256];
    float4 q_origin[256];
    float4 q_dir[256];
    int q_tri[256];
    float q_dist[256];
};

// Source density of a candidate hit, converted from the ray origin to the
// receiver through the area measure at the hit (cs_restir_temporal.sc:163-184),
// and its stream into the canonical reservoir.
RGI_DEV void stream_candidate(const GiPass& p, F3 position, F3 n, F3 albedo, F3 ray_origin, F3 dir, bool hit,
                              F3 hit_position, float hit_id, float rsel, Res& canonical)
{
    F3 contribution = splat3(0.0f), hit_normal = splat3(0.0f);
    bool evaluated = false;
    if (hit) evaluated = shade_sample(p, position, n, albedo, hit_position, hit_id, contribution, hit_normal);
    if (evaluated) {
        float target = luminance(contribution);
        float source_pdf = dot(n, dir) / kPi;
        if (!is_sky(hit_id)) {
            F3 from_receiver = hit_position - position;
            F3 from_origin = hit_position - ray_origin;
            float lim = ((0.0001f * p.sc) * p.sc) * 0.01f;
            float rs = maxf_(dot(from_receiver, from_receiver), lim);
            float os = maxf_(dot(from_origin, from_origin), lim);
            float hco = maxf_(-dot(hit_normal, from_origin), 0.0f) * rsqrt_(os);
            float hcr = -dot(hit_normal, from_receiver) * rsqrt_(rs);
            source_pdf = hcr > 0.0001f ? (source_pdf * (hco / os)) * (rs / hcr) : 0.0f;
        }
        if (target > 0.0f && source_pdf > 0.0001f)
            stream_in(canonical, contribution, hit_position, hit_normal, hit_id, target, target / source_pdf, rsel);
    }
}

template <bool kSameSize>
RGI_DEV void gi_temporal_body(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                              const rgi_gi_reservoirs& out, F2 texel, const TileMap& tiles, TemporalLds& lds)
{
    int* ring = lds.ring;
    float4* q_origin = lds.q_origin;
    float4* q_dir = lds.q_dir;
    int* q_tri = lds.q_tri;
    float* q_dist = lds.q_dist;

    GiPass p;
    p.f = &f;
    p.v = scene_view(s);
    p.g = g;
    p.in = in;
    set_gi_size<kSameSize>(p, f, g, texel);
    p.grid_camera = F3{f.prev_camera_position[0], f.prev_camera_position[1], f.prev_camera_position[2]};
    p.sc = f.scene_scale;

    int x, y;
    if (!group_pixel(tiles, x, y)) return;
    uint32_t lane = threadIdx.x & 63u, qbase = threadIdx.x & ~63u;
    bool inside = x < p.hw && y < p.hh;
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    float depth = inside ? fetch_depth(p, uv) : 1.0f;
    bool working = inside && depth < kSkyDepth;
    // A wave with no surface pixel writes its empty reservoirs and leaves
    // (the rest of the pass is wave-local: ballots, the LDS queue and the
    // 4 x 8 group sums never reach another wave).
    if (__ballot(working) == 0ull) {
        if (inside) {
            uint32_t o = (uint32_t)y * (uint32_t)p.hw + (uint32_t)x;
            F4 zero = F4{0.0f, 0.0f, 0.0f, 0.0f};
            store_f4(out.sample, o, zero);
            store_f4(out.receiver, o, zero);
            store_f4(out.normal, o, zero);
        }
        return;
    }
    F3 position = reconstruct_position(f, uv, depth);
    F3 n = decode_normal(p, uv);
    F3 albedo = diffuse_albedo(p, uv);
    F3 ray_origin = position + n * (0.02f * p.sc);
    float gx = floorf(uv.x * p.size.x), gy = floorf(uv.y * p.size.y);
    uint32_t rng = pcg((uint32_t)to_int(gx) + pcg((uint32_t)to_int(gy) + pcg(f.frame_index)));

    Res canonical = empty_res();
    for (int k = 0; k < f.initial_candidates; ++k) {
        float r0 = next_random(rng);
        float r1 = next_random(rng);
        float rsel = next_random(rng);
        F3 dir = cosine_lobe(n, r0, r1);
        canonical.count += 1.0f;
        F3 hit_position = splat3(0.0f);
        float hit_id = 0.0f;
        bool hit = false;
        if (working && f.ray_mode != RGI_RAY_BVH) hit = march(p, ray_origin, dir, 3.6f * p.sc, hit_position);
        bool need = working && !hit && f.ray_mode != RGI_RAY_SCREEN;
        uint64_t queue = __ballot(need);
        if (queue != 0ull) {
            uint32_t slot = rank_in(queue);
            if (need) {
                q_origin[qbase + slot] = make_float4(ray_origin.x, ray_origin.y, ray_origin.z, 0.0f);
                q_dir[qbase + slot] = make_float4(dir.x, dir.y, dir.z, 0.0f);
            }
            wave_sync();
            uint32_t queued = (uint32_t)__popcll(queue);
            if (lane < queued) {
                float4 qo = q_origin[qbase + lane], qd = q_dir[qbase + lane];
                float dist;
                int tri = trace_closest(p.v, F3{qo.x, qo.y, qo.z}, F3{qd.x, qd.y, qd.z}, 4.0f * p.sc, dist,
                                        ring + threadIdx.x, 256);
                q_tri[qbase + lane] = tri;
                q_dist[qbase + lane] = dist;
            }
            wave_sync();
            if (need) {
                int tri = q_tri[qbase + slot];
                float dist = q_dist[qbase + slot];
                if (tri >= 0) {
                    hit = true;
                    hit_position = ray_origin + dir * dist;
                    hit_id = sample_id_of(p.v, tri, dir);
                } else if (sky_on(f)) {
                    hit = true;
                    hit_position = dir;
                    hit_id = kSkySampleId;
                }
            }
            wave_sync();
        }
        stream_candidate(p, position, n, albedo, ray_origin, dir, hit, hit_position, hit_id, rsel, canonical);
    }
    float canonical_weight = canonical.target > 0.000001f ? canonical.wsum / (canonical.count * canonical.target) : 0.0f;

    bool history_valid = f.history_valid != 0;
    F2 previous_uv = F2{0.0f, 0.0f};
    F4 clip_prev = mul(f.prev_view_proj, F4{position.x, position.y, position.z, 1.0f});
    if (clip_prev.w > 0.0001f) previous_uv = clip_to_uv(f, clip_prev);

    float radius = reuse_radius(p, position);
    Pairwise mis = begin_pairwise(p, position, n, canonical, canonical.count, f.temporal_taps);
    float combined_weight = 0.0f;
    if (working) {
        for (int tap = 0; tap < f.temporal_taps; ++tap) {
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            float r2 = next_random(rng);
            float rsel = next_random(rng);
            F2 huv = F2{-1.0f, -1.0f};
            if (history_valid) {
                if (f.world_space_reuse) {
                    uint32_t entry = grid_pick(f, in.grid, p.grid_camera, position, n, r0, r1, r2);
                    if (entry != kGridInvalid) huv = pixel_uv(p, entry);
                } else if (tap == 0 && clip_prev.w > 0.0001f) {
                    huv = previous_uv;
                }
            }
            Nb nb;
            if (!on_screen(huv) || !accept(p, fetch_f4(in.history_sample, p, huv), fetch_f4(in.history_receiver, p, huv),
                                           fetch_f4(in.history_normal, p, huv), position, n, radius, nb)) {
                mis.c_share += mis.pair_c;
                continue;
            }
            add_pair(p, mis, position, n, albedo, nb, rsel);
        }
    }
    if (working) {
        float rfinal = next_random(rng);
        combined_weight = finish_pairwise(p, mis, position, n, canonical, canonical_weight, rfinal);
    }

    // Boiling filter (cs_restir_temporal.sc:269-289) with the 4 x 8 group mean.
    float estimate = working ? mis.combined.target * combined_weight : 0.0f;
    float sum_e = estimate, sum_w = working ? 1.0f : 0.0f;
    for (int stride = 16; stride > 0; stride >>= 1) {
        sum_e = sum_e + __shfl_down(sum_e, stride, 32);
        sum_w = sum_w + __shfl_down(sum_w, stride, 32);
    }
    sum_e = __shfl(sum_e, 0, 32);
    sum_w = __shfl(sum_w, 0, 32);
    float tile_mean = sum_e / maxf_(sum_w, 1.0f);
    if (working && f.boiling_filter > 0.0f && combined_weight > 0.0f) {
        float reference = canonical.target * canonical_weight;
        bool have_history = false;
        F3 history_color = splat3(0.0f);
        if (history_valid && on_screen(previous_uv) && same_history_surface(p, previous_uv, position, n)) {
            history_color = xyz(fetch_h4(in.history_color, p, previous_uv));
            have_history = true;
        }
        if (have_history) reference = maxf_(reference, luminance(history_color * albedo));
        else reference = maxf_(reference, tile_mean);
        float limit = f.boiling_filter * reference;
        if (estimate > limit) combined_weight = combined_weight * (limit / estimate);
    }
    if (!inside) return;
    uint32_t o = (uint32_t)y * (uint32_t)p.hw + (uint32_t)x;
    if (working) {
        store_f4(out.sample, o, F4{mis.combined.position.x, mis.combined.position.y, mis.combined.position.z, mis.total});
        store_f4(out.receiver, o, F4{position.x, position.y, position.z, combined_weight});
        store_f4(out.normal, o, F4{n.x, n.y, n.z, mis.combined.id});
    } else {
        F4 zero = F4{0.0f, 0.0f, 0.0f, 0.0f};
        store_f4(out.sample, o, zero);
        store_f4(out.receiver, o, zero);
        store_f4(out.normal, o, zero);
    }
}

template <bool kSameSize>
RGI_DEV void gi_spatial_body(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                             const rgi_gi_reservoirs& res, uint16_t* out_color, uint16_t* out_moments, F2 texel,
                             const TileMap& tiles, VisLds& vis)
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
    if (!group_pixel(tiles, x, y)) return;
    if (x >= p.hw || y >= p.hh) return;
    uint32_t o = (uint32_t)y * (uint32_t)p.hw + (uint32_t)x;
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    float depth = fetch_depth(p, uv);
    if (depth >= kSkyDepth) {
        store_h4(out_color, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        store_h4(out_moments, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    F3 position = reconstruct_position(f, uv, depth);
    F3 n = decode_normal(p, uv);
    F3 albedo = diffuse_albedo(p, uv);
    float px = floorf(uv.x * p.size.x), py = floorf(uv.y * p.size.y);
    uint32_t pixel_index = (uint32_t)to_int(py) * (uint32_t)p.hw + (uint32_t)to_int(px);
    uint32_t rng = pcg((uint32_t)to_int(px) + pcg((uint32_t)to_int(py) + pcg(f.frame_index ^ 0x68e31da4u)));

    uint32_t t0 = hix(p, uv);
    F4 own_sample = load_f4(res.sample, t0);
    float cweight = load_f4(res.receiver, t0).w;
    float own_id = load_f4(res.normal, t0).w;
    float ccount = maxf_(minf_(own_sample.w, 20.0f), 1.0f);
    Res canonical = empty_res();
    F3 ccontrib = splat3(0.0f), cnormal = F3{0.0f, 0.0f, 1.0f};
    bool cvalid = false;
    if (cweight > 0.0f) cvalid = shade_sample(p, position, n, albedo, xyz(own_sample), own_id, ccontrib, cnormal);
    if (cvalid) {
        canonical.contribution = ccontrib;
        canonical.position = xyz(own_sample);
        canonical.normal = cnormal;
        canonical.id = own_id;
        canonical.target = luminance(ccontrib);
    }
    float radius = reuse_radius(p, position);
    Pairwise mis = begin_pairwise(p, position, n, canonical, ccount, f.spatial_taps);
    // Taps in chunks whose visibility marches run together (march_queue).
    for (int t0 = 0; t0 < f.spatial_taps; t0 += kVisChunk) {
        const int taps = min(kVisChunk, f.spatial_taps - t0);
        uint32_t marches = 0u, walks = 0u;
#pragma unroll
        for (int j = 0; j < kVisChunk; ++j) {
            if (j >= taps) break;
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            float r2 = next_random(rng);
            float rsel = next_random(rng);
            F2 nuv = F2{-1.0f, -1.0f};
            if (f.world_space_reuse) {
                uint32_t entry = grid_pick(f, in.grid, p.grid_camera, position, n, r0, r1, r2);
                if (entry != kGridInvalid && entry != pixel_index) nuv = pixel_uv(p, entry);
            } else {
                float disk = 12.0f * sqrtf(r0);
                float angle = (2.0f * kPi) * r1;
                float sa, ca;
                rgi_sincos(angle, &sa, &ca);
                float nx = floorf(px + ca * disk), ny = floorf(py + sa * disk);
                if (nx != px || ny != py) nuv = F2{(nx + 0.5f) * texel.x, (ny + 0.5f) * texel.y};
            }
            float ncount = -1.0f, ng = 0.0f, q_max = 0.0f;
            int queue = kVisKnown;
            F3 q_origin = splat3(0.0f), q_dir = splat3(0.0f);
            Nb nb;
            if (on_screen(nuv) && accept(p, fetch_f4(res.sample, p, nuv), fetch_f4(res.receiver, p, nuv),
                                         fetch_f4(res.normal, p, nuv), position, n, radius, nb)) {
                ncount = nb.count;
                add_pair_deferred(p, mis, position, n, albedo, nb, rsel, ng, queue, q_origin, q_dir, q_max);
            }
            queue_tap(vis, j, marches, walks, queue, q_origin, q_dir, q_max, ncount, ng);
        }
        finish_chunk(p, vis, mis, taps, marches, walks);
    }
    float rfinal = next_random(rng);
    float cw = finish_pairwise(p, mis, position, n, canonical, cweight, rfinal);
    F3 estimate = mis.combined.contribution * cw;
    if (f.outlier_clamp > 0.0f) estimate = min3(estimate, splat3(f.outlier_clamp));
    F3 irradiance = estimate / max3(albedo, splat3(0.001f));
    float lum = luminance(irradiance);

    F3 prev_color = splat3(0.0f), prev_moments = splat3(0.0f);
    F4 clip_prev = mul(f.prev_view_proj, F4{position.x, position.y, position.z, 1.0f});
    if (f.history_valid && clip_prev.w > 0.0001f) {
        // Bilinear reprojection over same-surface texels (cs_restir_spatial.sc:49-93).
        F2 puv = clip_to_uv(f, clip_prev);
        float tcx = puv.x * p.size.x - 0.5f, tcy = puv.y * p.size.y - 0.5f;
        float bx = floorf(tcx), by = floorf(tcy);
        float blx = tcx - bx, bly = tcy - by;
        float wsum = 0.0f;
        F3 col = splat3(0.0f), mom = splat3(0.0f);
        for (int k = 0; k < 4; ++k) {
            float ox = (float)(k & 1), oy = (float)(k >> 1);
            F2 tuv = F2{((bx + ox) + 0.5f) * texel.x, ((by + oy) + 0.5f) * texel.y};
            if (!on_screen(tuv)) continue;
            if (!same_history_surface(p, tuv, position, n)) continue;
            float w = mixf_(1.0f - blx, blx, ox) * mixf_(1.0f - bly, bly, oy);
            col = col + xyz(fetch_h4(in.history_color, p, tuv)) * w;
            mom = mom + xyz(fetch_h4(in.history_moments, p, tuv)) * w;
            wsum += w;
        }
        if (!(wsum < 0.01f)) {
            prev_color = col / wsum;
            prev_moments = mom / wsum;
        }
    }
    float history_length = minf_(prev_moments.z + 1.0f, 16.0f);
    float bw = 1.0f / history_length;
    F3 color = mix3(prev_color, irradiance, bw);
    float m0 = mixf_(prev_moments.x, lum, bw);
    float m1 = mixf_(prev_moments.y, lum * lum, bw);
    store_h4(out_color, o, F4{color.x, color.y, color.z, history_length});
    store_h4(out_moments, o, F4{m0, m1, history_length, 0.0f});
}

// The kernels: each pass with the compiler's occupancy, and capped at
// 128 VGPRs for 4 waves per SIMD; each for a G-buffer of the GI size and for
// a larger (window-size) one.
#define RGI_WAVES4 __attribute__((amdgpu_waves_per_eu(4, 4)))

template <bool kSameSize>
__global__ void __launch_bounds__(256) k_gi_temporal(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in,
                                                     rgi_gi_reservoirs out, F2 texel, TileMap tiles)
{
    __shared__ TemporalLds lds;
    gi_temporal_body<kSameSize>(f, s, g, in, out, texel, tiles, lds);
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) RGI_WAVES4
k_gi_temporal_w4(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in, rgi_gi_reservoirs out, F2 texel,
                 TileMap tiles)
{
    __shared__ TemporalLds lds;
    gi_temporal_body<kSameSize>(f, s, g, in, out, texel, tiles, lds);
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) k_gi_spatial(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in,
                                                    rgi_gi_reservoirs res, uint16_t* out_color, uint16_t* out_moments,
                                                    F2 texel, TileMap tiles)
{
    __shared__ VisLds vis;
    gi_spatial_body<kSameSize>(f, s, g, in, res, out_color, out_moments, texel, tiles, vis);
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) RGI_WAVES4
k_gi_spatial_w4(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in, rgi_gi_reservoirs res, uint16_t* out_color,
                uint16_t* out_moments, F2 texel, TileMap tiles)
{
    __shared__ VisLds vis;
    gi_spatial_body<kSameSize>(f, s, g, in, res, out_color, out_moments, texel, tiles, vis);
}

} // namespace

int gi_temporal(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                const rgi_gi_reservoirs& out, const Tuning& tuning, hipStream_t stream)
{
    TileMap tiles = tile_map(tuning.tile_order, tuning.tile_shape, f.gi_width, f.gi_height);
    F2 texel = F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height};
    const bool same = g.width == f.gi_width && g.height == f.gi_height;
    auto kernel = tuning.gi_temporal_waves == 4 ? (same ? k_gi_temporal_w4<true> : k_gi_temporal_w4<false>)
                                                : (same ? k_gi_temporal<true> : k_gi_temporal<false>);
    hipLaunchKernelGGL(kernel, tile_grid(tiles), dim3(256), 0, stream, f, s, g, in, out, texel, tiles);
    return check_launch("gi temporal");
}

int gi_spatial(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
               const rgi_gi_reservoirs& res, uint16_t* color, uint16_t* moments, const Tuning& tuning,
               hipStream_t stream)
{
    TileMap tiles = tile_map(tuning.tile_order, tuning.tile_shape, f.gi_width, f.gi_height);
    F2 texel = F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height};
    const bool same = g.width == f.gi_width && g.height == f.gi_height;
    auto kernel = tuning.gi_spatial_waves == 4 ? (same ? k_gi_spatial_w4<true> : k_gi_spatial_w4<false>)
                                               : (same ? k_gi_spatial<true> : k_gi_spatial<false>);
    hipLaunchKernelGGL(kernel, tile_grid(tiles), dim3(256), 0, stream, f, s, g, in, res, color,
                       moments, texel, tiles);
    return check_launch("gi spatial");
}

} // namespace rgi
