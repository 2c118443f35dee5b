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
#include "batch.h"
#include "internal.h"

namespace rgi {
namespace {

// Per-block LDS of the temporal pass for a workgroup of kWaves waves: the
// closest-hit stack rings, the per-wave ray queue (the candidates' rays,
// then the taps' visibility rays) and the per-wave list of candidates left
// to the BVH.
// kTrace: the candidates are traced in the pass (the LDS below); otherwise
// only the taps' visibility queue.
template <int kWaves, bool kTrace>
struct TemporalLds;
template <int kWaves>
struct TemporalLds<kWaves, true> {
    int ring[kStackRing * 64 * kWaves];
    VisLds<kWaves> vis;
    uint8_t walk[kWaves][kVisQueue];
};
template <int kWaves>
struct TemporalLds<kWaves, false> {
    VisLds<kWaves> vis;
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

// The fresh candidates traced in the pass itself, kVisChunk per pixel at a
// time: each lane draws its rays in candidate order and queues them; the
// wave marches the queue, walks the rays the march left open through the
// BVH, then streams each pixel's candidates into its reservoir in order.
template <int kWaves>
RGI_DEV void candidates_in_place(const GiPass& p, F3 position, F3 n, F3 albedo, F3 ray_origin, bool working,
                                 uint32_t& rng, TemporalLds<kWaves, true>& lds, Res& canonical)
{
    const rgi_frame& f = *p.f;
    const uint32_t wave = threadIdx.x >> 6;
    float4* q_origin = lds.vis.origin[wave];
    float4* q_dir = lds.vis.dir[wave];
    uint8_t* walk_list = lds.walk[wave];
    for (int k0 = 0; k0 < f.initial_candidates; k0 += kVisChunk) {
        const int cands = min(kVisChunk, f.initial_candidates - k0);
        uint32_t slot[kVisChunk];
        float rsel[kVisChunk];
        uint32_t queued = 0u;
#pragma unroll
        for (int j = 0; j < kVisChunk; ++j) {
            slot[j] = kNoSlot;
            rsel[j] = 0.0f;
            if (j >= cands) continue;
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            rsel[j] = next_random(rng);
            F3 dir = cosine_lobe(n, r0, r1);
            canonical.count += 1.0f;
            uint64_t q = __ballot(working);
            if (working) {
                slot[j] = queued + rank_in(q);
                q_origin[slot[j]] = make_float4(ray_origin.x, ray_origin.y, ray_origin.z, 0.0f);
                q_dir[slot[j]] = make_float4(dir.x, dir.y, dir.z, kCandOpen);
            }
            queued += (uint32_t)__popcll(q);
        }
        if (queued != 0u) {
            wave_sync();
            if (f.ray_mode != RGI_RAY_BVH) march_candidates(p, q_origin, q_dir, queued, 3.6f * p.sc);
            if (f.ray_mode != RGI_RAY_SCREEN) {
                wave_sync();
                uint32_t walks = 0u;
#pragma unroll
                for (int j = 0; j < kVisChunk; ++j) {
                    bool need = slot[j] != kNoSlot && q_dir[slot[j]].w == kCandOpen;
                    uint64_t q = __ballot(need);
                    if (need) walk_list[walks + rank_in(q)] = (uint8_t)slot[j];
                    walks += (uint32_t)__popcll(q);
                }
                wave_sync();
                if (walks != 0u)
                    walk_candidates(p, q_origin, q_dir, walk_list, walks, 4.0f * p.sc, lds.ring + threadIdx.x, 64 * kWaves);
            }
            wave_sync();
        }
#pragma unroll
        for (int j = 0; j < kVisChunk; ++j) {
            if (j >= cands) continue;
            F3 hit_position = splat3(0.0f), dir = splat3(0.0f);
            float hit_id = 0.0f;
            bool hit = false;
            if (slot[j] != kNoSlot) {
                float4 a = q_origin[slot[j]], d = q_dir[slot[j]];
                dir = F3{d.x, d.y, d.z};
                if (d.w == kCandMarched) {
                    hit = true;
                    hit_position = F3{a.x, a.y, a.z};
                } else if (d.w == kCandWalked) {
                    int tri = __float_as_int(a.y);
                    if (tri >= 0) {
                        hit = true;
                        hit_position = ray_origin + dir * a.x;
                        hit_id = sample_id_of(p.v, tri, dir);
                    } else if (sky_on(f)) {
                        hit = true;
                        hit_position = dir;
                        hit_id = kSkySampleId;
                    }
                }
            }
            stream_candidate(p, position, n, albedo, ray_origin, dir, hit, hit_position, hit_id, rsel[j], canonical);
        }
        wave_sync();
    }
}

// The fresh candidates from the records k_gi_candidates and k_gi_walks left
// for GI pixel o, streamed into its reservoir in candidate order.
RGI_DEV void candidates_from_records(const GiPass& p, F3 position, F3 n, F3 albedo, F3 ray_origin, bool working,
                                     uint32_t o, uint32_t& rng, const CandRays& c, Res& canonical)
{
    const rgi_frame& f = *p.f;
    for (int k = 0; k < f.initial_candidates; ++k) {
        float r0 = next_random(rng);
        float r1 = next_random(rng);
        float rsel = next_random(rng);
        F3 dir = cosine_lobe(n, r0, r1);
        canonical.count += 1.0f;
        F3 hit_position = splat3(0.0f);
        float hit_id = 0.0f;
        bool hit = false;
        if (working) {
            float4 a = c.rec[(uint32_t)k * c.pixels + o];
            if (a.w == kCandMarched) {
                hit = true;
                hit_position = F3{a.x, a.y, a.z};
            } else if (f.ray_mode != RGI_RAY_SCREEN) {
                uint32_t slot = __float_as_uint(a.x);
                int tri = c.hit[slot];
                if (tri >= 0) {
                    hit = true;
                    hit_position = ray_origin + dir * c.dist[slot];
                    hit_id = sample_id_of(p.v, tri, dir);
                } else if (sky_on(f)) {
                    hit = true;
                    hit_position = dir;
                    hit_id = kSkySampleId;
                }
            }
        }
        stream_candidate(p, position, n, albedo, ray_origin, dir, hit, hit_position, hit_id, rsel, canonical);
    }
}

// One pixel (x, y) of the temporal pass per lane; the wave holds two of the
// reference's 4 x 8 groups, one per half-wave. kRecords: the candidates come
// from `cand` (the listed form); otherwise they are traced here with `lds`.
template <bool kSameSize, int kWaves, bool kRecords>
RGI_DEV void gi_temporal_body(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                              const rgi_gi_reservoirs& out, F2 texel, int x, int y,
                              TemporalLds<kWaves, !kRecords>& lds, const CandRays& cand)
{
    GiPass p;
    p.f = &f;
    p.v = scene_view(s);
    p.g = g;
    p.in = in;
    set_gi_size<kSameSize>(p, f, g, texel);
    p.grid_camera = F3{f.prev_camera_position[0], f.prev_camera_position[1], f.prev_camera_position[2]};
    p.sc = f.scene_scale;

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
    if constexpr (kRecords)
        candidates_from_records(p, position, n, albedo, ray_origin, working, (uint32_t)y * (uint32_t)p.hw + (uint32_t)x,
                                rng, cand, canonical);
    else
        candidates_in_place<kWaves>(p, position, n, albedo, ray_origin, working, rng, lds, canonical);
    float canonical_weight = canonical.target > 0.000001f ? canonical.wsum / (canonical.count * canonical.target) : 0.0f;

    bool history_valid = f.history_valid != 0;
    F2 previous_uv = F2{0.0f, 0.0f};
    F4 clip_prev = mul(f.prev_view_proj, F4{position.x, position.y, position.z, 1.0f});
    if (clip_prev.w > 0.0001f) previous_uv = clip_to_uv(f, clip_prev);

    float radius = reuse_radius(p, position);
    Pairwise mis = begin_pairwise(p, position, n, canonical, canonical.count, f.temporal_taps);
    float combined_weight = 0.0f;
    // Temporal taps, kVisChunk at a time, their visibility rays answered
    // per wave (finish_chunk); every lane takes part in the wave's queue,
    // pixels without a surface with no ray.
    for (int t0 = 0; t0 < f.temporal_taps; t0 += kVisChunk) {
        const int taps = min(kVisChunk, f.temporal_taps - t0);
        uint32_t marches = 0u, walks = 0u;
#pragma unroll
        for (int j = 0; j < kVisChunk; ++j) {
            if (j >= taps) break;
            float ncount = -1.0f, ng = 0.0f, v_max = 0.0f;
            int queue = kVisKnown;
            F3 v_origin = splat3(0.0f), v_dir = splat3(0.0f);
            if (working) {
                float r0 = next_random(rng);
                float r1 = next_random(rng);
                float r2 = next_random(rng);
                float rsel = next_random(rng);
                F2 huv = F2{-1.0f, -1.0f};
                if (history_valid) {
                    if (f.world_space_reuse) {
                        uint32_t entry = grid_pick(f, in.grid, p.grid_camera, position, n, r0, r1, r2);
                        if (entry != kGridInvalid) huv = pixel_uv(p, entry);
                    } else if (t0 + j == 0 && clip_prev.w > 0.0001f) {
                        huv = previous_uv;
                    }
                }
                Nb nb;
                if (on_screen(huv) && accept(p, fetch_f4(in.history_sample, p, huv), fetch_f4(in.history_receiver, p, huv),
                                             fetch_f4(in.history_normal, p, huv), position, n, radius, nb)) {
                    ncount = nb.count;
                    add_pair_deferred(p, mis, position, n, albedo, nb, rsel, ng, queue, v_origin, v_dir, v_max);
                }
            }
            queue_tap(lds.vis, j, marches, walks, queue, v_origin, v_dir, v_max, ncount, ng);
        }
        finish_chunk(p, lds.vis, mis, taps, marches, walks);
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
                             const TileMap& tiles, const BlockAt& at, VisLds<4>& vis)
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

// The kernels: each pass with the compiler's occupancy, and capped: GI
// temporal at 128 VGPRs for 4 waves per SIMD, GI spatial at 96 VGPRs for 5
// (its surface waves then fit the chip in one round on a mostly-sky frame;
// the spills to scratch cost less than the second round); each for a
// G-buffer of the GI size and for a larger (window-size) one.
#define RGI_WAVES4 __attribute__((amdgpu_waves_per_eu(4, 4)))
#define RGI_WAVES5 __attribute__((amdgpu_waves_per_eu(5, 5)))

template <bool kSameSize>
__global__ void __launch_bounds__(256) k_gi_temporal(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in,
                                                     rgi_gi_reservoirs out, F2 texel, TileMap tiles)
{
    __shared__ TemporalLds<4, true> lds;
    int x, y;
    if (!group_pixel(tiles, x, y)) return;
    gi_temporal_body<kSameSize, 4, false>(f, s, g, in, out, texel, x, y, lds, CandRays{});
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) RGI_WAVES4
k_gi_temporal_w4(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in, rgi_gi_reservoirs out, F2 texel,
                 TileMap tiles)
{
    __shared__ TemporalLds<4, true> lds;
    int x, y;
    if (!group_pixel(tiles, x, y)) return;
    gi_temporal_body<kSameSize, 4, false>(f, s, g, in, out, texel, x, y, lds, CandRays{});
}

// The fresh candidates' rays of the listed tiles, recorded for the listed
// form of the pass (cs_restir_temporal.sc:85-150): each lane draws its
// pixel's candidate directions in order, marches each ray through the depth
// buffer and records the hit; the rays the march leaves open are listed for
// the BVH (one list append per workgroup). Apart from the resampling, the
// march needs few registers, so 8 waves per SIMD (64 VGPRs) hide its chain
// of dependent depth loads. Every workgroup runs the same number of rounds
// of 4 tiles (one per wave).
template <bool kSameSize>
RGI_DEV void gi_candidates(const rgi_frame& f, const rgi_gbuffer& g, F2 texel, const TileList& list, const CandRays& c,
                           uint32_t first, uint32_t stride, uint32_t* wave_base)
{
    GiPass p;
    p.f = &f;
    p.g = g;
    set_gi_size<kSameSize>(p, f, g, texel);
    p.sc = f.scene_scale;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t total = list.words[0] + list.words[1];
    const int K = f.initial_candidates;
    const bool trace_open = f.ray_mode != RGI_RAY_SCREEN;
    for (uint32_t round = first * 4u; round < total; round += stride * 4u) {
        const uint32_t item = round + wave;
        int x = 0, y = 0;
        if (item < total) tile_pixel(list, listed_tile(list, item), x, y);
        const bool inside = item < total && x < p.hw && y < p.hh;
        F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
        const float depth = inside ? fetch_depth(p, uv) : 1.0f;
        const bool working = inside && depth < kSkyDepth;
        const uint32_t o = (uint32_t)y * (uint32_t)p.hw + (uint32_t)x;
        F3 n = splat3(0.0f), ray_origin = splat3(0.0f);
        uint32_t rng = 0u;
        if (working) {
            F3 position = reconstruct_position(f, uv, depth);
            n = decode_normal(p, uv);
            ray_origin = position + n * (0.02f * p.sc);
            float gx = floorf(uv.x * p.size.x), gy = floorf(uv.y * p.size.y);
            rng = pcg((uint32_t)to_int(gx) + pcg((uint32_t)to_int(gy) + pcg(f.frame_index)));
        }
        const uint32_t rng0 = rng;
        uint32_t open = 0u; // bit k: candidate k left to the BVH
        for (int k = 0; k < K; ++k) {
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            next_random(rng);
            if (!working) continue;
            F3 dir = cosine_lobe(n, r0, r1);
            F3 hit_position = splat3(0.0f);
            bool hit = false;
            if (f.ray_mode != RGI_RAY_BVH) hit = march(p, ray_origin, dir, 3.6f * p.sc, hit_position);
            if (hit) {
                c.rec[(uint32_t)k * c.pixels + o] = make_float4(hit_position.x, hit_position.y, hit_position.z, kCandMarched);
            } else if (trace_open) {
                open |= 1u << k;
            } else {
                c.rec[(uint32_t)k * c.pixels + o] = make_float4(0.0f, 0.0f, 0.0f, kCandOpen);
            }
        }
        // List the open rays: ranks per candidate within the wave, one
        // append for the workgroup.
        uint32_t wave_open = 0u;
        for (int k = 0; k < K; ++k) wave_open += (uint32_t)__popcll(__ballot((open >> k) & 1u));
        if ((threadIdx.x & 63u) == 0u) wave_base[wave] = wave_open;
        __syncthreads();
        if (threadIdx.x == 0u) {
            uint32_t sum = 0u;
            for (int w = 0; w < 4; ++w) {
                uint32_t v = wave_base[w];
                wave_base[w] = sum;
                sum += v;
            }
            uint32_t base = sum != 0u ? atomicAdd(list.words + kTileRayWord, sum) : 0u;
            for (int w = 0; w < 4; ++w) wave_base[w] += base;
        }
        __syncthreads();
        uint32_t next = wave_base[wave];
        rng = rng0;
        for (int k = 0; k < K; ++k) {
            float r0 = next_random(rng);
            float r1 = next_random(rng);
            next_random(rng);
            const bool mine = ((open >> k) & 1u) != 0u;
            const uint64_t b = __ballot(mine);
            if (mine) {
                uint32_t slot = next + rank_in(b);
                F3 dir = cosine_lobe(n, r0, r1);
                c.rays[2u * slot] = make_float4(ray_origin.x, ray_origin.y, ray_origin.z, 4.0f * p.sc);
                c.rays[2u * slot + 1u] = make_float4(dir.x, dir.y, dir.z, 0.0f);
                c.rec[(uint32_t)k * c.pixels + o] = make_float4(__uint_as_float(slot), 0.0f, 0.0f, kCandOpen);
            }
            next += (uint32_t)__popcll(b);
        }
        __syncthreads();
    }
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_gi_candidates(rgi_frame f, rgi_gbuffer g, F2 texel, TileList list, CandRays c)
{
    __shared__ uint32_t wave_base[4];
    gi_candidates<kSameSize>(f, g, texel, list, c, blockIdx.x, gridDim.x, wave_base);
}

// Closest hits of the candidates' listed rays (traceBvh, bvh.sh:216-295),
// answered by persistent waves with per-lane refill, in chunks of 64 rays (a
// frame's few hundred thousand rays then still reach every wave).
__global__ void __launch_bounds__(256) k_gi_walks(rgi_scene s, TileList list, CandRays c)
{
    __shared__ int ring[kStackRing * 256];
    trace_persistent<MollerTrumbore, false>(scene_view(s), c.rays, list.words[kTileRayWord], c.hit, c.dist,
                                            list.words + kWalkClaimWord, ring + threadIdx.x, 256, 64u);
}

__global__ void __launch_bounds__(256) k_gi_walks_batch(const BatchView* __restrict__ views, uint32_t count)
{
    __shared__ int ring[kStackRing * 256];
    for (uint32_t i = 0; i < count; ++i) {
        const BatchView& v = views[i];
        if (!v.split || v.f.initial_candidates <= 0 || v.f.ray_mode == RGI_RAY_SCREEN) continue;
        trace_persistent<MollerTrumbore, false>(scene_view(v.s), v.cand.rays, v.list.words[kTileRayWord], v.cand.hit,
                                                v.cand.dist, v.list.words + kWalkClaimWord, ring + threadIdx.x, 256,
                                                64u);
    }
}

// The listed form: one-wave workgroups that claim the listed tiles
// (heavy ones first) until the list is used up; every wave reaches that
// exit test. A wave's LDS is its own, so a long tile holds no other wave's.
template <bool kSameSize>
RGI_DEV void gi_temporal_listed(const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g, const rgi_gi_inputs& in,
                                const rgi_gi_reservoirs& out, F2 texel, const TileList& list, const CandRays& c,
                                TemporalLds<1, false>& lds)
{
    uint32_t t;
    while (claim_tile(list, t)) {
        int x, y;
        tile_pixel(list, t, x, y);
        gi_temporal_body<kSameSize, 1, true>(f, s, g, in, out, texel, x, y, lds, c);
        wave_sync();
    }
}

template <bool kSameSize>
__global__ void __launch_bounds__(64) k_gi_temporal_listed(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in,
                                                           rgi_gi_reservoirs out, F2 texel, TileList list, CandRays c)
{
    __shared__ TemporalLds<1, false> lds;
    gi_temporal_listed<kSameSize>(f, s, g, in, out, texel, list, c, lds);
}

template <bool kSameSize>
__global__ void __launch_bounds__(64) RGI_WAVES4
k_gi_temporal_listed_w4(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in, rgi_gi_reservoirs out, F2 texel,
                        TileList list, CandRays c)
{
    __shared__ TemporalLds<1, false> lds;
    gi_temporal_listed<kSameSize>(f, s, g, in, out, texel, list, c, lds);
}

template <bool kSameSize>
RGI_DEV void gi_temporal_listed_views(const BatchView* __restrict__ views, uint32_t count, TemporalLds<1, false>& lds)
{
    for (uint32_t i = 0; i < count; ++i) {
        const BatchView& v = views[i];
        if (!v.split || v.same != (kSameSize ? 1u : 0u)) continue;
        gi_temporal_listed<kSameSize>(v.fg, v.s, v.g, v.tin, v.res, v.gi_texel, v.list, v.cand, lds);
    }
}

template <bool kSameSize>
__global__ void __launch_bounds__(64) k_gi_temporal_listed_batch(const BatchView* __restrict__ views, uint32_t count)
{
    __shared__ TemporalLds<1, false> lds;
    gi_temporal_listed_views<kSameSize>(views, count, lds);
}

template <bool kSameSize>
__global__ void __launch_bounds__(64) RGI_WAVES4
k_gi_temporal_listed_batch_w4(const BatchView* __restrict__ views, uint32_t count)
{
    __shared__ TemporalLds<1, false> lds;
    gi_temporal_listed_views<kSameSize>(views, count, lds);
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) k_gi_spatial(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in,
                                                    rgi_gi_reservoirs res, uint16_t* out_color, uint16_t* out_moments,
                                                    F2 texel, TileMap tiles)
{
    __shared__ VisLds<4> vis;
    gi_spatial_body<kSameSize>(f, s, g, in, res, out_color, out_moments, texel, tiles, this_block(), vis);
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) RGI_WAVES5
k_gi_spatial_w5(rgi_frame f, rgi_scene s, rgi_gbuffer g, rgi_gi_inputs in, rgi_gi_reservoirs res, uint16_t* out_color,
                uint16_t* out_moments, F2 texel, TileMap tiles)
{
    __shared__ VisLds<4> vis;
    gi_spatial_body<kSameSize>(f, s, g, in, res, out_color, out_moments, texel, tiles, this_block(), vis);
}

template <bool kSameSize>
RGI_DEV void gi_spatial_view(const BatchView* __restrict__ views, const uint32_t* __restrict__ first, uint32_t count,
                             VisLds<4>& vis)
{
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    if (v.same != (kSameSize ? 1u : 0u)) return;
    gi_spatial_body<kSameSize>(v.fg, v.s, v.g, v.sin, v.res, v.color, v.moments, v.gi_texel, v.gi_tiles,
                               BlockAt{b, v.gi_gx, v.gi_gy}, vis);
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) k_gi_spatial_batch(const BatchView* __restrict__ views,
                                                          const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ VisLds<4> vis;
    gi_spatial_view<kSameSize>(views, first, count, vis);
}

template <bool kSameSize>
__global__ void __launch_bounds__(256) RGI_WAVES5
k_gi_spatial_batch_w5(const BatchView* __restrict__ views, const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ VisLds<4> vis;
    gi_spatial_view<kSameSize>(views, first, count, vis);
}

// Zeroes a list's counters before k_gi_tiles fills it.
__global__ void __launch_bounds__(kTileListWord) k_gi_tiles_reset(TileList list)
{
    list.words[threadIdx.x] = 0u;
}

// Lists the surface tiles of a GI pass, one tile per wave and
// kListTilesPerGroup per workgroup (one append per kind and workgroup), and writes the empty
// reservoirs of every GI pixel of the tiles without a surface: the pixels
// that the temporal pass writes as zeros without reading anything else.
template <bool kSameSize>
RGI_DEV void gi_tiles_body(const rgi_frame& f, const rgi_gbuffer& g, F2 texel, const TileList& list,
                           const rgi_gi_reservoirs& out, uint32_t block, uint32_t* counts, uint32_t* bases)
{
    if (threadIdx.x < 2u) counts[threadIdx.x] = 0u;
    __syncthreads();
    GiPass p;
    p.f = &f;
    p.g = g;
    set_gi_size<kSameSize>(p, f, g, texel);
    const uint32_t t = block * kListTilesPerGroup + (threadIdx.x >> 6);
    const bool valid = t < list.tiles;
    int x = 0, y = 0;
    if (valid) tile_pixel(list, t, x, y);
    const bool inside = valid && x < p.hw && y < p.hh;
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    const bool working = inside && fetch_depth(p, uv) < kSkyDepth;
    const uint64_t b = __ballot(working);
    const uint32_t kind = b == 0ull ? 2u : (__popcll(b) >= kHeavyTile ? 0u : 1u);
    const bool leader = (threadIdx.x & 63u) == 0u && valid && kind < 2u;
    uint32_t rank = 0u;
    if (leader) rank = atomicAdd(&counts[kind], 1u);
    if (kind == 2u && inside) {
        uint32_t o = (uint32_t)y * (uint32_t)p.hw + (uint32_t)x;
        F4 zero = F4{0.0f, 0.0f, 0.0f, 0.0f};
        store_f4(out.sample, o, zero);
        store_f4(out.receiver, o, zero);
        store_f4(out.normal, o, zero);
    }
    __syncthreads();
    if (threadIdx.x < 2u && counts[threadIdx.x] != 0u) bases[threadIdx.x] = atomicAdd(list.words + threadIdx.x, counts[threadIdx.x]);
    __syncthreads();
    if (leader) {
        uint32_t i = bases[kind] + rank;
        list.words[kTileListWord + (kind == 0u ? i : list.tiles - 1u - i)] = t;
    }
}

template <bool kSameSize>
__global__ void __launch_bounds__(64 * kListTilesPerGroup) k_gi_tiles(rgi_frame f, rgi_gbuffer g, F2 texel, TileList list,
                                                   rgi_gi_reservoirs out)
{
    __shared__ uint32_t counts[2], bases[2];
    gi_tiles_body<kSameSize>(f, g, texel, list, out, blockIdx.x, counts, bases);
}

// ------------------------------------------------------------------ batches
//
// The passes of a batch of views in one launch each, over the device table
// of the views (BatchView). Screen passes give each view a slice of the
// grid (workgroups [first[v], first[v + 1])); the listed and persistent
// passes loop over the views, each view's tiles or rays claimed from its own
// counters, so that a workgroup done with one view moves on to the next.
// Each kernel works on the views whose G-buffer size matches its kSameSize
// (a batch that mixes both launches both forms); the views' results are
// those of their single-view passes, the same bodies on the same inputs.

__global__ void __launch_bounds__(kTileListWord) k_gi_tiles_reset_batch(const BatchView* __restrict__ views)
{
    const BatchView& v = views[blockIdx.x];
    if (v.split) v.list.words[threadIdx.x] = 0u;
}

template <bool kSameSize>
__global__ void __launch_bounds__(64 * kListTilesPerGroup) k_gi_tiles_batch(const BatchView* __restrict__ views,
                                                         const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ uint32_t counts[2], bases[2];
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    if (v.same != (kSameSize ? 1u : 0u)) return; // the whole workgroup, before the barriers
    gi_tiles_body<kSameSize>(v.f, v.g, v.gi_texel, v.list, v.res, b, counts, bases);
}

// A view's rounds of 4 tiles continue the previous views' rounds, so that
// the workgroups share the whole batch's rounds evenly.
template <bool kSameSize>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_gi_candidates_batch(const BatchView* __restrict__ views, uint32_t count)
{
    __shared__ uint32_t wave_base[4];
    uint32_t before = 0u;
    for (uint32_t i = 0; i < count; ++i) {
        const BatchView& v = views[i];
        if (!v.split || v.same != (kSameSize ? 1u : 0u) || v.f.initial_candidates <= 0) continue;
        const uint32_t first = (blockIdx.x + gridDim.x - before % gridDim.x) % gridDim.x;
        gi_candidates<kSameSize>(v.f, v.g, v.gi_texel, v.list, v.cand, first, gridDim.x, wave_base);
        before += (v.list.words[0] + v.list.words[1] + 3u) / 4u;
    }
}

// Workgroups of a listed pass: as many one-wave workgroups as the chip holds
// at 4 waves per SIMD (16 per CU), at most one per tile.
uint32_t listed_grid(rgi_context* ctx, const TileList& list)
{
    uint32_t most = 16u * (uint32_t)context_compute_units(ctx);
    return list.tiles < most ? list.tiles : most;
}

} // namespace

int context_cand_rays(rgi_context* ctx, uint32_t candidates, CandRays& c)
{
    // rec: 16 B per candidate and pixel; rays: 32 B, hit and dist: 4 B each
    // per listed ray, at most one per candidate and pixel.
    const size_t n = (size_t)candidates * c.pixels;
    if (n == 0) {
        c.rec = c.rays = nullptr;
        c.hit = nullptr;
        c.dist = nullptr;
        return RGI_OK;
    }
    const size_t rec = align256(n * sizeof(float4)), rays = align256(2 * n * sizeof(float4)),
                 hits = align256(n * sizeof(int32_t)), dists = align256(n * sizeof(float));
    void* buffer = context_buffer(ctx, kBufCandidates, rec + rays + hits + dists);
    if (!buffer) return RGI_E_NOMEM;
    char* base = static_cast<char*>(buffer);
    c.rec = reinterpret_cast<float4*>(base);
    c.rays = reinterpret_cast<float4*>(base + rec);
    c.hit = reinterpret_cast<int32_t*>(base + rec + rays);
    c.dist = reinterpret_cast<float*>(base + rec + rays + hits);
    return RGI_OK;
}

bool gi_temporal_split(const rgi_context* ctx, const rgi_frame& f)
{
    return ctx != nullptr && f.initial_candidates <= kMaxCandRecords;
}

// The listed form's descriptors for frame f: the context's tile list and
// candidate records (neither is reallocated between the two halves of one
// frame, whose sizes are equal).
static int listed_buffers(rgi_context* ctx, const rgi_frame& f, TileList& list, CandRays& c)
{
    list.tiles_x = ((uint32_t)f.gi_width + 7u) / 8u;
    list.tiles = list.tiles_x * (((uint32_t)f.gi_height + 7u) / 8u);
    list.words = static_cast<uint32_t*>(
        context_buffer(ctx, kBufTileList, (size_t)(kTileListWord + list.tiles) * sizeof(uint32_t)));
    if (!list.words) return RGI_E_NOMEM;
    c.pixels = (uint32_t)f.gi_width * (uint32_t)f.gi_height;
    return context_cand_rays(ctx, (uint32_t)f.initial_candidates, c);
}

int gi_temporal_trace(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                      const rgi_gi_reservoirs& out, hipStream_t stream)
{
    TileList list;
    CandRays c;
    int e = listed_buffers(ctx, f, list, c);
    if (e != RGI_OK) return e;
    F2 texel = F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height};
    const bool same = g.width == f.gi_width && g.height == f.gi_height;
    hipLaunchKernelGGL(k_gi_tiles_reset, dim3(1), dim3(kTileListWord), 0, stream, list);
    hipLaunchKernelGGL(same ? k_gi_tiles<true> : k_gi_tiles<false>,
                       dim3((list.tiles + kListTilesPerGroup - 1u) / kListTilesPerGroup), dim3(64 * kListTilesPerGroup),
                       0, stream, f, g, texel, list, out);
    const uint32_t cus = (uint32_t)context_compute_units(ctx);
    if (f.initial_candidates > 0) {
        const uint32_t rounds = (list.tiles + 3u) / 4u, most = 8u * cus;
        hipLaunchKernelGGL(same ? k_gi_candidates<true> : k_gi_candidates<false>, dim3(rounds < most ? rounds : most),
                           dim3(256), 0, stream, f, g, texel, list, c);
        // Without a BVH every walk ends at once with no hit, as trace_closest
        // does. As many workgroups as the CUs hold at once: the walks wait
        // on chains of dependent node loads, so every resident wave counts
        // (8 instead of 4 per CU: atrium pass 1.98 -> 1.87 ms).
        if (f.ray_mode != RGI_RAY_SCREEN) {
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gi_walks, 256, 0) != hipSuccess || per_cu <= 0)
                per_cu = 4;
            hipLaunchKernelGGL(k_gi_walks, dim3((uint32_t)per_cu * cus), dim3(256), 0, stream, s, list, c);
        }
    }
    return check_launch("gi temporal tracing");
}

int gi_temporal_resample(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                         const rgi_gi_inputs& in, const rgi_gi_reservoirs& out, const Tuning& tuning,
                         hipStream_t stream)
{
    TileList list;
    CandRays c;
    int e = listed_buffers(ctx, f, list, c);
    if (e != RGI_OK) return e;
    F2 texel = F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height};
    const bool same = g.width == f.gi_width && g.height == f.gi_height;
    auto kernel = tuning.gi_temporal_waves == 4 ? (same ? k_gi_temporal_listed_w4<true> : k_gi_temporal_listed_w4<false>)
                                                : (same ? k_gi_temporal_listed<true> : k_gi_temporal_listed<false>);
    hipLaunchKernelGGL(kernel, dim3(listed_grid(ctx, list)), dim3(64), 0, stream, f, s, g, in, out, texel, list, c);
    return check_launch("gi temporal");
}

int gi_temporal(rgi_context* ctx, const rgi_frame& f, const rgi_scene& s, const rgi_gbuffer& g,
                const rgi_gi_inputs& in, const rgi_gi_reservoirs& out, const Tuning& tuning, hipStream_t stream)
{
    if (gi_temporal_split(ctx, f)) {
        int e = gi_temporal_trace(ctx, f, s, g, out, stream);
        return e != RGI_OK ? e : gi_temporal_resample(ctx, f, s, g, in, out, tuning, stream);
    }
    F2 texel = F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height};
    const bool same = g.width == f.gi_width && g.height == f.gi_height;
    TileMap tiles = tile_map(tuning.tile_order, tuning.tile_shape, f.gi_width, f.gi_height);
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
    auto kernel = tuning.gi_spatial_waves == 5 ? (same ? k_gi_spatial_w5<true> : k_gi_spatial_w5<false>)
                                               : (same ? k_gi_spatial<true> : k_gi_spatial<false>);
    hipLaunchKernelGGL(kernel, tile_grid(tiles), dim3(256), 0, stream, f, s, g, in, res, color,
                       moments, texel, tiles);
    return check_launch("gi spatial");
}

// ------------------------------------------------------------------ batches

int gi_trace_batch(const BatchLaunch& b, hipStream_t stream)
{
    if (!b.split) return RGI_OK;
    hipLaunchKernelGGL(k_gi_tiles_reset_batch, dim3(b.count), dim3(kTileListWord), 0, stream, b.views);
    const uint32_t* first = b.first[kPassTiles];
    const uint32_t blocks = b.blocks[kPassTiles];
    const dim3 threads(64 * kListTilesPerGroup);
    if (b.same) hipLaunchKernelGGL(k_gi_tiles_batch<true>, dim3(blocks), threads, 0, stream, b.views, first, b.count);
    if (b.larger) hipLaunchKernelGGL(k_gi_tiles_batch<false>, dim3(blocks), threads, 0, stream, b.views, first, b.count);
    const uint32_t cus = (uint32_t)b.compute_units;
    if (b.rounds > 0u) {
        const uint32_t grid = b.rounds < 8u * cus ? b.rounds : 8u * cus;
        if (b.same) hipLaunchKernelGGL(k_gi_candidates_batch<true>, dim3(grid), dim3(256), 0, stream, b.views, b.count);
        if (b.larger) hipLaunchKernelGGL(k_gi_candidates_batch<false>, dim3(grid), dim3(256), 0, stream, b.views, b.count);
        if (b.trace) {
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gi_walks_batch, 256, 0) != hipSuccess ||
                per_cu <= 0)
                per_cu = 4;
            hipLaunchKernelGGL(k_gi_walks_batch, dim3((uint32_t)per_cu * cus), dim3(256), 0, stream, b.views, b.count);
        }
    }
    return check_launch("gi temporal tracing (batch)");
}

int gi_resample_batch(const BatchLaunch& b, hipStream_t stream)
{
    if (!b.split) return RGI_OK;
    const uint32_t most = 16u * (uint32_t)b.compute_units, grid = b.tiles < most ? b.tiles : most;
    const bool w4 = b.gi_temporal_waves == 4;
    if (b.same)
        hipLaunchKernelGGL(w4 ? k_gi_temporal_listed_batch_w4<true> : k_gi_temporal_listed_batch<true>, dim3(grid),
                           dim3(64), 0, stream, b.views, b.count);
    if (b.larger)
        hipLaunchKernelGGL(w4 ? k_gi_temporal_listed_batch_w4<false> : k_gi_temporal_listed_batch<false>, dim3(grid),
                           dim3(64), 0, stream, b.views, b.count);
    return check_launch("gi temporal (batch)");
}

int gi_spatial_batch(const BatchLaunch& b, hipStream_t stream)
{
    const uint32_t* first = b.first[kPassGiSpatial];
    const uint32_t blocks = b.blocks[kPassGiSpatial];
    if (blocks == 0u) return RGI_OK;
    const bool w5 = b.gi_spatial_waves == 5;
    if (b.same)
        hipLaunchKernelGGL(w5 ? k_gi_spatial_batch_w5<true> : k_gi_spatial_batch<true>, dim3(blocks), dim3(256), 0,
                           stream, b.views, first, b.count);
    if (b.larger)
        hipLaunchKernelGGL(w5 ? k_gi_spatial_batch_w5<false> : k_gi_spatial_batch<false>, dim3(blocks), dim3(256), 0,
                           stream, b.views, first, b.count);
    return check_launch("gi spatial (batch)");
}

int gi_listed_buffers(rgi_context* ctx, const rgi_frame& f, TileList& list, CandRays& c)
{
    return listed_buffers(ctx, f, list, c);
}

} // namespace rgi
