// ReSTIR GI per-pixel logic on gfx950 (restir.sh, shadow.sh, sky.sh): sample
// shading, the depth-buffer march, visibility, the reconnection Jacobian and
// pairwise MIS, evaluated per lane in the same order as the reference's
// shader so results are bit-exact.
#pragma once

#include "bvh_device.h"
#include "shading.h"
#include "tiles.h"

namespace rgi {

// Everything a GI pass reads besides its per-pixel state. The G-buffer and
// the surface light are window-size planes; the reservoir, history and
// colour planes are GI-size planes (u_screenInfo.xy), smaller by the GI
// divisor (settings.h:18-23). Each is point-sampled at its own size.
struct GiPass {
    const rgi_frame* f;      // kernel-argument frame
    SceneView v;
    rgi_gbuffer g;
    rgi_gi_inputs in;
    F2 texel;                // 1 / GI size (u_screenInfo.zw)
    F2 size;                 // GI size (u_screenInfo.xy)
    int hw, hh;              // GI size as integers
    F3 grid_camera;          // u_gridCamera of this pass
    float sc;                // scene scale
};

// Texel of a G-buffer (window-size) plane and of a GI-size plane at uv.
RGI_DEV uint32_t tix(const GiPass& p, F2 uv) { return texel_index(uv, p.g.width, p.g.height); }
RGI_DEV uint32_t hix(const GiPass& p, F2 uv) { return texel_index(uv, p.hw, p.hh); }

// `texel` is 1 / GI size, computed once on the host and passed as a kernel
// argument (a uniform value: dividing it out per lane costs registers). With
// kSameSize (a G-buffer of the GI size, divisor 1) the GI-plane size is taken
// from the G-buffer's own fields, so the compiler sees one texel index where
// both kinds of plane are fetched at the same uv.
template <bool kSameSize>
RGI_DEV void set_gi_size(GiPass& p, const rgi_frame& f, const rgi_gbuffer& g, F2 texel)
{
    p.hw = kSameSize ? g.width : f.gi_width;
    p.hh = kSameSize ? g.height : f.gi_height;
    p.size = F2{(float)p.hw, (float)p.hh};
    p.texel = texel;
}
RGI_DEV float fetch_depth(const GiPass& p, F2 uv) { return p.g.depth[tix(p, uv)]; }

RGI_DEV F3 decode_normal(const GiPass& p, F2 uv)
{
    F4 n = unorm8(p.g.normal[tix(p, uv)]);
    return normalize(F3{n.x * 2.0f - 1.0f, n.y * 2.0f - 1.0f, n.z * 2.0f - 1.0f});
}

RGI_DEV F3 diffuse_albedo(const GiPass& p, F2 uv)
{
    uint32_t t = tix(p, uv);
    F4 a = unorm8(p.g.albedo[t]);
    float metallic = unorm8(p.g.normal[t]).w;
    return F3{a.x, a.y, a.z} * (1.0f - metallic);
}

// ------------------------------------------------------------ shadow, sky

// lightVisibilityHard (shadow.sh:99-108): one hard compare.
RGI_DEV float light_visibility_hard(const GiPass& p, F3 position, F3 n)
{
    F2 uv;
    float reference;
    if (!shadow_lookup(*p.f, p.in.shadow_map, position, n, 0.5f, uv, reference)) return 1.0f;
    float stored = p.in.shadow_map[texel_index(uv, p.f->shadow_width, p.f->shadow_height)];
    return reference <= stored ? 1.0f : 0.0f;
}

// ------------------------------------------------------------ shading

RGI_DEV F3 cosine_lobe(F3 n, float r0, float r1)
{
    float radius = sqrtf(r0);
    float angle = (2.0f * kPi) * r1;
    float sa, ca;
    rgi_sincos(angle, &sa, &ca);
    float lx = radius * ca, ly = radius * sa, lz = sqrtf(maxf_(1.0f - r0, 0.0f));
    F3 t, b;
    tangent_frame(n, t, b);
    return normalize((t * lx + b * ly) + n * lz);
}

RGI_DEV F3 point_light_direct(const GiPass& p, F3 albedo, F3 position, F3 n)
{
    const rgi_frame& f = *p.f;
    F3 to_light = F3{f.light_position[0], f.light_position[1], f.light_position[2]} - position;
    float d2 = maxf_(dot(to_light, to_light), (0.01f * p.sc) * p.sc);
    float nl = maxf_(dot(n, to_light * rsqrt_(d2)), 0.0f);
    F3 direct = splat3(0.0f);
    if (nl > 0.0f) {
        F3 color = F3{f.light_color[0], f.light_color[1], f.light_color[2]};
        direct = ((albedo * (color / d2)) * (nl / kPi)) * light_visibility_hard(p, position, n);
    }
    return direct;
}

RGI_DEV bool is_sky(float id) { return id > kSkySampleId * 0.5f; }

RGI_DEV float sample_id_of(const SceneView& v, int tri, F3 dir)
{
    F3 v0 = vertex(v, tri, 0);
    F3 pn = cross(vertex(v, tri, 1) - v0, vertex(v, tri, 2) - v0);
    return dot(pn, dir) <= 0.0f ? (float)(tri + 1) : -(float)(tri + 1);
}

// GI-size planes (reservoirs, history receiver/normal/colour/moments).
RGI_DEV F4 fetch_f4(const float* plane, const GiPass& p, F2 uv) { return load_f4(plane, hix(p, uv)); }
RGI_DEV F4 fetch_h4(const uint16_t* plane, const GiPass& p, F2 uv) { return load_h4(plane, hix(p, uv)); }
// The surface light (DI diffuse or the G-buffer emission): window size.
RGI_DEV F3 fetch_light(const GiPass& p, F2 uv) { return xyz(load_h4(p.in.surface_light, tix(p, uv))); }

// Whether history texel uv held the same surface as (position, n).
RGI_DEV bool same_history_surface(const GiPass& p, F2 uv, F3 position, F3 n)
{
    F3 rp = xyz(fetch_f4(p.in.history_receiver, p, uv));
    F3 rn = xyz(fetch_f4(p.in.history_normal, p, uv));
    F3 off = rp - position;
    float bounce = 0.1f * p.sc;
    return !(dot(rn, rn) < 0.5f || dot(off, off) > bounce * bounce || fabsf(dot(off, n)) > 0.03f * p.sc ||
             dot(normalize(rn), n) < 0.9f);
}

RGI_DEV F3 previous_indirect(const GiPass& p, F3 position, F3 n, F3 albedo)
{
    const rgi_frame& f = *p.f;
    if (!f.history_valid || !f.multiple_bounces) return splat3(0.0f);
    F4 clip = mul(f.prev_view_proj, F4{position.x, position.y, position.z, 1.0f});
    if (clip.w <= 0.0001f) return splat3(0.0f);
    F2 uv = clip_to_uv(f, clip);
    if (!on_screen(uv)) return splat3(0.0f);
    if (!same_history_surface(p, uv, position, n)) return splat3(0.0f);
    return xyz(fetch_h4(p.in.history_color, p, uv)) * albedo;
}

// ------------------------------------------------------------ waves

// Orders this wave's LDS writes before its reads (the wave runs in lockstep).
RGI_DEV void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pixel of this thread: two 4 x 8 groups per wave, eight per block, as a
// 4 x 2 arrangement of groups (16 x 16 tile) or 8 x 1 (32 x 8 wide tile);
// every group stays on the reference's 4 x 8 grid. False when the workgroup
// has no tile.
RGI_DEV bool group_pixel(const TileMap& tiles, const BlockAt& at, int& x, int& y)
{
    uint32_t group = threadIdx.x >> 5, t = threadIdx.x & 31u, tx, ty;
    if (!block_tile(tiles, at, tx, ty)) return false;
    if (tiles.wide) {
        x = (int)(tx * 32u + group * 4u + (t & 3u));
        y = (int)(ty * 8u + (t >> 2));
    } else {
        x = (int)(tx * 16u + (group & 3u) * 4u + (t & 3u));
        y = (int)(ty * 16u + (group >> 2) * 8u + (t >> 2));
    }
    return true;
}
RGI_DEV bool group_pixel(const TileMap& tiles, int& x, int& y) { return group_pixel(tiles, this_block(), x, y); }

// A marched ray of traceScreenSpaceRay: the point at distance t is
// origin + dir * t; its clip position is clip_origin + t * clip_dir (x, y
// and w kept), the projection distributed over the ray.
struct ClipRay {
    F3 origin;
    F3 dir;
    float co_x, co_y, co_w, cd_x, cd_y, cd_w;
};

RGI_DEV ClipRay clip_ray(const GiPass& p, F3 origin, F3 dir)
{
    F4 co = mul(p.f->view_proj, F4{origin.x, origin.y, origin.z, 1.0f});
    F4 cd = mul_dir(p.f->view_proj, dir);
    return ClipRay{origin, dir, co.x, co.y, co.w, cd.x, cd.y, cd.w};
}

// depthBehindSurface (restir.sh:327-358) at distance t along the ray.
RGI_DEV float depth_behind(const GiPass& p, const ClipRay& m, float t, F3& surface, bool& visible)
{
    surface = splat3(0.0f);
    visible = false;
    float cw = fmaf(m.cd_w, t, m.co_w);
    if (cw <= 0.0001f) return -1.0f;
    F2 uv = clip_to_uv(*p.f, F4{fmaf(m.cd_x, t, m.co_x), fmaf(m.cd_y, t, m.co_y), 0.0f, cw});
    visible = on_screen(uv);
    if (!visible) return -1.0f;
    uint32_t ti = tix(p, uv);
    float depth = p.g.depth[ti];
    if (depth >= kSkyDepth) return -1.0f;
    // The view depth of the surface behind the texel. For a perspective
    // camera the w row of the inverse view-projection is (0, 0, c, d), so
    // the point reconstructed at any uv with this depth has view depth
    // 1 / (c z + d): reconstructWorldPosition then the view-projection's w,
    // without the matrix products. The position itself is needed only when
    // the ray is behind the surface.
    const float* inv = p.f->inv_view_proj;
    float behind = cw - 1.0f / fmaf(inv[11], to_clip_depth(*p.f, depth), inv[15]);
    if (behind > 0.0f) {
        surface = reconstruct_position(*p.f, uv, depth);
        F4 nr = unorm8(p.g.normal[ti]);
        F3 sn = normalize(F3{nr.x * 2.0f - 1.0f, nr.y * 2.0f - 1.0f, nr.z * 2.0f - 1.0f});
        F3 ray_position = m.origin + m.dir * t;
        if (dot(ray_position - surface, sn) > -(0.01f * p.sc)) return -1.0f;
    }
    return behind;
}

// traceScreenSpaceRay's march (restir.sh:364-436): quadratic steps; on the
// first step behind the surface after one in front, five bisection steps,
// and a hit if the bisected depth is within the thickness; the ray stops on
// a hit, on leaving the frustum, or past its max distance.
//
// Written as one depth lookup per step: a lane is either stepping
// (round < 0) or in bisection round `round`. Each lane evaluates the same
// lookups in the same order as the nested step/bisection loops, so the
// result is the same; a wave then iterates as often as its busiest lane,
// instead of running every lane's bisections one after another. The state
// is resumable, so a lane can also take over another lane's ray between
// steps (march_queue). With kPosition the hit surface point is kept.
struct MarchState {
    ClipRay m;
    float max_distance, previous, near_d, far_d, step_d, behind_b;
    int i, round;
    bool previous_front, hit, active;
    F3 surface_b;
};

RGI_DEV float march_step_scale(const GiPass& p) { return 32.0f / (float)p.f->trace_steps; }

RGI_DEV void march_begin(const GiPass& p, MarchState& s, F3 origin, F3 dir, float max_distance)
{
    s.m = clip_ray(p, origin, dir);
    s.max_distance = max_distance;
    s.i = 1;
    s.round = -1;
    s.previous = 0.0f;
    s.near_d = 0.0f;
    s.far_d = 0.0f;
    s.step_d = 0.0f;
    s.behind_b = 0.0f;
    s.previous_front = true;
    s.hit = false;
    s.surface_b = splat3(0.0f);
    s.active = s.i <= p.f->trace_steps && !(s.previous >= max_distance);
}

// One depth lookup of an active ray.
template <bool kPosition>
RGI_DEV void march_step(const GiPass& p, MarchState& s, float step_scale)
{
    const int steps = p.f->trace_steps;
    const float sc = p.sc;
    float t;
    if (s.round < 0) {
        float sp = (float)s.i * step_scale;
        t = minf_((0.02f + (sp * sp) * 0.0035f) * sc, s.max_distance);
    } else {
        t = 0.5f * (s.near_d + s.far_d);
    }
    F3 surface;
    bool visible;
    float behind = depth_behind(p, s.m, t, surface, visible);
    bool next_step = false;
    if (s.round < 0) {
        if (!visible) {
            s.active = false;
        } else {
            bool front = behind <= 0.0f;
            if (!front && s.previous_front) {
                s.near_d = s.previous;
                s.far_d = t;
                s.step_d = t;
                s.behind_b = behind;
                if (kPosition) s.surface_b = surface;
                s.round = 0;
            } else {
                s.previous_front = front;
                s.previous = t;
                next_step = true;
            }
        }
    } else {
        if (visible && behind > 0.0f) {
            s.far_d = t;
            s.behind_b = behind;
            if (kPosition) s.surface_b = surface;
        } else {
            s.near_d = t;
        }
        if (++s.round == 5) {
            if (s.behind_b < 0.03f * sc + 0.03f * s.far_d) {
                s.hit = true;
                s.active = false;
            } else {
                s.round = -1;
                s.previous_front = false;
                s.previous = s.step_d;
                next_step = true;
            }
        }
    }
    if (next_step) {
        ++s.i;
        s.active = s.i <= steps && !(s.previous >= s.max_distance);
    }
}

// The march of one ray to its end, per lane: the same lookups and decisions
// as march_begin / march_step, written as one loop (the per-lane form keeps
// the GI temporal pass at 4 waves per SIMD).
RGI_DEV bool march(const GiPass& p, F3 origin, F3 dir, float max_distance, F3& hit_position)
{
    ClipRay m = clip_ray(p, origin, dir);
    const int steps = p.f->trace_steps;
    const float step_scale = 32.0f / (float)steps;
    const float sc = p.sc;
    int i = 1, round = -1;
    float previous = 0.0f, near_d = 0.0f, far_d = 0.0f, step_d = 0.0f, behind_b = 0.0f;
    bool previous_front = true, hit = false;
    F3 surface_b = splat3(0.0f);
    bool active = i <= steps && !(previous >= max_distance);
    while (active) {
        float t;
        if (round < 0) {
            float sp = (float)i * step_scale;
            t = minf_((0.02f + (sp * sp) * 0.0035f) * sc, max_distance);
        } else {
            t = 0.5f * (near_d + far_d);
        }
        F3 surface;
        bool visible;
        float behind = depth_behind(p, m, t, surface, visible);
        bool next_step = false;
        if (round < 0) {
            if (!visible) {
                active = false;
            } else {
                bool front = behind <= 0.0f;
                if (!front && previous_front) {
                    near_d = previous;
                    far_d = t;
                    step_d = t;
                    behind_b = behind;
                    surface_b = surface;
                    round = 0;
                } else {
                    previous_front = front;
                    previous = t;
                    next_step = true;
                }
            }
        } else {
            if (visible && behind > 0.0f) {
                far_d = t;
                behind_b = behind;
                surface_b = surface;
            } else {
                near_d = t;
            }
            if (++round == 5) {
                if (behind_b < 0.03f * sc + 0.03f * far_d) {
                    hit = true;
                    active = false;
                } else {
                    round = -1;
                    previous_front = false;
                    previous = step_d;
                    next_step = true;
                }
            }
        }
        if (next_step) {
            ++i;
            active = i <= steps && !(previous >= max_distance);
        }
    }
    hit_position = hit ? surface_b : splat3(0.0f);
    return hit;
}

// evaluateSample: the integrand of a sample point at a receiver.
RGI_DEV bool shade_sample(const GiPass& p, F3 rpos, F3 rn, F3 ralbedo, F3 spos, float sid, F3& contribution, F3& snormal)
{
    const rgi_frame& f = *p.f;
    contribution = splat3(0.0f);
    snormal = F3{0.0f, 0.0f, 1.0f};
    if (is_sky(sid)) {
        snormal = -spos;
        contribution = (ralbedo * sky_radiance(f, spos)) * (maxf_(dot(rn, spos), 0.0f) / kPi);
        return sky_on(f);
    }
    bool visible_surface = false;
    F2 suv;
    float sview;
    if (project_to_screen(f, spos, suv, sview)) {
        float sdepth = fetch_depth(p, suv);
        visible_surface = sdepth < kSkyDepth && length(reconstruct_position(f, suv, sdepth) - spos) <= 0.1f * p.sc;
    }
    int tri = to_int(fabsf(sid) + 0.5f) - 1;
    if (!visible_surface && (tri < 0 || !bvh_ready(p.v))) return false;
    if (visible_surface) {
        snormal = decode_normal(p, suv);
    } else {
        F3 v0 = vertex(p.v, tri, 0);
        snormal = normalize(cross(vertex(p.v, tri, 1) - v0, vertex(p.v, tri, 2) - v0)) * (sid < 0.0f ? -1.0f : 1.0f);
    }
    F3 to_receiver = rpos - spos;
    float d2 = dot(to_receiver, to_receiver);
    if (d2 < ((0.0001f * p.sc) * p.sc) * 0.01f) return true;
    F3 dir = to_receiver * rsqrt_(d2);
    float scos = dot(snormal, dir);
    float rcos = -dot(rn, dir);
    if (scos <= 0.0f || rcos <= 0.0f) return true;
    F3 radiance;
    if (visible_surface) {
        F3 albedo = diffuse_albedo(p, suv);
        radiance = (point_light_direct(p, albedo, spos, snormal) + previous_indirect(p, spos, snormal, albedo)) +
                   fetch_light(p, suv) * f.lighting_surface;
    } else {
        F3 albedo = tri_albedo(p.v, tri);
        radiance = (point_light_direct(p, albedo, spos, snormal) + previous_indirect(p, spos, snormal, albedo)) +
                   tri_emission(p.v, tri) * f.lighting_emission;
    }
    contribution = (ralbedo * radiance) * (rcos / kPi);
    return true;
}

RGI_DEV float geometry_term(const GiPass& p, F3 rpos, F3 rn, F3 spos, F3 sn, float sid)
{
    if (is_sky(sid)) return maxf_(dot(rn, spos), 0.0f);
    F3 off = spos - rpos;
    float d2 = dot(off, off);
    if (d2 < ((0.0001f * p.sc) * p.sc) * 0.01f) return 0.0f;
    F3 dir = off * rsqrt_(d2);
    float rc = dot(rn, dir);
    float scos = -dot(sn, dir);
    return rc > 0.0f && scos > 0.0f ? (rc * scos) / d2 : 0.0f;
}

// isVisible (restir.sh:564-606) as a ray to answer: kVisBvh for a BVH
// any-hit ray (sky samples, and every ray in BVH mode; a hit means occluded),
// kVisMarch for a march (a hit means occluded), or kVisKnown with `visible`
// already set (a sky sample without a BVH is occluded; a march that takes no
// step misses). `origin`, `dir` and `max_distance` are the ray's.
constexpr int kVisKnown = 0, kVisMarch = 1, kVisBvh = 2;

RGI_DEV int visibility_ray(const GiPass& p, F3 rpos, F3 rn, F3 spos, float sid, bool& visible, F3& origin, F3& dir,
                           float& max_distance)
{
    float bias = 0.02f * p.sc;
    float max_bvh = 4.0f * p.sc;
    origin = rpos + rn * bias;
    bool sky = is_sky(sid);
    F3 to_sample = sky ? spos * max_bvh : spos - origin;
    float sd = length(to_sample);
    dir = to_sample / maxf_(sd, 0.0001f);
    visible = true;
    max_distance = 0.0f;
    if (sky) {
        visible = false;
        max_distance = max_bvh;
        return bvh_ready(p.v) ? kVisBvh : kVisKnown;
    }
    if (p.f->ray_mode == RGI_RAY_BVH && bvh_ready(p.v)) {
        max_distance = sd - bias;
        return kVisBvh;
    }
    max_distance = sd * 0.95f - 0.05f * p.sc;
    // march_begin's first test: such a ray ends before its first lookup.
    return 1 <= p.f->trace_steps && !(0.0f >= max_distance) ? kVisMarch : kVisKnown;
}

RGI_DEV bool visible_between(const GiPass& p, F3 rpos, F3 rn, F3 spos, float sid)
{
    float bias = 0.02f * p.sc;
    float max_bvh = 4.0f * p.sc;
    F3 origin = rpos + rn * bias;
    bool sky = is_sky(sid);
    F3 to_sample = sky ? spos * max_bvh : spos - origin;
    float sd = length(to_sample);
    F3 dir = to_sample / maxf_(sd, 0.0001f);
    bool visible = true;
    if (sky) {
        visible = false;
        if (bvh_ready(p.v)) visible = trace_any(p.v, origin, dir, max_bvh) < 0;
    } else if (p.f->ray_mode == RGI_RAY_BVH && bvh_ready(p.v)) {
        visible = trace_any(p.v, origin, dir, sd - bias) < 0;
    } else {
        F3 occ;
        visible = !march(p, origin, dir, sd * 0.95f - 0.05f * p.sc, occ);
    }
    return visible;
}

RGI_DEV float jacobian(const GiPass& p, F3 here, F3 there, F3 spos, F3 sn, float sid)
{
    if (is_sky(sid)) return 1.0f;
    F3 th = here - spos;
    F3 tt = there - spos;
    float hs = dot(th, th), ts = dot(tt, tt);
    float lim = ((0.0001f * p.sc) * p.sc) * 0.01f;
    if (hs < lim || ts < lim) return 0.0f;
    float ch = dot(sn, th) * rsqrt_(hs);
    float ct = dot(sn, tt) * rsqrt_(ts);
    if (ct <= 0.001f) return 0.0f;
    float j = maxf_((ch / ct) * (ts / hs), 0.0f);
    float clamp_value = p.f->jacobian_clamp;
    return clamp_value > 0.0f ? minf_(j, clamp_value) : j;
}

struct Res {
    F3 contribution, position, normal;
    float id, target, wsum, count;
};

RGI_DEV Res empty_res()
{
    Res r;
    r.contribution = splat3(0.0f);
    r.position = splat3(0.0f);
    r.normal = F3{0.0f, 0.0f, 1.0f};
    r.id = 0.0f;
    r.target = 0.0f;
    r.wsum = 0.0f;
    r.count = 0.0f;
    return r;
}

RGI_DEV bool stream_in(Res& r, F3 contribution, F3 position, F3 n, float id, float target, float weight, float random)
{
    r.wsum += weight;
    if (random * r.wsum < weight) {
        r.contribution = contribution;
        r.position = position;
        r.normal = n;
        r.id = id;
        r.target = target;
        return true;
    }
    return false;
}

struct Nb {
    F3 sample_position;
    float sample_id;
    F3 receiver_position, receiver_normal;
    float count, weight;
};

RGI_DEV bool accept(const GiPass& p, F4 sample, F4 receiver, F4 rnormal, F3 position, F3 n, float radius, Nb& nb)
{
    nb.sample_position = xyz(sample);
    nb.sample_id = rnormal.w;
    nb.receiver_position = xyz(receiver);
    nb.receiver_normal = F3{0.0f, 0.0f, 1.0f};
    nb.count = minf_(sample.w, 20.0f);
    nb.weight = receiver.w;
    F3 rn = xyz(rnormal);
    if (sample.w <= 0.0f || dot(rn, rn) < 0.5f) return false;
    nb.receiver_normal = normalize(rn);
    F3 off = xyz(receiver) - position;
    return dot(off, off) <= radius * radius && fabsf(dot(off, n)) <= 0.03f * p.sc && dot(nb.receiver_normal, n) >= 0.9f;
}

RGI_DEV float reuse_radius(const GiPass& p, F3 position)
{
    float r = 0.4f * p.sc;
    return p.f->world_space_reuse ? maxf_(r, 2.0f * grid_cell_size(*p.f, p.grid_camera, position)) : r;
}

struct Pairwise {
    Res combined;
    F3 c_position, c_normal;
    float c_id, pair_c, c_geometry, c_share, total;
    bool has_c, neighbour_selected;
};

RGI_DEV Pairwise begin_pairwise(const GiPass& p, F3 position, F3 n, const Res& c, float ccount, int taps)
{
    Pairwise m;
    m.combined = empty_res();
    m.c_position = c.position;
    m.c_normal = c.normal;
    m.c_id = c.id;
    m.pair_c = ccount / (float)(taps > 1 ? taps : 1);
    m.total = ccount;
    m.has_c = c.target > 0.000001f;
    m.c_geometry = m.has_c ? geometry_term(p, position, n, c.position, c.normal, c.id) : 0.0f;
    m.c_share = taps > 0 ? 0.0f : ccount;
    m.neighbour_selected = false;
    return m;
}

// One accepted neighbour of the pairwise MIS (addNeighbour, restir.sh:833-893),
// with the visibility of the canonical sample at the neighbour (the
// canonical share's isVisible) left open: `ng` returns that term's geometry
// factor, and when its visibility needs a ray `queue` is its kind (kVisMarch
// or kVisBvh) with the ray (otherwise `ng` already includes the visibility). finish_pair_share adds
// the term once the visibility is known. The shift of the neighbour's
// sample to this pixel does not depend on it.
RGI_DEV void add_pair_deferred(const GiPass& p, Pairwise& m, F3 position, F3 n, F3 albedo, const Nb& nb, float random,
                               float& ng_out, int& queue, F3& q_origin, F3& q_dir, float& q_max)
{
    float ncount = nb.count;
    float pair_weight = ncount + m.pair_c;
    m.total += ncount;
    // The shift first: it touches neither the canonical share nor the
    // canonical sample, so the order does not change a result, and the
    // visibility ray is then built just before it is queued.
    F3 contribution = splat3(0.0f);
    F3 sn = F3{0.0f, 0.0f, 1.0f};
    bool shifted = false;
    if (nb.weight > 0.0f) shifted = shade_sample(p, position, n, albedo, nb.sample_position, nb.sample_id, contribution, sn);
    if (shifted) {
        float target = luminance(contribution);
        float j = jacobian(p, position, nb.receiver_position, nb.sample_position, sn, nb.sample_id);
        float hg = geometry_term(p, position, n, nb.sample_position, sn, nb.sample_id);
        float tg = geometry_term(p, nb.receiver_position, nb.receiver_normal, nb.sample_position, sn, nb.sample_id);
        float nd = ncount * tg;
        float den = nd + m.pair_c * hg;
        float mis = den > 0.0f ? (pair_weight * nd) / den : 0.0f;
        float w = ((mis * target) * nb.weight) * j;
        if (w > 0.0f && stream_in(m.combined, contribution, nb.sample_position, sn, nb.sample_id, target, w, random))
            m.neighbour_selected = true;
    }
    ng_out = 0.0f;
    queue = kVisKnown;
    if (m.has_c) {
        float ng = geometry_term(p, nb.receiver_position, nb.receiver_normal, m.c_position, m.c_normal, m.c_id);
        if (ng > 0.0f) {
            bool visible;
            queue = visibility_ray(p, nb.receiver_position, nb.receiver_normal, m.c_position, m.c_id, visible, q_origin,
                                   q_dir, q_max);
            if (queue == kVisKnown && !visible) ng = 0.0f;
        }
        ng_out = ng;
    }
}

// The canonical sample's MIS share from one tap, in tap order: a rejected tap
// (ncount < 0) gives its whole pair weight; an accepted one the balance
// between the canonical and the neighbour's density, with `ng` zero when the
// canonical sample is occluded at the neighbour.
RGI_DEV void finish_pair_share(Pairwise& m, float ncount, float ng)
{
    if (ncount < 0.0f) {
        m.c_share += m.pair_c;
        return;
    }
    if (!m.has_c) return;
    float pair_weight = ncount + m.pair_c;
    float cd = m.pair_c * m.c_geometry;
    float den = ncount * ng + cd;
    m.c_share += den > 0.0f ? (pair_weight * cd) / den : pair_weight;
}

// One accepted neighbour of the pairwise MIS (addNeighbour,
// restir.sh:833-893) with its visibility answered at once: the temporal pass,
// whose waves hold too few visibility rays to gain from answering them
// together (add_pair_deferred).
RGI_DEV void add_pair(const GiPass& p, Pairwise& m, F3 position, F3 n, F3 albedo, const Nb& nb, float random)
{
    float ncount = nb.count;
    float pair_weight = ncount + m.pair_c;
    m.total += ncount;
    if (m.has_c) {
        float ng = geometry_term(p, nb.receiver_position, nb.receiver_normal, m.c_position, m.c_normal, m.c_id);
        if (ng > 0.0f) {
            if (!visible_between(p, nb.receiver_position, nb.receiver_normal, m.c_position, m.c_id)) ng = 0.0f;
        }
        float cd = m.pair_c * m.c_geometry;
        float den = ncount * ng + cd;
        m.c_share += den > 0.0f ? (pair_weight * cd) / den : pair_weight;
    }
    F3 contribution = splat3(0.0f);
    F3 sn = F3{0.0f, 0.0f, 1.0f};
    bool shifted = false;
    if (nb.weight > 0.0f) shifted = shade_sample(p, position, n, albedo, nb.sample_position, nb.sample_id, contribution, sn);
    if (shifted) {
        float target = luminance(contribution);
        float j = jacobian(p, position, nb.receiver_position, nb.sample_position, sn, nb.sample_id);
        float hg = geometry_term(p, position, n, nb.sample_position, sn, nb.sample_id);
        float tg = geometry_term(p, nb.receiver_position, nb.receiver_normal, nb.sample_position, sn, nb.sample_id);
        float nd = ncount * tg;
        float den = nd + m.pair_c * hg;
        float mis = den > 0.0f ? (pair_weight * nd) / den : 0.0f;
        float w = ((mis * target) * nb.weight) * j;
        if (w > 0.0f && stream_in(m.combined, contribution, nb.sample_position, sn, nb.sample_id, target, w, random))
            m.neighbour_selected = true;
    }
}

// The visibility rays of up to kVisChunk taps of a wave, answered together:
// march rays fill each wave's queue from the front, BVH rays from the back.
constexpr int kVisChunk = 2;
constexpr int kVisQueue = 64 * kVisChunk;
constexpr uint16_t kNoSlot = 0xffffu;

// For a workgroup of kWaves waves.
template <int kWaves>
struct VisLds {
    float4 origin[kWaves][kVisQueue];         // per wave of the block: ray origin, max distance
    float4 dir[kWaves][kVisQueue];            // direction; w = hit (written when the ray is answered)
    float2 rec[kVisChunk][64 * kWaves];       // per tap and thread: neighbour count (-1: rejected), geometry factor
    uint16_t slot[kVisChunk][64 * kWaves];    // per tap and thread: the queued ray, or kNoSlot
};

// Answers this wave's queued rays [first, first + count) of one kind with
// every active lane: a lane whose ray ends takes the next ray not yet
// started, so the wave takes about (all the rays' steps) / 64 iterations
// instead of the sum over taps of each tap's longest ray. Each ray's answer
// is the same whichever lane computes it; its hit goes to dir[k].w. `first`
// and `count` are wave-uniform.
template <int kKind>
RGI_DEV void answer_queue(const GiPass& p, float4* origin, float4* dir, uint32_t first, uint32_t count)
{
    const float step_scale = march_step_scale(p);
    MarchState s;
    AnyWalk<MollerTrumbore> w;
    s.active = false;
    s.hit = false;
    w.active = false;
    w.hit = -1;
    uint32_t k = 0u, next = 0u;
    bool busy = false;
    for (;;) {
        bool active = kKind == kVisMarch ? s.active : w.active;
        if (busy && !active) {
            dir[first + k].w = (kKind == kVisMarch ? s.hit : w.hit >= 0) ? 1.0f : 0.0f;
            busy = false;
        }
        if (next < count) {
            uint64_t idle = __ballot(!busy);
            uint32_t slot = next + rank_in(idle);
            if (!busy && slot < count) {
                k = slot;
                busy = true;
                float4 o = origin[first + k], d = dir[first + k];
                if (kKind == kVisMarch) march_begin(p, s, F3{o.x, o.y, o.z}, F3{d.x, d.y, d.z}, o.w);
                else any_begin(p.v, w, MollerTrumbore{F3{o.x, o.y, o.z}, F3{d.x, d.y, d.z}}, F3{d.x, d.y, d.z}, o.w);
            }
            next += (uint32_t)__popcll(idle);
        }
        if (__ballot(busy) == 0ull) break;
        if (kKind == kVisMarch) {
            if (busy && s.active) march_step<false>(p, s, step_scale);
        } else {
            if (busy && w.active) any_step(p.v, w);
        }
    }
}

// Queues one tap's visibility ray (every active lane of the wave calls it,
// with or without a ray; `marches` and `walks` are the wave's queue counts,
// uniform) and records the tap.
template <int kWaves>
RGI_DEV void queue_tap(VisLds<kWaves>& vis, int j, uint32_t& marches, uint32_t& walks, int kind, F3 o, F3 d,
                       float max_distance, float ncount, float ng)
{
    const uint32_t wave = threadIdx.x >> 6;
    uint64_t want_march = __ballot(kind == kVisMarch), want_walk = __ballot(kind == kVisBvh);
    uint32_t slot = kNoSlot;
    if (kind == kVisMarch) slot = marches + rank_in(want_march);
    if (kind == kVisBvh) slot = (uint32_t)kVisQueue - 1u - (walks + rank_in(want_walk));
    if (slot != kNoSlot) {
        vis.origin[wave][slot] = make_float4(o.x, o.y, o.z, max_distance);
        vis.dir[wave][slot] = make_float4(d.x, d.y, d.z, 0.0f);
    }
    marches += (uint32_t)__popcll(want_march);
    walks += (uint32_t)__popcll(want_walk);
    vis.rec[j][threadIdx.x] = make_float2(ncount, ng);
    vis.slot[j][threadIdx.x] = (uint16_t)slot;
}

// Answers the chunk's queued rays and adds the chunk's MIS shares in tap
// order.
template <int kWaves>
RGI_DEV void finish_chunk(const GiPass& p, VisLds<kWaves>& vis, Pairwise& m, int taps, uint32_t marches, uint32_t walks)
{
    const uint32_t wave = threadIdx.x >> 6;
    if (marches + walks != 0u) {
        wave_sync();
        if (walks != 0u)
            answer_queue<kVisBvh>(p, vis.origin[wave], vis.dir[wave], (uint32_t)kVisQueue - walks, walks);
        if (marches != 0u) answer_queue<kVisMarch>(p, vis.origin[wave], vis.dir[wave], 0u, marches);
        wave_sync();
    }
#pragma unroll
    for (int j = 0; j < kVisChunk; ++j) {
        if (j < taps) {
            float2 r = vis.rec[j][threadIdx.x];
            uint32_t slot = vis.slot[j][threadIdx.x];
            float ng = r.y;
            if (slot != kNoSlot && vis.dir[wave][slot].w != 0.0f) ng = 0.0f;
            finish_pair_share(m, r.x, ng);
        }
    }
    wave_sync();
}

// The temporal pass's fresh candidates, up to kVisChunk per pixel at a time,
// traced per wave from the VisLds queue: every working pixel's ray is
// queued, marched with lane refill (march_candidates), and the rays the
// march leaves open are walked through the BVH with lane refill
// (walk_candidates). Per queued ray k: origin[k] holds the origin, then the
// march's hit point or the walk's (distance, triangle bits); dir[k] holds the
// direction, with w the ray's state below. Whichever lane answers a ray, its
// answer is the one march() / trace_closest() give.
constexpr float kCandOpen = 0.0f, kCandMarched = 1.0f, kCandWalked = 2.0f;

RGI_DEV void march_candidates(const GiPass& p, float4* origin, float4* dir, uint32_t count, float max_distance)
{
    const float step_scale = march_step_scale(p);
    MarchState s;
    s.active = false;
    s.hit = false;
    uint32_t k = 0u, next = 0u;
    bool busy = false;
    for (;;) {
        if (busy && !s.active) {
            if (s.hit) {
                origin[k] = make_float4(s.surface_b.x, s.surface_b.y, s.surface_b.z, 0.0f);
                dir[k].w = kCandMarched;
            }
            busy = false;
        }
        if (next < count) {
            uint64_t idle = __ballot(!busy);
            uint32_t slot = next + rank_in(idle);
            if (!busy && slot < count) {
                k = slot;
                busy = true;
                float4 o = origin[k], d = dir[k];
                march_begin(p, s, F3{o.x, o.y, o.z}, F3{d.x, d.y, d.z}, max_distance);
            }
            next += (uint32_t)__popcll(idle);
        }
        if (__ballot(busy) == 0ull) break;
        if (busy && s.active) march_step<true>(p, s, step_scale);
    }
}

// Closest-hit walks of the queued rays list[0, count); `ring` is this lane's
// ring stack, its entries `stride` ints apart (slot-major over the block's
// lanes).
RGI_DEV void walk_candidates(const GiPass& p, float4* origin, float4* dir, const uint8_t* list, uint32_t count,
                             float max_distance, int* ring, int stride)
{
    ClosestWalk<MollerTrumbore> w;
    w.active = false;
    w.hit = -1;
    w.dist = 0.0f;
    uint32_t k = 0u, next = 0u;
    bool busy = false;
    for (;;) {
        if (busy && !w.active) {
            origin[k] = make_float4(w.dist, __int_as_float(w.hit), 0.0f, 0.0f);
            dir[k].w = kCandWalked;
            busy = false;
        }
        if (next < count) {
            uint64_t idle = __ballot(!busy);
            uint32_t i = next + rank_in(idle);
            if (!busy && i < count) {
                k = list[i];
                busy = true;
                float4 o = origin[k], d = dir[k];
                F3 d3 = F3{d.x, d.y, d.z};
                closest_begin(p.v, w, MollerTrumbore{F3{o.x, o.y, o.z}, d3}, d3, max_distance);
            }
            next += (uint32_t)__popcll(idle);
        }
        if (__ballot(busy) == 0ull) break;
        if (busy && w.active) closest_step(p.v, w, ring, stride);
    }
}

RGI_DEV float finish_pairwise(const GiPass& p, Pairwise& m, F3 position, F3 n, const Res& c, float cweight, float random)
{
    if (m.has_c) {
        float w = (m.c_share * c.target) * cweight;
        if (stream_in(m.combined, c.contribution, c.position, c.normal, c.id, c.target, w, random))
            m.neighbour_selected = false;
    }
    float cw = m.combined.target > 0.000001f ? m.combined.wsum / (m.total * m.combined.target) : 0.0f;
    if (cw > 0.0f && m.neighbour_selected) {
        if (!visible_between(p, position, n, m.combined.position, m.combined.id)) cw = 0.0f;
    }
    return cw;
}

RGI_DEV F2 pixel_uv(const GiPass& p, uint32_t index)
{
    uint32_t w = (uint32_t)p.hw;
    return F2{((float)(index % w) + 0.5f) * p.texel.x, ((float)(index / w) + 0.5f) * p.texel.y};
}

} // namespace rgi
