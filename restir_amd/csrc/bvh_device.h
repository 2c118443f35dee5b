// BVH traversal on gfx950 over the node layout of src/bvh.h:9-22 (nodes of
// two float4 in depth-first order with skip links, then 64-byte triangle
// records), reproducing the reference's visit orders exactly so that hit
// triangle ids and distances are bit-exact:
//   - any hit: the stackless skip-link walk (bvh.sh:140-174);
//   - closest hit: nearer child first with a short ring stack of 8 entries per
//     lane in LDS, dropping the oldest on overflow and finishing with a
//     stackless pass when anything was dropped (bvh.sh:216-295).
// Both stop after 2048 node visits (kBvhMaxSteps, bvh.sh:23).
#pragma once

#include "device.h"

namespace rgi {

constexpr int kStackRing = 8;
constexpr int kMaxSteps = 2048;

struct SceneView {
    const float4* data;
    uint32_t node_count;
    uint32_t tri_offset;
};

RGI_DEV SceneView scene_view(const rgi_scene& s) { return SceneView{reinterpret_cast<const float4*>(s.bvh), s.node_count, s.triangle_offset}; }
RGI_DEV bool bvh_ready(const SceneView& v) { return v.node_count > 0u && v.data != nullptr; }
RGI_DEV F4 sv(const SceneView& v, uint32_t i)
{
    float4 q = v.data[i];
    return F4{q.x, q.y, q.z, q.w};
}
RGI_DEV F3 vertex(const SceneView& v, int tri, int corner) { return xyz(sv(v, v.tri_offset + (uint32_t)tri * 4u + (uint32_t)corner)); }

// Slab test: entry distance, or -1 on a miss within max_distance.
RGI_DEV float box_enter(F3 o, F3 inv, F4 lo, F4 hi, float max_distance)
{
    float t0x = (lo.x - o.x) * inv.x, t1x = (hi.x - o.x) * inv.x;
    float t0y = (lo.y - o.y) * inv.y, t1y = (hi.y - o.y) * inv.y;
    float t0z = (lo.z - o.z) * inv.z, t1z = (hi.z - o.z) * inv.z;
    float enter = maxf_(maxf_(minf_(t0x, t1x), minf_(t0y, t1y)), maxf_(minf_(t0z, t1z), 0.0f));
    float leave = minf_(minf_(maxf_(t0x, t1x), maxf_(t0y, t1y)), minf_(maxf_(t0z, t1z), max_distance));
    return enter <= leave ? enter : -1.0f;
}

// The slab test made conservative for the watertight triangle test (Pharr,
// Jakob and Humphreys, PBR 3rd ed., 3.9.2): each axis's far distance is
// scaled by 1 + 2 gamma(3), gamma(n) = n u / (1 - n u), u = 2^-24, so that
// rounding never rejects a box that the ray grazes.
RGI_DEV float box_enter_conservative(F3 o, F3 inv, F4 lo, F4 hi, float max_distance)
{
    const float grow = 1.0f + 2.0f * (3.0f * 0x1p-24f / (1.0f - 3.0f * 0x1p-24f));
    float t0x = (lo.x - o.x) * inv.x, t1x = (hi.x - o.x) * inv.x;
    float t0y = (lo.y - o.y) * inv.y, t1y = (hi.y - o.y) * inv.y;
    float t0z = (lo.z - o.z) * inv.z, t1z = (hi.z - o.z) * inv.z;
    float enter = maxf_(maxf_(minf_(t0x, t1x), minf_(t0y, t1y)), maxf_(minf_(t0z, t1z), 0.0f));
    float leave = minf_(minf_(maxf_(t0x, t1x) * grow, maxf_(t0y, t1y) * grow), minf_(maxf_(t0z, t1z) * grow, max_distance));
    return enter <= leave ? enter : -1.0f;
}

// Moeller-Trumbore, both sides, |det| < 1e-12 rejected.
RGI_DEV float tri_hit(F3 o, F3 d, F3 v0, F3 v1, F3 v2)
{
    F3 e1 = v1 - v0;
    F3 e2 = v2 - v0;
    F3 p = cross(d, e2);
    float det = dot(e1, p);
    if (fabsf(det) < 1e-12f) return -1.0f;
    float inv = 1.0f / det;
    F3 s = o - v0;
    float u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    F3 q = cross(s, e1);
    float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    return dot(e2, q) * inv;
}

// Watertight ray/triangle test (Woop, Benthin and Wald 2013, JCGT 2(1)),
// both sides. The ray's dominant axis becomes z (x and y swapped when the
// ray points down it, to keep the winding), the triangle is sheared into ray
// space once per vertex, and the signs of the three edge functions decide
// the hit. An edge function that comes out exactly zero in FP32 is
// recomputed from the same sheared coordinates in FP64, so a ray through a
// shared edge or vertex hits at least one of the triangles that meet there.
struct WatertightRay {
    F3 o;
    int kx, ky, kz;
    float sx, sy, sz;
};

RGI_DEV float axis(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

RGI_DEV WatertightRay watertight_ray(F3 o, F3 d)
{
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = ax >= ay ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = axis(d, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    return WatertightRay{o, kx, ky, kz, axis(d, kx) / dz, axis(d, ky) / dz, 1.0f / dz};
}

RGI_DEV float watertight_hit(const WatertightRay& r, F3 v0, F3 v1, F3 v2)
{
    F3 a = v0 - r.o, b = v1 - r.o, c = v2 - r.o;
    float az = axis(a, r.kz), bz = axis(b, r.kz), cz = axis(c, r.kz);
    float ax = axis(a, r.kx) - r.sx * az, ay = axis(a, r.ky) - r.sy * az;
    float bx = axis(b, r.kx) - r.sx * bz, by = axis(b, r.ky) - r.sy * bz;
    float cx = axis(c, r.kx) - r.sx * cz, cy = axis(c, r.ky) - r.sy * cz;
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
    float t = ((u * (r.sz * az) + v * (r.sz * bz)) + w * (r.sz * cz)) / det;
    return t;
}

// The two triangle tests as traversal functors, each with its box test.
struct MollerTrumbore {
    F3 o, d;
    RGI_DEV float hit(F3 v0, F3 v1, F3 v2) const { return tri_hit(o, d, v0, v1, v2); }
    RGI_DEV float box(F3 inv, F4 lo, F4 hi, float max_distance) const { return box_enter(o, inv, lo, hi, max_distance); }
};
struct Watertight {
    WatertightRay r;
    RGI_DEV float hit(F3 v0, F3 v1, F3 v2) const { return watertight_hit(r, v0, v1, v2); }
    RGI_DEV float box(F3 inv, F4 lo, F4 hi, float max_distance) const
    {
        return box_enter_conservative(r.o, inv, lo, hi, max_distance);
    }
};

// A triangle test set up for one ray.
template <class Tri> RGI_DEV Tri triangle_test(F3 o, F3 d);
template <> RGI_DEV MollerTrumbore triangle_test<MollerTrumbore>(F3 o, F3 d) { return MollerTrumbore{o, d}; }
template <> RGI_DEV Watertight triangle_test<Watertight>(F3 o, F3 d) { return Watertight{watertight_ray(o, d)}; }

RGI_DEV float safe_dir(float d) { return fabsf(d) > 1e-8f ? d : (d >= 0.0f ? 1e-8f : -1e-8f); }
RGI_DEV F3 inverse_dir(F3 d) { return F3{1.0f / safe_dir(d.x), 1.0f / safe_dir(d.y), 1.0f / safe_dir(d.z)}; }

template <class Tri>
RGI_DEV int leaf_hits(const SceneView& v, const Tri& test, int first, int count, float& dist, int hit)
{
    for (int t = first; t < first + count; ++t) {
        uint32_t b = v.tri_offset + (uint32_t)t * 4u;
        float h = test.hit(xyz(sv(v, b)), xyz(sv(v, b + 1u)), xyz(sv(v, b + 2u)));
        if (h > 0.0f && h < dist) {
            dist = h;
            hit = t;
        }
    }
    return hit;
}

template <class Tri>
RGI_DEV int walk_stackless(const SceneView& v, const Tri& test, F3 o, F3 d, bool any_hit, float& dist)
{
    F3 inv = inverse_dir(d);
    int nodes = (int)v.node_count;
    int hit = -1;
    int node = 0;
    for (int step = 0; step < kMaxSteps; ++step) {
        if (node >= nodes) break;
        F4 n0 = sv(v, (uint32_t)node * 2u);
        F4 n1 = sv(v, (uint32_t)node * 2u + 1u);
        bool leaf = n1.w >= 0.0f;
        if (test.box(inv, n0, n1, dist) < 0.0f) {
            node = leaf ? node + 1 : to_int(-n1.w) - 1;
            continue;
        }
        if (leaf) {
            hit = leaf_hits(v, test, to_int(n0.w), to_int(n1.w), dist, hit);
            if (any_hit && hit >= 0) return hit;
        }
        node = node + 1;
    }
    return hit;
}

// `ring` points at this lane's first ring entry; entries are `stride` ints
// apart (the ring is laid out slot-major across the lanes of a block).
template <class Tri>
RGI_DEV int walk_short_stack(const SceneView& v, const Tri& test, F3 o, F3 d, float& dist, int* ring, int stride)
{
    F3 inv = inverse_dir(d);
    if (test.box(inv, sv(v, 0), sv(v, 1), dist) < 0.0f) return -1;
    int top = 0, held = 0;
    bool dropped = false;
    int hit = -1;
    int node = 0;
    for (int step = 0; step < kMaxSteps; ++step) {
        F4 n0 = sv(v, (uint32_t)node * 2u);
        F4 n1 = sv(v, (uint32_t)node * 2u + 1u);
        if (n1.w >= 0.0f) {
            hit = leaf_hits(v, test, to_int(n0.w), to_int(n1.w), dist, hit);
        } else {
            int left = node + 1;
            int right = to_int(n0.w);
            float ld = test.box(inv, sv(v, (uint32_t)left * 2u), sv(v, (uint32_t)left * 2u + 1u), dist);
            float rd = test.box(inv, sv(v, (uint32_t)right * 2u), sv(v, (uint32_t)right * 2u + 1u), dist);
            bool vl = ld >= 0.0f, vr = rd >= 0.0f;
            if (vl && vr) {
                bool left_first = ld <= rd;
                ring[(top & (kStackRing - 1)) * stride] = left_first ? right : left;
                ++top;
                if (held == kStackRing) dropped = true;
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
        node = ring[(top & (kStackRing - 1)) * stride];
    }
    if (dropped) {
        int again = walk_stackless(v, test, o, d, false, dist);
        if (again >= 0) hit = again;
    }
    return hit;
}

// The two walks as resumable states, so that a lane can take up another
// ray between node visits (a queue of rays answered by every lane of a wave).
// Each visits the same nodes in the same order, and keeps the same distance
// and hit, as walk_stackless / walk_short_stack, whichever lane runs it.
//
// AnyWalk: walk_stackless's any-hit walk. any_begin sets a ray up, each
// any_step visits one node; `active` goes false when the walk ends, `hit` is
// the triangle found (or -1) and `dist` the distance as trace_any leaves it.
template <class Tri>
struct AnyWalk {
    Tri test;
    F3 inv;
    float dist;
    int node, step, hit;
    bool active;
};

template <class Tri>
RGI_DEV void any_begin(const SceneView& v, AnyWalk<Tri>& w, const Tri& test, F3 d, float max_distance)
{
    w.test = test;
    w.inv = inverse_dir(d);
    w.dist = max_distance;
    w.node = 0;
    w.step = 0;
    w.hit = -1;
    w.active = bvh_ready(v);
}

template <class Tri>
RGI_DEV void any_step(const SceneView& v, AnyWalk<Tri>& w)
{
    if (w.step >= kMaxSteps || w.node >= (int)v.node_count) {
        w.active = false;
        return;
    }
    F4 n0 = sv(v, (uint32_t)w.node * 2u);
    F4 n1 = sv(v, (uint32_t)w.node * 2u + 1u);
    bool leaf = n1.w >= 0.0f;
    ++w.step;
    if (w.test.box(w.inv, n0, n1, w.dist) < 0.0f) {
        w.node = leaf ? w.node + 1 : to_int(-n1.w) - 1;
        return;
    }
    if (leaf) {
        w.hit = leaf_hits(v, w.test, to_int(n0.w), to_int(n1.w), w.dist, w.hit);
        if (w.hit >= 0) {
            w.active = false;
            return;
        }
    }
    w.node = w.node + 1;
}

// ClosestWalk: walk_short_stack's closest-hit walk, then its stackless
// re-pass when the ring dropped an entry. closest_step visits one node of
// whichever phase is running; `ring` and `stride` as for walk_short_stack.
template <class Tri>
struct ClosestWalk {
    Tri test;
    F3 inv;
    float dist;
    int hit, node, step, top, held;
    bool dropped, repass, active;
};

template <class Tri>
RGI_DEV void closest_begin(const SceneView& v, ClosestWalk<Tri>& w, const Tri& test, F3 d, float max_distance)
{
    w.test = test;
    w.inv = inverse_dir(d);
    w.dist = max_distance;
    w.hit = -1;
    w.node = 0;
    w.step = 0;
    w.top = 0;
    w.held = 0;
    w.dropped = false;
    w.repass = false;
    w.active = bvh_ready(v) && test.box(w.inv, sv(v, 0), sv(v, 1), w.dist) >= 0.0f;
}

template <class Tri>
RGI_DEV void closest_end_stack(ClosestWalk<Tri>& w)
{
    if (w.dropped) {
        w.repass = true;
        w.node = 0;
        w.step = 0;
    } else {
        w.active = false;
    }
}

template <class Tri>
RGI_DEV void closest_step(const SceneView& v, ClosestWalk<Tri>& w, int* ring, int stride)
{
    if (w.repass) {
        // walk_stackless, closest hit, continuing from the stack walk's
        // distance and hit.
        if (w.step >= kMaxSteps || w.node >= (int)v.node_count) {
            w.active = false;
            return;
        }
        F4 n0 = sv(v, (uint32_t)w.node * 2u);
        F4 n1 = sv(v, (uint32_t)w.node * 2u + 1u);
        bool leaf = n1.w >= 0.0f;
        ++w.step;
        if (w.test.box(w.inv, n0, n1, w.dist) < 0.0f) {
            w.node = leaf ? w.node + 1 : to_int(-n1.w) - 1;
            return;
        }
        if (leaf) w.hit = leaf_hits(v, w.test, to_int(n0.w), to_int(n1.w), w.dist, w.hit);
        w.node = w.node + 1;
        return;
    }
    if (w.step >= kMaxSteps) {
        closest_end_stack(w);
        return;
    }
    ++w.step;
    F4 n0 = sv(v, (uint32_t)w.node * 2u);
    F4 n1 = sv(v, (uint32_t)w.node * 2u + 1u);
    if (n1.w >= 0.0f) {
        w.hit = leaf_hits(v, w.test, to_int(n0.w), to_int(n1.w), w.dist, w.hit);
    } else {
        int left = w.node + 1;
        int right = to_int(n0.w);
        float ld = w.test.box(w.inv, sv(v, (uint32_t)left * 2u), sv(v, (uint32_t)left * 2u + 1u), w.dist);
        float rd = w.test.box(w.inv, sv(v, (uint32_t)right * 2u), sv(v, (uint32_t)right * 2u + 1u), w.dist);
        bool vl = ld >= 0.0f, vr = rd >= 0.0f;
        if (vl && vr) {
            bool left_first = ld <= rd;
            ring[(w.top & (kStackRing - 1)) * stride] = left_first ? right : left;
            ++w.top;
            if (w.held == kStackRing) w.dropped = true;
            else ++w.held;
            w.node = left_first ? left : right;
            return;
        }
        if (vl || vr) {
            w.node = vl ? left : right;
            return;
        }
    }
    if (w.held == 0) {
        closest_end_stack(w);
        return;
    }
    --w.top;
    --w.held;
    w.node = ring[(w.top & (kStackRing - 1)) * stride];
}

// Persistent form. Group g (workgroups b = g mod kTraceGroups) claims
// chunks g, g + kTraceGroups, ... of `chunk` rays from its own counter
// (each on a 128-byte line) until they pass `count`; every wave reaches that
// exit test, so the grid always drains. The counters are zeroed before the
// launch.
constexpr uint32_t kTraceGroups = 8, kTraceChunk = 256, kTraceLine = 32;

// Lanes keep traversing while at least kRefillBelow of them are busy; below
// that the wave stops to hand the idle lanes new rays (Aila and Laine 2009's
// dynamic ray fetch), so the hand-out costs little per node visited.
constexpr int kRefillBelow = 40;

// Answers rays[0, count) (two float4 per ray: origin and max distance,
// direction) into hit[] and dist[] with every lane of every wave that calls
// it; `ring` is this lane's closest-hit ring entry, entries `stride` ints
// apart. `counters` holds kTraceGroups counters kTraceLine words apart,
// zeroed before the launch. Smaller chunks spread fewer rays over more waves.
template <class Tri, bool kAny>
RGI_DEV void trace_persistent(const SceneView& v, const float4* rays, uint32_t count, int32_t* hit, float* dist,
                              uint32_t* counters, int* ring, int stride, uint32_t chunk)
{
    const uint32_t group = blockIdx.x % kTraceGroups;
    uint32_t* counter = counters + group * kTraceLine;
    AnyWalk<Tri> aw;
    ClosestWalk<Tri> cw;
    aw.active = false;
    cw.active = false;
    uint32_t k = 0u, cur = 0u, end = 0u;
    bool busy = false, exhausted = false;
    for (;;) {
        // Hand the idle lanes the next rays of the wave's chunk (claiming a
        // new chunk when it is used up).
        uint64_t idle = __ballot(!busy);
        while (idle != 0ull && !exhausted) {
            if (cur >= end) {
                uint32_t c = 0u;
                if ((threadIdx.x & 63u) == 0u) c = atomicAdd(counter, 1u);
                uint32_t first = ((uint32_t)__shfl((int)c, 0) * kTraceGroups + group) * chunk;
                if (first >= count) {
                    exhausted = true;
                    break;
                }
                cur = first;
                end = count - first < chunk ? count : first + chunk;
            }
            uint32_t ray = cur + rank_in(idle);
            bool take = !busy && ray < end;
            if (take) {
                k = ray;
                busy = true;
                float4 a = rays[2u * k], b = rays[2u * k + 1u];
                F3 o = F3{a.x, a.y, a.z}, d = F3{b.x, b.y, b.z};
                Tri test = triangle_test<Tri>(o, d);
                if (kAny) any_begin(v, aw, test, d, a.w);
                else closest_begin(v, cw, test, d, a.w);
            }
            uint32_t taken = (uint32_t)__popcll(idle);
            cur = end - cur < taken ? end : cur + taken;
            idle = __ballot(!busy);
        }
        const int keep = exhausted ? 1 : kRefillBelow;
        if (__popcll(__ballot(busy)) == 0) break;
        // Traverse until fewer than `keep` lanes are busy.
        for (;;) {
            if (busy) {
                if (kAny) {
                    if (aw.active) any_step(v, aw);
                } else {
                    if (cw.active) closest_step(v, cw, ring, stride);
                }
                if (!(kAny ? aw.active : cw.active)) {
                    hit[k] = kAny ? aw.hit : cw.hit;
                    dist[k] = kAny ? aw.dist : cw.dist;
                    busy = false;
                }
            }
            if (__popcll(__ballot(busy)) < keep) break;
        }
    }
}

// traceBvh's any-hit form: -1 when nothing is hit within max_distance;
// `dist` ends at the first hit found (or max_distance). The reference's
// Moeller-Trumbore test unless `watertight`.
RGI_DEV int trace_any(const SceneView& v, F3 o, F3 d, float max_distance, float& dist, bool watertight = false)
{
    dist = max_distance;
    if (!bvh_ready(v)) return -1;
    if (watertight) return walk_stackless(v, Watertight{watertight_ray(o, d)}, o, d, true, dist);
    return walk_stackless(v, MollerTrumbore{o, d}, o, d, true, dist);
}

RGI_DEV int trace_any(const SceneView& v, F3 o, F3 d, float max_distance)
{
    float dist;
    return trace_any(v, o, d, max_distance, dist);
}

RGI_DEV int trace_closest(const SceneView& v, F3 o, F3 d, float max_distance, float& dist, int* ring, int stride,
                          bool watertight = false)
{
    dist = max_distance;
    if (!bvh_ready(v)) return -1;
    if (watertight) return walk_short_stack(v, Watertight{watertight_ray(o, d)}, o, d, dist, ring, stride);
    return walk_short_stack(v, MollerTrumbore{o, d}, o, d, dist, ring, stride);
}

RGI_DEV F3 tri_albedo(const SceneView& v, int tri) { return xyz(sv(v, v.tri_offset + (uint32_t)tri * 4u + 3u)); }
RGI_DEV F3 tri_emission(const SceneView& v, int tri)
{
    uint32_t b = v.tri_offset + (uint32_t)tri * 4u;
    return F3{sv(v, b + 1u).w, sv(v, b + 2u).w, sv(v, b + 3u).w};
}

} // namespace rgi
