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

RGI_DEV float safe_dir(float d) { return fabsf(d) > 1e-8f ? d : (d >= 0.0f ? 1e-8f : -1e-8f); }
RGI_DEV F3 inverse_dir(F3 d) { return F3{1.0f / safe_dir(d.x), 1.0f / safe_dir(d.y), 1.0f / safe_dir(d.z)}; }

RGI_DEV int leaf_hits(const SceneView& v, F3 o, F3 d, int first, int count, float& dist, int hit)
{
    for (int t = first; t < first + count; ++t) {
        uint32_t b = v.tri_offset + (uint32_t)t * 4u;
        float h = tri_hit(o, d, xyz(sv(v, b)), xyz(sv(v, b + 1u)), xyz(sv(v, b + 2u)));
        if (h > 0.0f && h < dist) {
            dist = h;
            hit = t;
        }
    }
    return hit;
}

RGI_DEV int walk_stackless(const SceneView& v, F3 o, F3 d, bool any_hit, float& dist)
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
        if (box_enter(o, inv, n0, n1, dist) < 0.0f) {
            node = leaf ? node + 1 : to_int(-n1.w) - 1;
            continue;
        }
        if (leaf) {
            hit = leaf_hits(v, o, d, to_int(n0.w), to_int(n1.w), dist, hit);
            if (any_hit && hit >= 0) return hit;
        }
        node = node + 1;
    }
    return hit;
}

// `ring` points at this lane's first ring entry; entries are `stride` ints
// apart (the ring is laid out slot-major across the lanes of a block).
RGI_DEV int walk_short_stack(const SceneView& v, F3 o, F3 d, float& dist, int* ring, int stride)
{
    F3 inv = inverse_dir(d);
    if (box_enter(o, inv, sv(v, 0), sv(v, 1), dist) < 0.0f) return -1;
    int top = 0, held = 0;
    bool dropped = false;
    int hit = -1;
    int node = 0;
    for (int step = 0; step < kMaxSteps; ++step) {
        F4 n0 = sv(v, (uint32_t)node * 2u);
        F4 n1 = sv(v, (uint32_t)node * 2u + 1u);
        if (n1.w >= 0.0f) {
            hit = leaf_hits(v, o, d, to_int(n0.w), to_int(n1.w), dist, hit);
        } else {
            int left = node + 1;
            int right = to_int(n0.w);
            float ld = box_enter(o, inv, sv(v, (uint32_t)left * 2u), sv(v, (uint32_t)left * 2u + 1u), dist);
            float rd = box_enter(o, inv, sv(v, (uint32_t)right * 2u), sv(v, (uint32_t)right * 2u + 1u), dist);
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
        int again = walk_stackless(v, o, d, false, dist);
        if (again >= 0) hit = again;
    }
    return hit;
}

// traceBvh's any-hit form: -1 when nothing is hit within max_distance;
// `dist` ends at the first hit found (or max_distance).
RGI_DEV int trace_any(const SceneView& v, F3 o, F3 d, float max_distance, float& dist)
{
    dist = max_distance;
    if (!bvh_ready(v)) return -1;
    return walk_stackless(v, o, d, true, dist);
}

RGI_DEV int trace_any(const SceneView& v, F3 o, F3 d, float max_distance)
{
    float dist;
    return trace_any(v, o, d, max_distance, dist);
}

RGI_DEV int trace_closest(const SceneView& v, F3 o, F3 d, float max_distance, float& dist, int* ring, int stride)
{
    dist = max_distance;
    if (!bvh_ready(v)) return -1;
    return walk_short_stack(v, o, d, dist, ring, stride);
}

RGI_DEV F3 tri_albedo(const SceneView& v, int tri) { return xyz(sv(v, v.tri_offset + (uint32_t)tri * 4u + 3u)); }
RGI_DEV F3 tri_emission(const SceneView& v, int tri)
{
    uint32_t b = v.tri_offset + (uint32_t)tri * 4u;
    return F3{sv(v, b + 1u).w, sv(v, b + 2u).w, sv(v, b + 3u).w};
}

} // namespace rgi
