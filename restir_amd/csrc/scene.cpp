// Scene data layer of the hot path (host): seeded procedural scenes, the
// binned-SAH BVH in the reference's GPU layout (src/bvh.h:9-22), the light
// list (src/lights.h:7-14) and camera matrices. See restir_amd_scene.h.
#include "../../include/restir_amd.h"
#include "../../include/restir_amd_scene.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

struct Vec { float x, y, z; };
inline Vec operator+(Vec a, Vec b) { return Vec{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec operator-(Vec a, Vec b) { return Vec{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec operator*(Vec a, float s) { return Vec{a.x * s, a.y * s, a.z * s}; }
inline Vec cross(Vec a, Vec b) { return Vec{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float axis_of(const Vec& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// A PCG stream for the generators: same permutation as the passes' hash.
struct Rng {
    uint32_t state;
    float next()
    {
        state = state * 747796405u + 2891336453u;
        uint32_t w = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        w = (w >> 22u) ^ w;
        return (float)(w >> 8) * (1.0f / 16777216.0f);
    }
    float range(float lo, float hi) { return lo + (hi - lo) * next(); }
};

struct Builder {
    std::vector<float> pos, alb, emi;
    std::vector<uint32_t> mat;
    std::vector<float> materials;

    uint32_t add_material(float roughness, float metallic)
    {
        materials.insert(materials.end(), {roughness, metallic, 0.0f, 0.0f});
        return (uint32_t)(materials.size() / 4 - 1);
    }
    void tri(Vec a, Vec b, Vec c, Vec albedo, Vec emission, uint32_t material)
    {
        pos.insert(pos.end(), {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z});
        alb.insert(alb.end(), {albedo.x, albedo.y, albedo.z});
        emi.insert(emi.end(), {emission.x, emission.y, emission.z});
        mat.push_back(material);
    }
    // Quad a-b-c-d (counter-clockwise seen from its front).
    void quad(Vec a, Vec b, Vec c, Vec d, Vec albedo, Vec emission, uint32_t material)
    {
        tri(a, b, c, albedo, emission, material);
        tri(a, c, d, albedo, emission, material);
    }
    // A quad split into nu x nv quads.
    void grid_quad(Vec o, Vec u, Vec v, int nu, int nv, Vec albedo, Vec emission, uint32_t material, Rng* per_tri)
    {
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nu; ++i) {
                Vec a = o + u * ((float)i / nu) + v * ((float)j / nv);
                Vec b = o + u * ((float)(i + 1) / nu) + v * ((float)j / nv);
                Vec c = o + u * ((float)(i + 1) / nu) + v * ((float)(j + 1) / nv);
                Vec d = o + u * ((float)i / nu) + v * ((float)(j + 1) / nv);
                Vec al = albedo;
                if (per_tri) al = Vec{per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f)};
                quad(a, b, c, d, al, emission, material);
            }
    }
    // Axis-aligned box of half extents h centred at c, turned by yaw about +y.
    void box(Vec c, Vec h, float yaw, Vec albedo, Vec emission, uint32_t material)
    {
        float cs = cosf(yaw), sn = sinf(yaw);
        auto p = [&](float sx, float sy, float sz) {
            Vec l = Vec{sx * h.x, sy * h.y, sz * h.z};
            return Vec{c.x + cs * l.x + sn * l.z, c.y + l.y, c.z - sn * l.x + cs * l.z};
        };
        quad(p(-1, -1, 1), p(1, -1, 1), p(1, 1, 1), p(-1, 1, 1), albedo, emission, material);     // +z
        quad(p(1, -1, -1), p(-1, -1, -1), p(-1, 1, -1), p(1, 1, -1), albedo, emission, material); // -z
        quad(p(1, -1, 1), p(1, -1, -1), p(1, 1, -1), p(1, 1, 1), albedo, emission, material);     // +x
        quad(p(-1, -1, -1), p(-1, -1, 1), p(-1, 1, 1), p(-1, 1, -1), albedo, emission, material); // -x
        quad(p(-1, 1, 1), p(1, 1, 1), p(1, 1, -1), p(-1, 1, -1), albedo, emission, material);     // +y
        quad(p(-1, -1, -1), p(1, -1, -1), p(1, -1, 1), p(-1, -1, 1), albedo, emission, material); // -y
    }
    // Vertical cylinder (no caps) of `segments` around and `rings` up.
    void cylinder(Vec base, float radius, float height, int segments, int rings, Vec albedo, uint32_t material, Rng* per_tri)
    {
        const float two_pi = 6.28318530718f;
        for (int r = 0; r < rings; ++r)
            for (int s = 0; s < segments; ++s) {
                float a0 = two_pi * (float)s / segments, a1 = two_pi * (float)(s + 1) / segments;
                float y0 = base.y + height * (float)r / rings, y1 = base.y + height * (float)(r + 1) / rings;
                Vec p00 = Vec{base.x + radius * cosf(a0), y0, base.z + radius * sinf(a0)};
                Vec p10 = Vec{base.x + radius * cosf(a1), y0, base.z + radius * sinf(a1)};
                Vec p11 = Vec{base.x + radius * cosf(a1), y1, base.z + radius * sinf(a1)};
                Vec p01 = Vec{base.x + radius * cosf(a0), y1, base.z + radius * sinf(a0)};
                Vec al = albedo;
                if (per_tri) al = Vec{per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f)};
                quad(p00, p01, p11, p10, al, Vec{0, 0, 0}, material);
            }
    }
    // A semicircular arch band from a to b (horizontal span), springing at
    // height y, of radial thickness `thick` and depth `depth` along the
    // horizontal normal of the span.
    void arch(Vec a, Vec b, float y, float depth, float thick, int segments, Vec albedo, uint32_t material, Rng* per_tri)
    {
        Vec span = b - a;
        float len = sqrtf(dot(span, span));
        Vec ux = span * (1.0f / len);
        Vec nz = Vec{-ux.z, 0.0f, ux.x};
        Vec mid = (a + b) * 0.5f;
        float radius = 0.5f * len;
        const float pi = 3.14159265359f;
        for (int s = 0; s < segments; ++s) {
            float t0 = pi * (float)s / segments, t1 = pi * (float)(s + 1) / segments;
            for (int side = 0; side < 2; ++side) {
                Vec off = nz * (side ? depth * 0.5f : -depth * 0.5f);
                auto at = [&](float t, float r) {
                    return mid + off + ux * (r * cosf(t)) + Vec{0.0f, y + r * sinf(t), 0.0f};
                };
                Vec al = albedo;
                if (per_tri) al = Vec{per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f)};
                quad(at(t0, radius), at(t1, radius), at(t1, radius + thick), at(t0, radius + thick), al, Vec{0, 0, 0},
                     material);
            }
            // The intrados, under the band.
            auto under = [&](float t, float side) {
                return mid + nz * (side * depth * 0.5f) + ux * (radius * cosf(t)) + Vec{0.0f, y + radius * sinf(t), 0.0f};
            };
            Vec al = albedo;
            if (per_tri) al = Vec{per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f), per_tri->range(0.1f, 0.9f)};
            quad(under(t0, -1.0f), under(t0, 1.0f), under(t1, 1.0f), under(t1, -1.0f), al, Vec{0, 0, 0}, material);
        }
    }
};

// The open-front Cornell box: walls x in [-1, 1], y in [0, 2], z in [-1, 1],
// a short and a tall box, and a 0.5 x 0.4 emissive quad of radiance 17 just
// below the ceiling; albedos and roughnesses drawn from the seed.
void make_cornell(Builder& b, uint32_t seed)
{
    Rng rng{seed * 2654435761u + 0x9e3779b9u};
    auto albedo = [&]() { return Vec{rng.range(0.05f, 0.8f), rng.range(0.05f, 0.8f), rng.range(0.05f, 0.8f)}; };
    auto material = [&]() { return b.add_material(rng.range(0.3f, 0.9f), 0.0f); };
    Vec zero{0, 0, 0};
    Vec a;
    uint32_t m;
    a = albedo(); m = material();
    b.quad(Vec{-1, 0, 1}, Vec{1, 0, 1}, Vec{1, 0, -1}, Vec{-1, 0, -1}, a, zero, m);   // floor
    a = albedo(); m = material();
    b.quad(Vec{-1, 2, -1}, Vec{1, 2, -1}, Vec{1, 2, 1}, Vec{-1, 2, 1}, a, zero, m);   // ceiling
    a = albedo(); m = material();
    b.quad(Vec{-1, 0, -1}, Vec{1, 0, -1}, Vec{1, 2, -1}, Vec{-1, 2, -1}, a, zero, m); // back
    a = albedo(); m = material();
    b.quad(Vec{-1, 0, 1}, Vec{-1, 0, -1}, Vec{-1, 2, -1}, Vec{-1, 2, 1}, a, zero, m); // left
    a = albedo(); m = material();
    b.quad(Vec{1, 0, -1}, Vec{1, 0, 1}, Vec{1, 2, 1}, Vec{1, 2, -1}, a, zero, m);     // right
    const float deg = 3.14159265359f / 180.0f;
    a = albedo(); m = material();
    b.box(Vec{0.35f, 0.3f, 0.3f}, Vec{0.3f, 0.3f, 0.3f}, -17.0f * deg, a, zero, m);
    a = albedo(); m = material();
    b.box(Vec{-0.35f, 0.6f, -0.3f}, Vec{0.3f, 0.6f, 0.3f}, 17.0f * deg, a, zero, m);
    a = albedo(); m = material();
    Vec e{17.0f, 17.0f, 17.0f};
    b.quad(Vec{-0.25f, 1.98f, -0.2f}, Vec{0.25f, 1.98f, -0.2f}, Vec{0.25f, 1.98f, 0.2f}, Vec{-0.25f, 1.98f, 0.2f}, a, e, m);
}

// A three-storey courtyard: tessellated ground, gallery floors and outer
// walls, 24-segment columns along the courtyard edge with arches between
// them, and 32 emissive quads under the gallery ceilings. The tessellation
// is scaled so the whole comes to about `budget` triangles.
void make_atrium(Builder& b, uint32_t seed, uint32_t budget)
{
    Rng rng{seed * 2654435761u + 0x7f4a7c15u};
    uint32_t stone = b.add_material(0.8f, 0.0f);
    uint32_t column_material = b.add_material(0.5f, 0.0f);
    uint32_t lamp = b.add_material(0.9f, 0.0f);
    const float inner = 10.0f, outer = 16.0f, storey = 4.0f;
    const int storeys = 3;
    const int columns_per_side = 6;                 // per storey and side
    const int column_count = storeys * 4 * columns_per_side;
    // Fixed part per column: 24 segments x rings x 2, arches 24 segments x 3 quads x 2.
    // Choose rings and the slab tessellation n from the budget.
    double target = budget > 4096 ? (double)budget : 4096.0;
    int rings = 8;
    int n = 8;
    for (int iter = 0; iter < 100000; ++iter) {
        double cols = (double)column_count * (24.0 * rings * 2.0 + 24.0 * 3.0 * 2.0);
        double slabs = (double)(storeys + 1) * 4.0 * n * n * 2.0 * 0.5 + (double)n * n * 2.0; // gallery rings + courtyard
        double walls = 4.0 * n * n * 2.0;
        double total = cols + slabs + walls + 64.0;
        if (total < target * 0.97) {
            if (cols < 0.5 * total && rings < 64) ++rings;
            else ++n;
        } else {
            break;
        }
    }
    Vec zero{0, 0, 0};
    // Courtyard ground.
    b.grid_quad(Vec{-inner, 0, inner}, Vec{2 * inner, 0, 0}, Vec{0, 0, -2 * inner}, n, n, zero, zero, stone, &rng);
    // Gallery floors (and the roof) as four slabs per level between inner and outer.
    for (int level = 0; level <= storeys; ++level) {
        float y = storey * (float)level;
        int half = n / 2 > 1 ? n / 2 : 1;
        b.grid_quad(Vec{-outer, y, outer}, Vec{2 * outer, 0, 0}, Vec{0, 0, -(outer - inner)}, n, half, zero, zero, stone, &rng);
        b.grid_quad(Vec{-outer, y, -inner}, Vec{2 * outer, 0, 0}, Vec{0, 0, -(outer - inner)}, n, half, zero, zero, stone, &rng);
        b.grid_quad(Vec{-outer, y, inner}, Vec{outer - inner, 0, 0}, Vec{0, 0, -2 * inner}, half, n, zero, zero, stone, &rng);
        b.grid_quad(Vec{inner, y, inner}, Vec{outer - inner, 0, 0}, Vec{0, 0, -2 * inner}, half, n, zero, zero, stone, &rng);
        if (level > 0) {
            // Gallery ceiling underside, facing down.
            float yc = y - 0.01f;
            b.grid_quad(Vec{-outer, yc, -outer}, Vec{2 * outer, 0, 0}, Vec{0, 0, outer - inner}, n / 4 + 1, 1, zero, zero, stone, &rng);
            b.grid_quad(Vec{-outer, yc, inner}, Vec{2 * outer, 0, 0}, Vec{0, 0, outer - inner}, n / 4 + 1, 1, zero, zero, stone, &rng);
        }
    }
    // Outer walls.
    float height = storey * storeys;
    b.grid_quad(Vec{-outer, 0, -outer}, Vec{2 * outer, 0, 0}, Vec{0, height, 0}, n, n, zero, zero, stone, &rng);
    b.grid_quad(Vec{outer, 0, outer}, Vec{-2 * outer, 0, 0}, Vec{0, height, 0}, n, n, zero, zero, stone, &rng);
    b.grid_quad(Vec{-outer, 0, outer}, Vec{0, 0, -2 * outer}, Vec{0, height, 0}, n, n, zero, zero, stone, &rng);
    b.grid_quad(Vec{outer, 0, -outer}, Vec{0, 0, 2 * outer}, Vec{0, height, 0}, n, n, zero, zero, stone, &rng);
    // Columns and arches along the courtyard edge.
    const float radius = 0.4f, spring = 2.6f;
    for (int level = 0; level < storeys; ++level) {
        float y = storey * (float)level;
        for (int side = 0; side < 4; ++side) {
            for (int c = 0; c < columns_per_side; ++c) {
                float t0 = -inner + 2.0f * inner * (float)c / columns_per_side;
                float t1 = -inner + 2.0f * inner * (float)(c + 1) / columns_per_side;
                Vec p0, p1;
                if (side == 0) { p0 = Vec{t0, y, inner}; p1 = Vec{t1, y, inner}; }
                else if (side == 1) { p0 = Vec{inner, y, -t0}; p1 = Vec{inner, y, -t1}; }
                else if (side == 2) { p0 = Vec{-t0, y, -inner}; p1 = Vec{-t1, y, -inner}; }
                else { p0 = Vec{-inner, y, t0}; p1 = Vec{-inner, y, t1}; }
                b.cylinder(p0, radius, spring, 24, rings, zero, column_material, &rng);
                b.arch(p0, p1, y + spring, 0.8f, 0.3f, 24, zero, column_material, &rng);
            }
        }
    }
    // 32 emissive quads: 8 per side hung under the gallery ceilings of the
    // three storeys and the roof edge.
    int lamps = 0;
    for (int level = 0; level < storeys + 1 && lamps < 32; ++level)
        for (int k = 0; k < 8 && lamps < 32; ++k, ++lamps) {
            float y = storey * (float)(level + 1) - 0.05f;
            if (level == storeys) y = storey * storeys + 0.6f;
            float angle = 6.28318530718f * ((float)k + 0.5f) / 8.0f;
            float r = 0.5f * (inner + outer);
            Vec c{r * cosf(angle), y, r * sinf(angle)};
            if (level == storeys) c = Vec{0.55f * inner * cosf(angle), y + 4.0f, 0.55f * inner * sinf(angle)};
            Vec e{rng.range(5.0f, 20.0f), rng.range(5.0f, 20.0f), rng.range(5.0f, 20.0f)};
            Vec al{0.8f, 0.8f, 0.8f};
            b.quad(c + Vec{-0.4f, 0, 0.3f}, c + Vec{-0.4f, 0, -0.3f}, c + Vec{0.4f, 0, -0.3f}, c + Vec{0.4f, 0, 0.3f}, al, e, lamp);
        }
}

// A floor with `count` random boxes, one in eight of them emissive on top.
void make_boxfield(Builder& b, uint32_t seed, uint32_t count)
{
    Rng rng{seed * 2654435761u + 0x3c6ef372u};
    uint32_t floor_m = b.add_material(0.7f, 0.0f);
    uint32_t box_m = b.add_material(0.5f, 0.0f);
    float extent = 2.0f + sqrtf((float)count) * 0.8f;
    Vec zero{0, 0, 0};
    b.quad(Vec{-extent, 0, extent}, Vec{extent, 0, extent}, Vec{extent, 0, -extent}, Vec{-extent, 0, -extent},
           Vec{0.6f, 0.6f, 0.6f}, zero, floor_m);
    for (uint32_t i = 0; i < count; ++i) {
        Vec h{rng.range(0.1f, 0.5f), rng.range(0.1f, 1.0f), rng.range(0.1f, 0.5f)};
        Vec c{rng.range(-extent, extent), h.y, rng.range(-extent, extent)};
        Vec al{rng.range(0.1f, 0.9f), rng.range(0.1f, 0.9f), rng.range(0.1f, 0.9f)};
        float yaw = rng.range(0.0f, 3.14159265f);
        Vec e = (i % 8u) == 0u ? Vec{rng.range(2.0f, 10.0f), rng.range(2.0f, 10.0f), rng.range(2.0f, 10.0f)} : zero;
        b.box(c, h, yaw, al, e, box_m);
    }
}

// ------------------------------------------------------------------ BVH

struct Box {
    Vec lo{3.4e38f, 3.4e38f, 3.4e38f};
    Vec hi{-3.4e38f, -3.4e38f, -3.4e38f};
    void add(Vec p)
    {
        lo = Vec{std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = Vec{std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    void add(const Box& o)
    {
        lo = Vec{std::min(lo.x, o.lo.x), std::min(lo.y, o.lo.y), std::min(lo.z, o.lo.z)};
        hi = Vec{std::max(hi.x, o.hi.x), std::max(hi.y, o.hi.y), std::max(hi.z, o.hi.z)};
    }
    float area() const
    {
        if (lo.x > hi.x) return 0.0f;
        Vec d = hi - lo;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
};

constexpr int kBins = 12;
constexpr uint32_t kLeafSize = 4;
constexpr uint32_t kForcedSplit = 16;
constexpr uint32_t kMaxDepth = 40;

// Builds the tree straight into depth-first order: a node's left child is
// written next, its right child after the whole left subtree, and the skip
// link is known once the right subtree is done.
struct BvhWriter {
    const std::vector<Box>& tri_box;
    const std::vector<Vec>& centroid;
    std::vector<uint32_t>& order;
    std::vector<float> nodes; // 8 floats per node
    uint32_t depth = 0;

    uint32_t build(uint32_t begin, uint32_t end, uint32_t level)
    {
        uint32_t index = (uint32_t)(nodes.size() / 8);
        nodes.resize(nodes.size() + 8, 0.0f);
        Box bounds, cbounds;
        for (uint32_t i = begin; i < end; ++i) {
            bounds.add(tri_box[order[i]]);
            cbounds.add(centroid[order[i]]);
        }
        float* n = &nodes[(size_t)index * 8];
        n[0] = bounds.lo.x; n[1] = bounds.lo.y; n[2] = bounds.lo.z;
        n[4] = bounds.hi.x; n[5] = bounds.hi.y; n[6] = bounds.hi.z;
        uint32_t count = end - begin;
        uint32_t middle = begin;
        bool split = false;
        if (count > 1 && level < kMaxDepth) {
            // Binned SAH over the centroid bounds of each axis.
            float best = 3.4e38f;
            int best_axis = -1, best_plane = 0;
            for (int axis = 0; axis < 3; ++axis) {
                float lo = axis_of(cbounds.lo, axis), extent = axis_of(cbounds.hi, axis) - lo;
                if (extent <= 0.0f) continue;
                float scale = (float)kBins / extent;
                Box bin_box[kBins];
                uint32_t bin_count[kBins] = {};
                for (uint32_t i = begin; i < end; ++i) {
                    int bin = std::min((int)((axis_of(centroid[order[i]], axis) - lo) * scale), kBins - 1);
                    bin_box[bin].add(tri_box[order[i]]);
                    ++bin_count[bin];
                }
                float right_area[kBins] = {};
                uint32_t right_count[kBins] = {};
                Box acc;
                uint32_t total = 0;
                for (int bin = kBins - 1; bin > 0; --bin) {
                    acc.add(bin_box[bin]);
                    total += bin_count[bin];
                    right_area[bin] = acc.area();
                    right_count[bin] = total;
                }
                Box left;
                uint32_t left_count = 0;
                for (int plane = 1; plane < kBins; ++plane) {
                    left.add(bin_box[plane - 1]);
                    left_count += bin_count[plane - 1];
                    if (left_count == 0 || right_count[plane] == 0) continue;
                    float cost = left.area() * (float)left_count + right_area[plane] * (float)right_count[plane];
                    if (cost < best) {
                        best = cost;
                        best_axis = axis;
                        best_plane = plane;
                    }
                }
            }
            float leaf_cost = bounds.area() * (float)count;
            if (best_axis >= 0 && (best < leaf_cost || count > kLeafSize)) {
                float lo = axis_of(cbounds.lo, best_axis);
                float scale = (float)kBins / (axis_of(cbounds.hi, best_axis) - lo);
                auto goes_left = [&](uint32_t t) {
                    return std::min((int)((axis_of(centroid[t], best_axis) - lo) * scale), kBins - 1) < best_plane;
                };
                uint32_t i = begin, j = end;
                while (true) {
                    while (i < j && goes_left(order[i])) ++i;
                    while (i < j && !goes_left(order[j - 1])) --j;
                    if (i >= j) break;
                    std::swap(order[i], order[j - 1]);
                    ++i;
                    --j;
                }
                middle = i;
                split = true;
            } else if (count > kForcedSplit) {
                middle = begin + count / 2;
                split = true;
            }
        }
        if (!split) {
            n[3] = (float)begin;
            n[7] = (float)count;
            depth = std::max(depth, level);
            return index;
        }
        build(begin, middle, level + 1);
        uint32_t right = build(middle, end, level + 1);
        n = &nodes[(size_t)index * 8];
        n[3] = (float)right;
        n[7] = -(float)(nodes.size() / 8 + 1);
        return index;
    }
};

double luminance(const float* c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

// A saturated colour of hue h in [0, 1): each channel 1 - 0.8 * a tent of the hue.
void hue_rgb(float h, float* rgb)
{
    for (int c = 0; c < 3; ++c) {
        float k = fmodf(h * 6.0f + (float)(4 - 2 * c) + 6.0f, 6.0f);
        rgb[c] = 1.0f - fmaxf(0.0f, fminf(fminf(k, 4.0f - k), 1.0f)) * 0.8f;
    }
}

void mat_mul(const double* a, const double* b, double* out) // out = a * b, column major
{
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            out[c * 4 + r] = s;
        }
}

bool mat_inverse(const double* m, double* out)
{
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int pivot = c;
        for (int r = c + 1; r < 4; ++r)
            if (fabs(a[r][c]) > fabs(a[pivot][c])) pivot = r;
        if (fabs(a[pivot][c]) < 1e-300) return false;
        for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[pivot][k]);
        double inv = 1.0 / a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] *= inv;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
            }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c * 4 + r] = a[r][c + 4];
    return true;
}

// A right-handed look-at (bx::mtxLookAt semantics: the camera looks down -z,
// right = normalize(up x (eye - target))) and a perspective with cot(half
// fov) `t`, depth in [0, 1] or [-1, 1], in double; vp = proj * view.
void look_at_projection(const float* eye3, const float* target3, const double* up, double t, double aspect,
                        double n, double fa, int homogeneous_depth, double* view, double* proj, double* vp,
                        double* inv)
{
    double e[3] = {eye3[0], eye3[1], eye3[2]};
    double f[3] = {target3[0] - e[0], target3[1] - e[1], target3[2] - e[2]};
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= fl;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    if (sl < 1e-12) {
        // Looking along `up`: take +z as the reference instead.
        double alt[3] = {0.0, 0.0, 1.0};
        s[0] = f[1] * alt[2] - f[2] * alt[1];
        s[1] = f[2] * alt[0] - f[0] * alt[2];
        s[2] = f[0] * alt[1] - f[1] * alt[0];
        sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    }
    for (double& v : s) v /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    const double vw[16] = {s[0], u[0], -f[0], 0.0, s[1], u[1], -f[1], 0.0, s[2], u[2], -f[2], 0.0,
                           -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]), -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]),
                           (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]), 1.0};
    memcpy(view, vw, sizeof(vw));
    for (int i = 0; i < 16; ++i) proj[i] = 0.0;
    proj[0] = t / aspect;
    proj[5] = t;
    proj[11] = -1.0;
    if (homogeneous_depth) {
        proj[10] = (fa + n) / (n - fa);
        proj[14] = 2.0 * fa * n / (n - fa);
    } else {
        proj[10] = fa / (n - fa);
        proj[14] = fa * n / (n - fa);
    }
    mat_mul(proj, view, vp);
    mat_inverse(vp, inv);
}

} // namespace

extern "C" {

int64_t rgi_scene_generate(int kind, uint32_t seed, uint32_t size, rgi_mesh* mesh, uint32_t capacity)
{
    Builder b;
    if (kind == RGI_SCENE_CORNELL) make_cornell(b, seed);
    else if (kind == RGI_SCENE_ATRIUM) make_atrium(b, seed, size);
    else if (kind == RGI_SCENE_BOXFIELD) make_boxfield(b, seed, size);
    else return -1;
    uint32_t n = (uint32_t)b.mat.size();
    if (mesh == nullptr || capacity < n) return n;
    memcpy(mesh->positions, b.pos.data(), sizeof(float) * 9 * n);
    memcpy(mesh->albedo, b.alb.data(), sizeof(float) * 3 * n);
    memcpy(mesh->emission, b.emi.data(), sizeof(float) * 3 * n);
    memcpy(mesh->material, b.mat.data(), sizeof(uint32_t) * n);
    uint32_t mc = (uint32_t)(b.materials.size() / 4);
    if (mesh->materials) memcpy(mesh->materials, b.materials.data(), sizeof(float) * 4 * std::min(mc, 16u));
    mesh->material_count = mc;
    mesh->triangle_count = n;
    return n;
}

void rgi_mesh_bounds(const rgi_mesh* mesh, float* bmin, float* bmax, float* center, float* radius)
{
    Box box;
    for (uint32_t i = 0; i < mesh->triangle_count * 3; ++i)
        box.add(Vec{mesh->positions[i * 3], mesh->positions[i * 3 + 1], mesh->positions[i * 3 + 2]});
    if (bmin) { bmin[0] = box.lo.x; bmin[1] = box.lo.y; bmin[2] = box.lo.z; }
    if (bmax) { bmax[0] = box.hi.x; bmax[1] = box.hi.y; bmax[2] = box.hi.z; }
    Vec c = (box.lo + box.hi) * 0.5f;
    if (center) { center[0] = c.x; center[1] = c.y; center[2] = c.z; }
    Vec d = box.hi - box.lo;
    if (radius) *radius = std::max(0.5f * sqrtf(dot(d, d)), 1e-3f);
}

uint64_t rgi_bvh_capacity(uint32_t triangle_count)
{
    uint64_t nodes = triangle_count > 0 ? 2ull * triangle_count - 1ull : 0ull;
    return (2ull * nodes + 4ull * triangle_count) * 4ull;
}

int64_t rgi_bvh_build(const rgi_mesh* mesh, float* out, uint32_t* out_depth)
{
    uint32_t n = mesh->triangle_count;
    if (n == 0) {
        if (out_depth) *out_depth = 0;
        return 0;
    }
    if (n >= (1u << 24) / 2) return -1; // indices must stay exact as floats
    std::vector<Box> tri_box(n);
    std::vector<Vec> centroid(n);
    for (uint32_t t = 0; t < n; ++t) {
        const float* p = mesh->positions + (size_t)t * 9;
        Box bb;
        for (int c = 0; c < 3; ++c) bb.add(Vec{p[c * 3], p[c * 3 + 1], p[c * 3 + 2]});
        tri_box[t] = bb;
        centroid[t] = (bb.lo + bb.hi) * 0.5f;
    }
    std::vector<uint32_t> order(n);
    for (uint32_t t = 0; t < n; ++t) order[t] = t;
    BvhWriter w{tri_box, centroid, order, {}, 0};
    w.nodes.reserve((size_t)(2 * n) * 8);
    w.build(0, n, 1);
    uint32_t nodes = (uint32_t)(w.nodes.size() / 8);
    memcpy(out, w.nodes.data(), sizeof(float) * w.nodes.size());
    float* rec = out + (size_t)nodes * 8;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t t = order[i];
        const float* p = mesh->positions + (size_t)t * 9;
        float* r = rec + (size_t)i * 16;
        for (int c = 0; c < 3; ++c) {
            r[c * 4 + 0] = p[c * 3 + 0];
            r[c * 4 + 1] = p[c * 3 + 1];
            r[c * 4 + 2] = p[c * 3 + 2];
        }
        r[3] = (float)mesh->material[t];
        r[7] = mesh->emission[t * 3 + 0];
        r[11] = mesh->emission[t * 3 + 1];
        r[15] = mesh->emission[t * 3 + 2];
        r[12] = mesh->albedo[t * 3 + 0];
        r[13] = mesh->albedo[t * 3 + 1];
        r[14] = mesh->albedo[t * 3 + 2];
    }
    if (out_depth) *out_depth = w.depth;
    return nodes;
}

int64_t rgi_lights_build(const float* bvh, uint32_t node_count, uint32_t triangle_count, int32_t extra_count,
                         float extra_intensity, uint32_t seed, const float* bmin, const float* bmax,
                         float intensity_scale, float* out, uint32_t capacity)
{
    std::vector<float> data;
    std::vector<float> power;
    const float* rec = bvh + (size_t)node_count * 8;
    for (uint32_t t = 0; t < triangle_count; ++t) {
        const float* r = rec + (size_t)t * 16;
        float e[3] = {r[7], r[11], r[15]};
        if (!(e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f)) continue;
        Vec a{r[0], r[1], r[2]}, b{r[4], r[5], r[6]}, c{r[8], r[9], r[10]};
        Vec cr = cross(b - a, c - a);
        float area = 0.5f * sqrtf(dot(cr, cr));
        if (area <= 0.0f) continue;
        data.insert(data.end(), {a.x, a.y, a.z, 0.0f, b.x, b.y, b.z, (float)t, c.x, c.y, c.z, 0.0f, e[0], e[1], e[2], 0.0f});
        power.push_back(3.14159265f * (float)luminance(e) * area);
    }
    Rng rng{seed * 2654435761u + 1u};
    for (int i = 0; i < extra_count; ++i) {
        float p[3];
        const float lo[3] = {bmin[0], bmin[1], bmin[2]}, hi[3] = {bmax[0], bmax[1], bmax[2]};
        for (int a = 0; a < 3; ++a) p[a] = lo[a] + (hi[a] - lo[a]) * (0.1f + 0.8f * rng.next());
        float rgb[3];
        hue_rgb(rng.next(), rgb);
        for (float& ch : rgb) ch *= extra_intensity * intensity_scale;
        data.insert(data.end(), {p[0], p[1], p[2], 1.0f, 0.0f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f, 0.0f, 0.0f, rgb[0], rgb[1], rgb[2], 0.0f});
        power.push_back(4.0f * 3.14159265f * (float)luminance(rgb));
    }
    uint32_t count = (uint32_t)power.size();
    double total = 0.0;
    for (float v : power) total += v;
    double cumulative = 0.0;
    for (uint32_t i = 0; i < count; ++i) {
        double pmf = total > 0.0 ? power[i] / total : 1.0 / (double)count;
        cumulative += pmf;
        data[(size_t)i * 16 + 11] = i + 1 == count ? 1.0f : (float)cumulative;
        data[(size_t)i * 16 + 15] = (float)pmf;
    }
    if (out != nullptr && capacity >= count) memcpy(out, data.data(), sizeof(float) * data.size());
    return count;
}

void rgi_camera_matrices(const float* eye3, const float* target3, float fovy_degrees, float aspect, float near_plane,
                         float far_plane, int homogeneous_depth, float* view_out, float* proj_out, float* vp_out,
                         float* inv_out)
{
    const double up[3] = {0.0, 1.0, 0.0};
    double view[16], proj[16], vp[16], inv[16];
    look_at_projection(eye3, target3, up, 1.0 / tan(0.5 * (double)fovy_degrees * 3.14159265358979323846 / 180.0),
                       aspect, near_plane, far_plane, homogeneous_depth, view, proj, vp, inv);
    for (int i = 0; i < 16; ++i) {
        if (view_out) view_out[i] = (float)view[i];
        if (proj_out) proj_out[i] = (float)proj[i];
        if (vp_out) vp_out[i] = (float)vp[i];
        if (inv_out) inv_out[i] = (float)inv[i];
    }
}

int rgi_shadow_setup(const float* light_position, float scene_radius, int32_t face_size, int homogeneous_depth,
                     int origin_bottom_left, rgi_shadow* out)
{
    if (!light_position || !out) return RGI_E_INVALID;
    if (face_size != 256 && face_size != 512 && face_size != 1024 && face_size != 2048) return RGI_E_INVALID;
    if (!(scene_radius > 0.0f)) return RGI_E_INVALID;
    memset(out, 0, sizeof(*out));
    // renderer.cpp:529-564: kShadowFaceBorder = 4 texels on each side.
    const float tan_half = (float)face_size / (float)(face_size - 2 * 4);
    const float far_plane = 2.5f * scene_radius;
    const float near_plane = 0.01f * scene_radius;
    static const double directions[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int face = 0; face < 6; ++face) {
        const double up[3] = {0.0, face == 2 || face == 3 ? 0.0 : 1.0, face == 2 || face == 3 ? 1.0 : 0.0};
        float target[3];
        for (int i = 0; i < 3; ++i) target[i] = (float)((double)light_position[i] + directions[face][i]);
        double view[16], proj[16], vp[16], inv[16];
        look_at_projection(light_position, target, up, 1.0 / (double)tan_half, 1.0, near_plane, far_plane,
                           homogeneous_depth, view, proj, vp, inv);
        for (int i = 0; i < 16; ++i) {
            out->view_proj[face * 16 + i] = (float)vp[i];
            out->inv_view_proj[face * 16 + i] = (float)inv[i];
        }
    }
    out->info[0] = far_plane;
    out->info[1] = tan_half;
    out->info[2] = 1.0f / (float)face_size;
    out->info[3] = origin_bottom_left ? 1.0f : 0.0f;
    for (int i = 0; i < 3; ++i) out->light_position[i] = light_position[i];
    out->light_position[3] = 1.0f;
    out->near_plane = near_plane;
    out->face_size = face_size;
    out->homogeneous_depth = homogeneous_depth ? 1 : 0;
    return RGI_OK;
}

} // extern "C"
