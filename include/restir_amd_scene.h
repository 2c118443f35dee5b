/*
 * restir_amd scene data layer — host C ABI.
 *
 * The producers of the hot path's scene inputs, run on the host once per
 * scene: seeded procedural scenes, the binned-SAH BVH in the reference's GPU
 * layout, the light list with its power CDF, and camera matrices. Everything
 * is plain arrays owned by the caller; functions that fill an array take its
 * capacity and return how many elements they need (call with NULL first).
 */
#ifndef RESTIR_AMD_SCENE_H
#define RESTIR_AMD_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Procedural scenes. */
#define RGI_SCENE_CORNELL 0  /* open-front Cornell box: 5 walls, 2 boxes, emissive ceiling quad (36 tris) */
#define RGI_SCENE_ATRIUM 1   /* three-storey courtyard of columns, arches and slabs, 32 emissive quads */
#define RGI_SCENE_BOXFIELD 2 /* random axis-aligned boxes on a floor, some emissive */

/* A triangle soup: per triangle 9 floats of positions (v0, v1, v2), 3 floats
 * of diffuse albedo, 3 floats of emitted radiance and a material index. */
typedef struct rgi_mesh {
    uint32_t triangle_count;
    float* positions;
    float* albedo;
    float* emission;
    uint32_t* material;
    uint32_t material_count;
    float* materials;      /* 4 floats per material: roughness, metallic, 0, 0 */
} rgi_mesh;

/* Fills `mesh` (arrays of at least `capacity` triangles and 16 materials) with
 * scene `kind`; `size` is the atrium's triangle budget or the box field's box
 * count (ignored for the Cornell box). Returns the triangle count (nothing is
 * written when capacity is too small), or a negative code. */
int64_t rgi_scene_generate(int kind, uint32_t seed, uint32_t size, rgi_mesh* mesh, uint32_t capacity);

/* Bounds of a mesh and the framing describeScene derives from them
 * (src/scene.cpp:810-816): centre and radius = half the diagonal. */
void rgi_mesh_bounds(const rgi_mesh* mesh, float* bounds_min3, float* bounds_max3, float* center3,
                     float* radius);

/* buildBvh (src/bvh.cpp:253-393): binned SAH with 12 bins per axis, leaves of
 * at most 4 triangles when the SAH allows, always split above 16, depth at
 * most 40; nodes written depth first with skip links, then the triangle
 * records in leaf order, in the layout of src/bvh.h:9-22. `out` must hold
 * rgi_bvh_capacity(triangle_count) floats. Returns the node count, or a
 * negative code. `out_depth` (may be NULL) receives the deepest leaf's depth
 * (root = 1). */
uint64_t rgi_bvh_capacity(uint32_t triangle_count);
int64_t rgi_bvh_build(const rgi_mesh* mesh, float* out, uint32_t* out_depth);

/* buildLightList (src/lights.cpp:86-165): every emissive triangle of the BVH
 * (in BVH order) with power pi * luminance * area, then `extra_count` point
 * lights scattered through the middle 80% of the bounds in saturated colours
 * from `seed`, with intensity extra_intensity * intensity_scale and power
 * 4 pi * luminance; CDF and pmf in the .w of rows 2 and 3 (lights.h:7-14).
 * `out` holds 16 floats per light; returns the light count (nothing written
 * when capacity is too small). */
int64_t rgi_lights_build(const float* bvh, uint32_t node_count, uint32_t triangle_count,
                         int32_t extra_count, float extra_intensity, uint32_t seed,
                         const float* bounds_min3, const float* bounds_max3, float intensity_scale,
                         float* out, uint32_t capacity);

/* Right-handed look-at view and perspective projection (fovy in degrees),
 * depth range [0, 1] or, with homogeneous_depth, [-1, 1]; outputs are
 * column-major: view, projection, view_proj = projection * view, and its
 * inverse (computed in double precision). Any output may be NULL. */
void rgi_camera_matrices(const float* eye3, const float* target3, float fovy_degrees, float aspect,
                         float near_plane, float far_plane, int homogeneous_depth, float* view,
                         float* projection, float* view_proj, float* inv_view_proj);

#ifdef __cplusplus
}
#endif

#endif /* RESTIR_AMD_SCENE_H */
