// Ray batches through the BVH (rgi_trace) and the primary-ray G-buffer
// producer (rgi_gbuffer_primary).
//
// rgi_trace has two launch forms, bit-exact with each other:
//   - one ray per lane;
//   - persistent waves (RGI_OPT_TRACE_PERSISTENT): each wave claims chunks of
//     rays and answers them with every lane, a lane whose ray ends taking the
//     next ray of the wave's chunk at once, so no lane idles behind a long
//     traversal while rays remain. Chunks are dealt to 8 counters (workgroup
//     b uses counter b mod 8), since one counter serves only about one claim
//     per 13 ns.
// Closest-hit rays keep their 8-entry ring stack in LDS, laid out slot-major
// so that the 64 lanes of a wave hit 64 different banks.
#include "bvh_device.h"
#include "internal.h"

namespace rgi {
namespace {

constexpr int kTraceBlock = 256;

__global__ void __launch_bounds__(kTraceBlock) k_trace(rgi_scene s, const float4* rays, uint32_t count, int32_t* hit,
                                                       float* dist, int any_hit, int watertight)
{
    __shared__ int ring[kStackRing * kTraceBlock];
    SceneView v = scene_view(s);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 a = rays[2u * i], b = rays[2u * i + 1u];
    F3 o = F3{a.x, a.y, a.z}, d = F3{b.x, b.y, b.z};
    float t = a.w;
    int h;
    if (any_hit) {
        h = trace_any(v, o, d, a.w, t, watertight != 0);
    } else {
        h = trace_closest(v, o, d, a.w, t, ring + threadIdx.x, kTraceBlock, watertight != 0);
    }
    hit[i] = h;
    dist[i] = t;
}

// Persistent form (trace_persistent): workgroups claim chunks of rays from
// kTraceGroups counters until the rays run out.
template <class Tri, bool kAny>
__global__ void __launch_bounds__(kTraceBlock) k_trace_persistent(rgi_scene s, const float4* rays, uint32_t count,
                                                                  int32_t* hit, float* dist, uint32_t* counters)
{
    __shared__ int ring[kStackRing * kTraceBlock];
    trace_persistent<Tri, kAny>(scene_view(s), rays, count, hit, dist, counters, ring + threadIdx.x, kTraceBlock,
                                kTraceChunk);
}

RGI_DEV uint32_t unorm_byte(float v)
{
    float q = floorf(clampf_(v, 0.0f, 1.0f) * 255.0f + 0.5f);
    return (uint32_t)to_int(q);
}

__global__ void __launch_bounds__(256) k_gbuffer(rgi_frame f, rgi_scene s, const float4* materials,
                                                 uint32_t material_count, F2 texel, float* depth, uint32_t* albedo,
                                                 uint32_t* normal, uint32_t* shading_normal, uint16_t* emission)
{
    __shared__ int ring[kStackRing * 256];
    SceneView v = scene_view(s);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = (uint32_t)f.width * (uint32_t)f.height;
    if (i >= n) return;
    int x = (int)(i % (uint32_t)f.width), y = (int)(i / (uint32_t)f.width);
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    F3 clip = F3{uv.x * 2.0f - 1.0f, to_clip_y(f, uv.y * 2.0f - 1.0f), 1.0f};
    F3 far_point = clip_to_world(f.inv_view_proj, clip);
    F3 eye = F3{f.camera_position[0], f.camera_position[1], f.camera_position[2]};
    F3 dir = normalize(far_point - eye);
    float dist;
    int tri = trace_closest(v, eye, dir, 1e30f, dist, ring + threadIdx.x, 256);
    float d = 1.0f;
    if (tri >= 0) {
        F3 p = eye + dir * dist;
        F4 c = mul(f.view_proj, F4{p.x, p.y, p.z, 1.0f});
        float z = c.z / c.w;
        if (f.homogeneous_depth) z = z * 0.5f + 0.5f;
        d = z;
        if (!(z >= 0.0f && z < kSkyDepth)) tri = -1;
    }
    if (tri < 0) {
        depth[i] = 1.0f;
        albedo[i] = 0u;
        normal[i] = 0u;
        shading_normal[i] = 0u;
        store_h4(emission, i, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    depth[i] = d;
    F3 v0 = vertex(v, tri, 0);
    F3 nn = normalize(cross(vertex(v, tri, 1) - v0, vertex(v, tri, 2) - v0));
    if (dot(nn, dir) > 0.0f) nn = -nn;
    uint32_t material = (uint32_t)to_int(sv(v, v.tri_offset + (uint32_t)tri * 4u).w);
    float roughness = 1.0f, metallic = 0.0f;
    if (material < material_count) {
        float4 m = materials[material];
        roughness = m.x;
        metallic = m.y;
    }
    F3 a = tri_albedo(v, tri);
    albedo[i] = unorm_byte(a.x) | (unorm_byte(a.y) << 8) | (unorm_byte(a.z) << 16) | (unorm_byte(roughness) << 24);
    uint32_t nw = unorm_byte(nn.x * 0.5f + 0.5f) | (unorm_byte(nn.y * 0.5f + 0.5f) << 8) |
                  (unorm_byte(nn.z * 0.5f + 0.5f) << 16);
    normal[i] = nw | (unorm_byte(metallic) << 24);
    shading_normal[i] = nw;
    F3 e = tri_emission(v, tri);
    // Emissive target: rgb = emitted radiance, a = ambient occlusion, 1
    // without an occlusion map (fs_gbuffer.sc:81-93).
    store_h4(emission, i, F4{e.x, e.y, e.z, 1.0f});
}

// The shadow pass: texel (X, Y) of the 3F x 2F atlas belongs to face
// column + 3 row, with column = X / F and row = Y / F (counted from the top,
// from the bottom when the origin is bottom-left), which inverts
// shadowAtlasUv (shadow.sh:52-69) at the texel centre. Its ray runs from the
// light through the texel's point on the far plane (NDC z = 1), between the
// near and far planes.
__global__ void __launch_bounds__(256) k_shadow_atlas(rgi_shadow sh, rgi_scene s, float* atlas)
{
    __shared__ int ring[kStackRing * 256];
    SceneView v = scene_view(s);
    int F = sh.face_size;
    int X = (int)(blockIdx.x * 16u + (threadIdx.x & 15u)), Y = (int)(blockIdx.y * 16u + (threadIdx.x >> 4));
    if (X >= 3 * F || Y >= 2 * F) return;
    bool obl = sh.info[3] > 0.5f;
    int column = X / F, texture_row = Y / F;
    int face = column + 3 * (obl ? 1 - texture_row : texture_row);
    float lx = ((float)(X - column * F) + 0.5f) / (float)F;
    float ly = ((float)(Y - texture_row * F) + 0.5f) / (float)F;
    F3 ndc = F3{lx * 2.0f - 1.0f, obl ? ly * 2.0f - 1.0f : 1.0f - ly * 2.0f, 1.0f};
    F3 far_point = clip_to_world(sh.inv_view_proj + face * 16, ndc);
    F3 light = F3{sh.light_position[0], sh.light_position[1], sh.light_position[2]};
    F3 to_far = far_point - light;
    float far_distance = length(to_far);
    F3 dir = to_far * (1.0f / far_distance);
    float near_distance = far_distance * (sh.near_plane / sh.info[0]);
    F3 origin = light + dir * near_distance;
    float dist;
    int tri = trace_closest(v, origin, dir, far_distance - near_distance, dist, ring + threadIdx.x, 256);
    float value = 1.0f;
    if (tri >= 0) value = length((origin + dir * dist) - light) / sh.info[0];
    atlas[(uint32_t)Y * (uint32_t)(3 * F) + (uint32_t)X] = value;
}

} // namespace

__global__ void k_zero_counters(uint32_t* counters)
{
    if (threadIdx.x < kTraceGroups) counters[threadIdx.x * kTraceLine] = 0u;
}

int shadow_atlas(const rgi_shadow& sh, const rgi_scene& s, float* atlas, hipStream_t stream)
{
    uint32_t F = (uint32_t)sh.face_size;
    hipLaunchKernelGGL(k_shadow_atlas, dim3(3u * F / 16u, 2u * F / 16u), dim3(256), 0, stream, sh, s, atlas);
    return check_launch("shadow atlas");
}

int trace_rays(const rgi_scene& s, const float* rays, uint32_t count, int32_t* hit, float* dist, int any_hit,
               const Tuning& tuning, uint32_t* counter, hipStream_t stream)
{
    if (count == 0) return RGI_OK;
    const float4* r = reinterpret_cast<const float4*>(rays);
    if (tuning.trace_persistent && counter) {
        // As many workgroups as run at once (at least one per group).
        auto kernel = any_hit ? (tuning.watertight ? k_trace_persistent<Watertight, true>
                                                   : k_trace_persistent<MollerTrumbore, true>)
                              : (tuning.watertight ? k_trace_persistent<Watertight, false>
                                                   : k_trace_persistent<MollerTrumbore, false>);
        int device = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&device) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTraceBlock, 0) != hipSuccess ||
            cus <= 0 || per_cu <= 0)
            return fail(RGI_E_DEVICE, "trace: occupancy query failed");
        uint32_t blocks = (uint32_t)cus * (uint32_t)per_cu;
        if (blocks < kTraceGroups) blocks = kTraceGroups;
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, stream, counter);
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kTraceBlock), 0, stream, s, r, count, hit, dist, counter);
    } else {
        uint32_t blocks = (count + kTraceBlock - 1u) / kTraceBlock;
        hipLaunchKernelGGL(k_trace, dim3(blocks), dim3(kTraceBlock), 0, stream, s, r, count, hit, dist, any_hit,
                           tuning.watertight);
    }
    return check_launch("trace");
}

int gbuffer_primary(const rgi_frame& f, const rgi_scene& s, const float* materials, uint32_t material_count,
                    float* depth, uint32_t* albedo, uint32_t* normal, uint32_t* shading_normal, uint16_t* emission,
                    hipStream_t stream)
{
    uint32_t n = (uint32_t)f.width * (uint32_t)f.height;
    F2 texel = F2{1.0f / (float)f.width, 1.0f / (float)f.height};
    hipLaunchKernelGGL(k_gbuffer, dim3((n + 255u) / 256u), dim3(256), 0, stream, f, s,
                       reinterpret_cast<const float4*>(materials), material_count, texel, depth, albedo, normal,
                       shading_normal, emission);
    return check_launch("gbuffer primary");
}

} // namespace rgi
