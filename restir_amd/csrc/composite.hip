// The composite pass on gfx950 (fs_composite.sc): per full-resolution pixel,
// the main point light through the BRDF and a 4 x 4 shadow-atlas compare,
// the ReSTIR DI light, the GI irradiance upsampled from the GI planes over
// same-surface texels and re-modulated by the albedo, indirect specular and
// emission; the sky where the depth is the far plane. One thread per pixel,
// 16 x 16 pixels per block. A resolved diffuse plane, when given, replaces
// the DI diffuse light and the re-modulated GI irradiance together.
#include "batch.h"
#include "internal.h"
#include "shading.h"

namespace rgi {
namespace {

RGI_DEV F4 h4_or_zero(const uint16_t* plane, uint32_t index)
{
    return plane ? load_h4(plane, index) : F4{0.0f, 0.0f, 0.0f, 0.0f};
}

RGI_DEV F3 gbuffer_normal(const uint32_t* plane, uint32_t index)
{
    F4 n = unorm8(plane[index]);
    return normalize(F3{n.x * 2.0f - 1.0f, n.y * 2.0f - 1.0f, n.z * 2.0f - 1.0f});
}

// upsampleIndirect (fs_composite.sc:48-80): the four GI texels around the
// pixel, bilinear weights plus 0.01, kept only on the pixel's surface. The GI
// and indirect specular planes are GI size (u_screenInfo), the G-buffer the
// window's size.
RGI_DEV void upsample_indirect(const rgi_frame& f, const rgi_gbuffer& g, const rgi_composite_inputs& in, F2 uv,
                               F3 position, F3 n, F3& diffuse, F3& specular)
{
    float tcx = uv.x * (float)f.gi_width - 0.5f, tcy = uv.y * (float)f.gi_height - 0.5f;
    float bx = floorf(tcx), by = floorf(tcy);
    float blx = tcx - bx, bly = tcy - by;
    F2 gi_texel = F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height};
    F3 sum = splat3(0.0f), specular_sum = splat3(0.0f);
    float weight_sum = 0.0f;
    for (int k = 0; k < 4; ++k) {
        float ox = (float)(k & 1), oy = (float)(k >> 1);
        F2 tuv = F2{((bx + ox) + 0.5f) * gi_texel.x, ((by + oy) + 0.5f) * gi_texel.y};
        uint32_t gt = texel_index(tuv, g.width, g.height);
        float tap_depth = g.depth[gt];
        if (tap_depth >= kSkyDepth) continue;
        F3 tap_position = reconstruct_position(f, tuv, tap_depth);
        F3 tap_normal = gbuffer_normal(g.normal, gt);
        float wx = mixf_(1.0f - blx, blx, ox), wy = mixf_(1.0f - bly, bly, oy);
        float plane_step = 0.03f * f.scene_scale >= fabsf(dot(tap_position - position, n)) ? 1.0f : 0.0f;
        float normal_step = dot(tap_normal, n) >= 0.9f ? 1.0f : 0.0f;
        float weight = ((wx * wy + 0.01f) * plane_step) * normal_step;
        uint32_t it = texel_index(tuv, f.gi_width, f.gi_height);
        sum = sum + xyz(load_h4(in.gi, it)) * weight;
        specular_sum = specular_sum + xyz(h4_or_zero(in.specular, it)) * weight;
        weight_sum += weight;
    }
    uint32_t own = texel_index(uv, f.gi_width, f.gi_height);
    diffuse = weight_sum > 0.0f ? sum / weight_sum : xyz(load_h4(in.gi, own));
    specular = weight_sum > 0.0f ? specular_sum / weight_sum : xyz(h4_or_zero(in.specular, own));
}

// fs_composite.sc:82-143, for the 16 x 16-pixel block (bx, by).
RGI_DEV void composite_body(const rgi_frame& f, const rgi_gbuffer& g, const rgi_composite_inputs& in, F2 texel,
                            uint16_t* out, uint32_t bx, uint32_t by)
{
    int x = (int)(bx * 16u + (threadIdx.x & 15u)), y = (int)(by * 16u + (threadIdx.x >> 4));
    if (x >= g.width || y >= g.height) return;
    uint32_t o = (uint32_t)y * (uint32_t)g.width + (uint32_t)x;
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    uint32_t t = texel_index(uv, g.width, g.height);
    F3 camera = F3{f.camera_position[0], f.camera_position[1], f.camera_position[2]};
    float depth = g.depth[t];
    if (depth >= kSkyDepth) {
        F3 far_position = reconstruct_position(f, uv, 1.0f);
        F3 direction = normalize(far_position - camera);
        F3 sky = sky_radiance(f, direction);
        store_h4(out, o, sky_on(f) ? F4{sky.x, sky.y, sky.z, 1.0f} : F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    F4 albedo_roughness = unorm8(g.albedo[t]);
    F4 normal_metallic = unorm8(g.normal[t]);
    F3 albedo = xyz(albedo_roughness);
    float roughness = maxf_(albedo_roughness.w, 0.05f);
    float metallic = normal_metallic.w;
    F3 surface_normal = normalize(F3{normal_metallic.x * 2.0f - 1.0f, normal_metallic.y * 2.0f - 1.0f,
                                     normal_metallic.z * 2.0f - 1.0f});
    F3 n = gbuffer_normal(g.shading_normal, t);
    F4 emissive_occlusion = load_h4(g.emission, t);
    F3 position = reconstruct_position(f, uv, depth);
    F3 view_direction = normalize(camera - position);
    F3 to_light = F3{f.light_position[0], f.light_position[1], f.light_position[2]} - position;
    float distance_squared = maxf_(dot(to_light, to_light), (0.01f * f.scene_scale) * f.scene_scale);
    F3 light_direction = to_light * rsqrt_(distance_squared);

    F3 diffuse, specular;
    brdf(n, view_direction, light_direction, albedo, roughness, metallic, diffuse, specular);
    float visibility = 0.0f;
    if (dot(n, light_direction) > 0.0f) visibility = light_visibility_soft(f, in.shadow_map, position, surface_normal);
    F3 light_color = F3{f.light_color[0], f.light_color[1], f.light_color[2]};
    F3 diffuse_light = in.resolved_diffuse ? xyz(load_h4(in.resolved_diffuse, t)) : xyz(h4_or_zero(in.di_diffuse, t));
    F3 direct = (((diffuse + specular) * (light_color / distance_squared)) * visibility + diffuse_light) +
                xyz(h4_or_zero(in.di_specular, t));

    F3 irradiance, specular_radiance;
    upsample_indirect(f, g, in, uv, position, surface_normal, irradiance, specular_radiance);
    F3 modulate = in.options[2] > 0.5f ? (albedo * (1.0f - metallic)) * emissive_occlusion.w : splat3(1.0f);
    F3 indirect = in.resolved_diffuse ? splat3(0.0f) : irradiance * modulate;
    F3 f0 = mix3(splat3(0.04f), albedo, metallic);
    indirect = indirect + ((specular_radiance * specular_albedo(f0, roughness, maxf_(dot(n, view_direction), 0.001f))) *
                           emissive_occlusion.w) *
                              in.specular_weight;
    F3 color = (direct * in.options[0] + indirect * in.options[1]) + xyz(emissive_occlusion) * in.options[3];
    store_h4(out, o, F4{color.x, color.y, color.z, 1.0f});
}

__global__ void __launch_bounds__(256) k_composite(rgi_frame f, rgi_gbuffer g, rgi_composite_inputs in, F2 texel,
                                                   uint16_t* out)
{
    composite_body(f, g, in, texel, out, blockIdx.x, blockIdx.y);
}

// The batched form: workgroups [first[v], first[v + 1]) run view v's blocks.
__global__ void __launch_bounds__(256) k_composite_batch(const BatchView* __restrict__ views,
                                                         const uint32_t* __restrict__ first, uint32_t count)
{
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    const uint32_t gx = ((uint32_t)v.g.width + 15u) / 16u;
    composite_body(v.f, v.g, v.ci, v.texel, v.hdr, b % gx, b / gx);
}

} // namespace

int composite_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, hipStream_t stream)
{
    if (blocks == 0u) return RGI_OK;
    hipLaunchKernelGGL(k_composite_batch, dim3(blocks), dim3(256), 0, stream, views, first, count);
    return check_launch("composite (batch)");
}

int composite(const rgi_frame& f, const rgi_gbuffer& g, const rgi_composite_inputs& in, uint16_t* out,
              hipStream_t stream)
{
    dim3 blocks((uint32_t)(g.width + 15) / 16u, (uint32_t)(g.height + 15) / 16u);
    F2 texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
    hipLaunchKernelGGL(k_composite, blocks, dim3(256), 0, stream, f, g, in, texel, out);
    return check_launch("composite");
}

} // namespace rgi
