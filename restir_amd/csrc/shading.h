// Shading shared by the DI, GI and composite passes on gfx950: the BRDF
// (brdf.sh), the point light's shadow-atlas lookup (shadow.sh) and the sky
// (sky.sh). Plain float arithmetic in the shaders' order, as in device.h.
#pragma once

#include "device.h"
#include "rgi_math.h"

namespace rgi {

// ---------------------------------------------------------------- BRDF

// GGX normal distribution (distributionGGX, brdf.sh:14-21).
RGI_DEV float ggx_d(F3 n, F3 h, float roughness)
{
    float alpha = roughness * roughness;
    float a2 = alpha * alpha;
    float nh = maxf_(dot(n, h), 0.0f);
    float den = (nh * nh) * (a2 - 1.0f) + 1.0f;
    return a2 / maxf_((kPi * den) * den, 0.0001f);
}

// Smith-Schlick G1 (geometrySchlickGGX, brdf.sh:23-27).
RGI_DEV float smith_g1(float nd, float roughness)
{
    float k = ((roughness + 1.0f) * (roughness + 1.0f)) / 8.0f;
    return nd / maxf_(nd * (1.0f - k) + k, 0.0001f);
}

// fresnelSchlick (brdf.sh:29-39).
RGI_DEV F3 fresnel_schlick(F3 h, F3 v, F3 f0)
{
    float c = clampf_(dot(h, v), 0.0f, 1.0f);
    float x = 1.0f - c;
    float w = ((x * x) * (x * x)) * x;
    return f0 + (splat3(1.0f) - f0) * w;
}

// evaluateBrdf (brdf.sh:41-66): Lambert + GGX with Schlick Fresnel, both
// times n.l.
RGI_DEV void brdf(F3 n, F3 v, F3 l, F3 albedo, float roughness_in, float metallic, F3& diffuse, F3& specular)
{
    float nl = maxf_(dot(n, l), 0.0f);
    float nv = maxf_(dot(n, v), 0.0f);
    F3 h = normalize(v + l);
    F3 f0 = mix3(splat3(0.04f), albedo, metallic);
    F3 fr = fresnel_schlick(h, v, f0);
    float roughness = maxf_(roughness_in, 0.05f);
    float d = ggx_d(n, h, roughness);
    float g = smith_g1(nl, roughness) * smith_g1(nv, roughness);
    specular = ((fr * (d * g)) / maxf_((4.0f * nl) * nv, 0.001f)) * nl;
    diffuse = (((splat3(1.0f) - fr) * (1.0f - metallic)) * albedo / kPi) * nl;
}

// specularAlbedo (brdf.sh:68-76): Karis' analytic fit of the split-sum
// specular albedo.
RGI_DEV F3 specular_albedo(F3 f0, float roughness, float nv)
{
    float rx = -1.0f * roughness + 1.0f;
    float ry = -0.0275f * roughness + 0.0425f;
    float rz = -0.572f * roughness + 1.04f;
    float rw = 0.022f * roughness + -0.04f;
    float a004 = minf_(rx * rx, rgi_exp2f(-9.28f * nv)) * rx + ry;
    float sx = -1.04f * a004 + rz;
    float sy = 1.04f * a004 + rw;
    return f0 * sx + splat3(sy);
}

// ---------------------------------------------------------- shadow atlas

// shadowFace (shadow.sh:31-45).
RGI_DEV int shadow_face(F3 d)
{
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    if (ax >= ay && ax >= az) return d.x >= 0.0f ? 0 : 1;
    if (ay >= az) return d.y >= 0.0f ? 2 : 3;
    return d.z >= 0.0f ? 4 : 5;
}

// shadowLookup (shadow.sh:72-97) with shadowAtlasUv (48-69): where a surface
// point falls in the atlas and the biased distance it is compared against;
// false when shadows are off or no atlas is bound.
RGI_DEV bool shadow_lookup(const rgi_frame& f, const float* atlas, F3 position, F3 n, float filter_radius, F2& uv,
                           float& reference)
{
    uv = F2{0.0f, 0.0f};
    reference = 0.0f;
    if (!f.shadows_enabled || atlas == nullptr) return false;
    F3 light = F3{f.light_position[0], f.light_position[1], f.light_position[2]};
    float light_distance = length(light - position);
    float texel_size = ((2.0f * light_distance) * f.shadow_info[1]) * f.shadow_info[2];
    F3 pos = position + n * (texel_size * 1.5f);
    F3 from_light = pos - light;
    int face = shadow_face(from_light);
    F4 clip = mul(f.shadow_view_proj + face * 16, F4{pos.x, pos.y, pos.z, 1.0f});
    float ndx = clip.x / clip.w, ndy = clip.y / clip.w;
    bool obl = f.shadow_info[3] > 0.5f;
    float lx = ndx * 0.5f + 0.5f;
    float ly = obl ? ndy * 0.5f + 0.5f : 0.5f - ndy * 0.5f;
    float row = face >= 3 ? 1.0f : 0.0f;
    float column = (float)face - 3.0f * row;
    if (obl) row = 1.0f - row;
    uv = F2{(column + lx) * (1.0f / 3.0f), (row + ly) * 0.5f};
    float cosine = clampf_(dot(n, -normalize(from_light)), 0.0f, 1.0f);
    float slope = minf_(sqrtf(1.0f - cosine * cosine) / maxf_(cosine, 0.0001f), 8.0f);
    float bias = texel_size * (1.0f + filter_radius * slope);
    reference = (length(from_light) - bias) / f.shadow_info[0];
    return true;
}

// lightVisibilitySoft (shadow.sh:113-134): a 4 x 4 grid of compares.
RGI_DEV float light_visibility_soft(const rgi_frame& f, const float* atlas, F3 position, F3 n)
{
    F2 uv;
    float reference;
    if (!shadow_lookup(f, atlas, position, n, 2.0f, uv, reference)) return 1.0f;
    F2 atlas_texel = F2{div3(f.shadow_info[2]), f.shadow_info[2] * 0.5f};
    float lit = 0.0f;
    for (int y = 0; y < 4; ++y) {
        for (int x = 0; x < 4; ++x) {
            F2 offset = F2{((float)x - 1.5f) * atlas_texel.x, ((float)y - 1.5f) * atlas_texel.y};
            F2 tap = F2{uv.x + offset.x, uv.y + offset.y};
            lit += reference <= atlas[texel_index(tap, f.shadow_width, f.shadow_height)] ? 1.0f : 0.0f;
        }
    }
    return lit / 16.0f;
}

// ------------------------------------------------------------------ sky

RGI_DEV bool sky_on(const rgi_frame& f) { return f.sky_zenith[3] > 0.5f; }

// skyEnabled, skyRadiance (sky.sh:15-31).
RGI_DEV F3 sky_radiance(const rgi_frame& f, F3 d)
{
    if (!sky_on(f)) return splat3(0.0f);
    F3 horizon = F3{f.sky_horizon[0], f.sky_horizon[1], f.sky_horizon[2]};
    float h = clampf_(d.y, -1.0f, 1.0f);
    if (h >= 0.0f) return mix3(horizon, F3{f.sky_zenith[0], f.sky_zenith[1], f.sky_zenith[2]}, sqrtf(h));
    return mix3(horizon, F3{f.sky_ground[0], f.sky_ground[1], f.sky_ground[2]}, sqrtf(-h));
}

} // namespace rgi
