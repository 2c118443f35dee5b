// The SVGF-style denoiser on gfx950 (svgf.sh, fs_svgf_variance.sc,
// fs_svgf_atrous.sc): a per-pixel variance estimate, then a-trous wavelet
// iterations with edge-stopping weights from the GI receiver position and
// normal planes. Full-screen passes over the GI planes: one thread per
// texel, 16 x 16 texels per block, every tap a point fetch with clamp.
#include "device.h"
#include "batch.h"
#include "internal.h"
#include "rgi_math.h"

namespace rgi {
namespace {

struct SvgfPlanes {
    int w, h;
    F2 ts;
    RGI_DEV uint32_t at(F2 uv) const { return texel_index(uv, w, h); }
    RGI_DEV F4 f4(const float* p, F2 uv) const { return load_f4(p, at(uv)); }
    RGI_DEV F4 h4(const uint16_t* p, F2 uv) const { return load_h4(p, at(uv)); }
};

// geometryWeight (svgf.sh:27-37).
RGI_DEV float geometry_weight(float scale, F3 p, F3 n, F3 np, F3 nn)
{
    if (dot(nn, nn) < 0.5f) return 0.0f;
    float plane = fabsf(dot(n, np - p)) / (0.02f * scale);
    float normal_weight = rgi_pow64(maxf_(dot(n, normalize(nn)), 0.0f));
    return rgi_expf(-plane) * normal_weight;
}

// isInside (svgf.sh:39-42).
RGI_DEV bool inside(F2 uv) { return uv.x >= 0.0f && uv.x <= 1.0f && uv.y >= 0.0f && uv.y <= 1.0f; }

// splineWeight (fs_svgf_atrous.sc:29-33).
RGI_DEV float spline_weight(int offset)
{
    int d = offset < 0 ? -offset : offset;
    return d == 0 ? 0.375f : (d == 1 ? 0.25f : 0.0625f);
}

// Texel of this thread in the 16 x 16-texel block (bx, by).
RGI_DEV bool block_texel(int w, int h, uint32_t bx, uint32_t by, int& x, int& y)
{
    x = (int)(bx * 16u + (threadIdx.x & 15u));
    y = (int)(by * 16u + (threadIdx.x >> 4));
    return x < w && y < h;
}

// fs_svgf_variance.sc:24-73.
RGI_DEV void variance_body(const SvgfPlanes& t, float scale, const uint16_t* color, const uint16_t* moments,
                           const float* receiver, const float* normal, uint16_t* out, uint32_t bx, uint32_t by)
{
    int x, y;
    if (!block_texel(t.w, t.h, bx, by, x, y)) return;
    uint32_t o = (uint32_t)y * (uint32_t)t.w + (uint32_t)x;
    F2 uv0 = F2{((float)x + 0.5f) * t.ts.x, ((float)y + 0.5f) * t.ts.y};
    F3 n = xyz(t.f4(normal, uv0));
    if (dot(n, n) < 0.5f) {
        store_h4(out, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    n = normalize(n);
    F3 c = xyz(t.h4(color, uv0));
    F3 m = xyz(t.h4(moments, uv0));
    float history = m.z;
    if (history >= 4.0f) {
        store_h4(out, o, F4{c.x, c.y, c.z, maxf_(m.y - m.x * m.x, 0.0f)});
        return;
    }
    F3 p = xyz(t.f4(receiver, uv0));
    float sx = 0.0f, sy = 0.0f, wsum = 0.0f;
    for (int dy = -3; dy <= 3; ++dy) {
        for (int dx = -3; dx <= 3; ++dx) {
            F2 uv = F2{uv0.x + (float)dx * t.ts.x, uv0.y + (float)dy * t.ts.y};
            if (!inside(uv)) continue;
            float wt = geometry_weight(scale, p, n, xyz(t.f4(receiver, uv)), xyz(t.f4(normal, uv)));
            F4 mm = t.h4(moments, uv);
            sx = sx + mm.x * wt;
            sy = sy + mm.y * wt;
            wsum = wsum + wt;
        }
    }
    float d = maxf_(wsum, 0.0001f);
    float mx = sx / d, my = sy / d;
    float variance = maxf_(my - mx * mx, 0.0f) * (4.0f / maxf_(history, 1.0f));
    store_h4(out, o, F4{c.x, c.y, c.z, variance});
}

// fs_svgf_atrous.sc:35-101.
RGI_DEV void atrous_body(const SvgfPlanes& t, float scale, const uint16_t* input, const float* receiver,
                         const float* normal, F2 spacing, uint16_t* out, uint32_t bx, uint32_t by)
{
    int x, y;
    if (!block_texel(t.w, t.h, bx, by, x, y)) return;
    uint32_t o = (uint32_t)y * (uint32_t)t.w + (uint32_t)x;
    F2 uv0 = F2{((float)x + 0.5f) * t.ts.x, ((float)y + 0.5f) * t.ts.y};
    F3 n = xyz(t.f4(normal, uv0));
    if (dot(n, n) < 0.5f) {
        store_h4(out, o, F4{0.0f, 0.0f, 0.0f, 0.0f});
        return;
    }
    n = normalize(n);
    F3 p = xyz(t.f4(receiver, uv0));
    F4 center = t.h4(input, uv0);
    float center_lum = luminance(xyz(center));
    float variance = 0.0f;
    for (int by = -1; by <= 1; ++by) {
        for (int bx = -1; bx <= 1; ++bx) {
            float wt = (bx == 0 ? 0.5f : 0.25f) * (by == 0 ? 0.5f : 0.25f);
            F2 uv = F2{uv0.x + (float)bx * t.ts.x, uv0.y + (float)by * t.ts.y};
            variance = variance + t.h4(input, uv).w * wt;
        }
    }
    float lum_scale = 1.0f / (4.0f * sqrtf(maxf_(variance, 0.0f)) + 0.0001f);
    F3 csum = splat3(0.0f);
    float vsum = 0.0f, wsum = 0.0f;
    for (int dy = -2; dy <= 2; ++dy) {
        for (int dx = -2; dx <= 2; ++dx) {
            F2 uv = F2{uv0.x + (float)dx * spacing.x, uv0.y + (float)dy * spacing.y};
            if (!inside(uv)) continue;
            F4 tap = t.h4(input, uv);
            float wt = spline_weight(dx) * spline_weight(dy);
            if (dx != 0 || dy != 0) {
                wt = wt * geometry_weight(scale, p, n, xyz(t.f4(receiver, uv)), xyz(t.f4(normal, uv)));
                wt = wt * rgi_expf(-fabsf(luminance(xyz(tap)) - center_lum) * lum_scale);
            }
            csum = csum + xyz(tap) * wt;
            vsum = vsum + (tap.w * wt) * wt;
            wsum = wsum + wt;
        }
    }
    F3 c = csum / wsum;
    store_h4(out, o, F4{c.x, c.y, c.z, vsum / (wsum * wsum)});
}

__global__ void __launch_bounds__(256) k_svgf_variance(SvgfPlanes t, float scale, const uint16_t* color,
                                                       const uint16_t* moments, const float* receiver,
                                                       const float* normal, uint16_t* out)
{
    variance_body(t, scale, color, moments, receiver, normal, out, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_svgf_atrous(SvgfPlanes t, float scale, const uint16_t* input,
                                                     const float* receiver, const float* normal, F2 spacing,
                                                     uint16_t* out)
{
    atrous_body(t, scale, input, receiver, normal, spacing, out, blockIdx.x, blockIdx.y);
}

__host__ __device__ inline SvgfPlanes planes_of(const rgi_frame& f)
{
    return SvgfPlanes{f.gi_width, f.gi_height, F2{1.0f / (float)f.gi_width, 1.0f / (float)f.gi_height}};
}

// The batched forms: workgroups [first[v], first[v + 1]) run view v's
// 16 x 16-texel blocks of its GI planes (views without the denoiser have no
// workgroups; a-trous iteration `it` skips views with fewer iterations).
__global__ void __launch_bounds__(256) k_svgf_variance_batch(const BatchView* __restrict__ views,
                                                             const uint32_t* __restrict__ first, uint32_t count)
{
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    const uint32_t gx = ((uint32_t)v.f.gi_width + 15u) / 16u;
    variance_body(planes_of(v.f), v.f.scene_scale, v.color, v.moments, v.res.receiver, v.res.normal, v.den_out[0],
                  b % gx, b / gx);
}

__global__ void __launch_bounds__(256) k_svgf_atrous_batch(const BatchView* __restrict__ views,
                                                           const uint32_t* __restrict__ first, uint32_t count, int it)
{
    uint32_t b;
    const BatchView& v = views[batch_view(first, count, blockIdx.x, b)];
    if (it >= (int)v.atrous) return;
    const SvgfPlanes t = planes_of(v.f);
    const float step = (float)(1 << it);
    const uint32_t gx = ((uint32_t)v.f.gi_width + 15u) / 16u;
    atrous_body(t, v.f.scene_scale, v.den_out[it], v.res.receiver, v.res.normal, F2{step * t.ts.x, step * t.ts.y},
                v.den_out[it + 1], b % gx, b / gx);
}

dim3 grid_of(const rgi_frame& f)
{
    return dim3((uint32_t)(f.gi_width + 15) / 16u, (uint32_t)(f.gi_height + 15) / 16u);
}

} // namespace

uint32_t svgf_blocks(const rgi_frame& f)
{
    const dim3 g = grid_of(f);
    return g.x * g.y;
}

int svgf_batch(const BatchView* views, const uint32_t* first, uint32_t count, uint32_t blocks, int iterations,
               hipStream_t stream)
{
    if (blocks == 0u) return RGI_OK;
    hipLaunchKernelGGL(k_svgf_variance_batch, dim3(blocks), dim3(256), 0, stream, views, first, count);
    for (int it = 0; it < iterations; ++it)
        hipLaunchKernelGGL(k_svgf_atrous_batch, dim3(blocks), dim3(256), 0, stream, views, first, count, it);
    return check_launch("svgf (batch)");
}

int svgf_variance(const rgi_frame& f, const uint16_t* color, const uint16_t* moments, const float* receiver,
                  const float* normal, uint16_t* out, hipStream_t stream)
{
    hipLaunchKernelGGL(k_svgf_variance, grid_of(f), dim3(256), 0, stream, planes_of(f), f.scene_scale, color, moments,
                       receiver, normal, out);
    return check_launch("svgf variance");
}

int svgf_atrous(const rgi_frame& f, const uint16_t* input, const float* receiver, const float* normal, int step,
                uint16_t* out, hipStream_t stream)
{
    SvgfPlanes t = planes_of(f);
    F2 spacing = F2{(float)step * t.ts.x, (float)step * t.ts.y};
    hipLaunchKernelGGL(k_svgf_atrous, grid_of(f), dim3(256), 0, stream, t, f.scene_scale, input, receiver, normal,
                       spacing, out);
    return check_launch("svgf atrous");
}

} // namespace rgi
