// Reference accumulation and the error metric (fs_reference.sc,
// cs_error.sc), in FP64 with a fixed reduction order so that the measured
// error is reproducible run to run:
//   1. one thread per pixel; a 256-pixel block sums its pixels with a tree in
//      LDS (strides 128 .. 1) into one partial per block;
//   2. one workgroup: thread t sums partials t, t + 256, ... in order, then
//      the same tree over the 256 threads.
#include "device.h"
#include "internal.h"

namespace rgi {
namespace {

constexpr int kBlock = 256;
constexpr int kTerms = 5; // relative, squared rgb, image luminance, reference luminance, count

// The compared image at pixel i: the composite itself (`composited`), or the
// GI re-modulated by the albedo plus the DI planes; geometry pixels only.
RGI_DEV bool image_at(const rgi_gbuffer& g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds,
                      bool composited, uint32_t i, F3& L)
{
    if (g.depth[i] >= kSkyDepth) return false;
    F4 c = load_h4(color, i);
    if (composited) {
        L = xyz(c);
        return true;
    }
    F4 a = unorm8(g.albedo[i]);
    float metallic = unorm8(g.normal[i]).w;
    L = F3{c.x, c.y, c.z} * (F3{a.x, a.y, a.z} * (1.0f - metallic));
    if (dd) L = L + xyz(load_h4(dd, i));
    if (ds) L = L + xyz(load_h4(ds, i));
    return true;
}

__global__ void __launch_bounds__(kBlock) k_ref_accumulate(rgi_gbuffer g, const uint16_t* color, const uint16_t* dd,
                                                           const uint16_t* ds, bool composited, double* accum)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)g.width * (uint32_t)g.height) return;
    F3 L;
    if (!image_at(g, color, dd, ds, composited, i, L)) return;
    if (L.x != L.x || L.y != L.y || L.z != L.z) return;
    accum[3u * i + 0u] += (double)L.x;
    accum[3u * i + 1u] += (double)L.y;
    accum[3u * i + 2u] += (double)L.z;
}

RGI_DEV double lum(double r, double g, double b) { return (0.2126 * r + 0.7152 * g) + 0.0722 * b; }

RGI_DEV void tree(double (*s)[kBlock], double* v)
{
    for (int k = 0; k < kTerms; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride)
            for (int k = 0; k < kTerms; ++k) s[k][threadIdx.x] = s[k][threadIdx.x] + s[k][threadIdx.x + stride];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) k_ref_error(rgi_gbuffer g, const uint16_t* color, const uint16_t* dd,
                                                      const uint16_t* ds, bool composited, const double* accum,
                                                      double inv_frames, double* partial)
{
    __shared__ double s[kTerms][kBlock];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[kTerms] = {0.0, 0.0, 0.0, 0.0, 0.0};
    F3 L;
    if (i < (uint32_t)g.width * (uint32_t)g.height && image_at(g, color, dd, ds, composited, i, L)) {
        double rr = accum[3u * i] * inv_frames, rg = accum[3u * i + 1u] * inv_frames, rb = accum[3u * i + 2u] * inv_frames;
        double lr = (double)L.x, lg = (double)L.y, lb = (double)L.z;
        double li = lum(lr, lg, lb), lref = lum(rr, rg, rb);
        double d = li - lref;
        v[0] = (d * d) / (lref * lref + 0.01);
        v[1] = ((lr - rr) * (lr - rr) + (lg - rg) * (lg - rg)) + (lb - rb) * (lb - rb);
        v[2] = li;
        v[3] = lref;
        v[4] = 1.0;
    }
    tree(s, v);
    if (threadIdx.x == 0)
        for (int k = 0; k < kTerms; ++k) partial[(size_t)blockIdx.x * kTerms + k] = s[k][0];
}

__global__ void __launch_bounds__(kBlock) k_ref_final(const double* partial, uint32_t blocks, double* out)
{
    __shared__ double s[kTerms][kBlock];
    double v[kTerms] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock)
        for (int k = 0; k < kTerms; ++k) v[k] = v[k] + partial[(size_t)b * kTerms + k];
    tree(s, v);
    if (threadIdx.x == 0) {
        double count = s[4][0];
        double n = count > 1.0 ? count : 1.0;
        out[0] = s[0][0] / n;
        out[1] = s[1][0] / (3.0 * n);
        out[2] = s[2][0] / n;
        out[3] = s[3][0] / n;
        out[4] = count;
    }
}

} // namespace
} // namespace rgi

using namespace rgi;

namespace {

int accumulate(const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds, bool composited,
               double* accum, void* stream)
{
    if (!g || !g->depth || (!composited && (!g->albedo || !g->normal)) || !color || !accum)
        return fail(RGI_E_INVALID, "reference_accumulate: G-buffer planes, image and accum are required");
    if (g->width <= 0 || g->height <= 0) return fail(RGI_E_INVALID, "reference_accumulate: empty frame");
    uint32_t n = (uint32_t)g->width * (uint32_t)g->height;
    hipLaunchKernelGGL(k_ref_accumulate, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, as_stream(stream), *g, color,
                       dd, ds, composited, accum);
    return check_launch("reference accumulate");
}

int error(rgi_context* ctx, const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd, const uint16_t* ds,
          bool composited, const double* accum, uint32_t frames, double* out, void* stream)
{
    if (!ctx || !g || !g->depth || (!composited && (!g->albedo || !g->normal)) || !color || !accum || !out)
        return fail(RGI_E_INVALID, "reference_error: context, planes, accum and out are required");
    if (frames == 0) return fail(RGI_E_INVALID, "reference_error: no reference frames");
    if (g->width <= 0 || g->height <= 0) return fail(RGI_E_INVALID, "reference_error: empty frame");
    uint32_t n = (uint32_t)g->width * (uint32_t)g->height;
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    double* partial = static_cast<double*>(context_scratch(ctx, (size_t)blocks * kTerms * sizeof(double)));
    if (!partial) return RGI_E_NOMEM;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_ref_error, dim3(blocks), dim3(kBlock), 0, s, *g, color, dd, ds, composited, accum,
                       1.0 / (double)frames, partial);
    hipLaunchKernelGGL(k_ref_final, dim3(1), dim3(kBlock), 0, s, partial, blocks, out);
    return check_launch("reference error");
}

} // namespace

extern "C" int rgi_reference_accumulate(const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd,
                                        const uint16_t* ds, double* accum, void* stream)
{
    return accumulate(g, color, dd, ds, false, accum, stream);
}

extern "C" int rgi_reference_error(rgi_context* ctx, const rgi_gbuffer* g, const uint16_t* color, const uint16_t* dd,
                                   const uint16_t* ds, const double* accum, uint32_t frames, double* out, void* stream)
{
    return error(ctx, g, color, dd, ds, false, accum, frames, out, stream);
}

extern "C" int rgi_reference_accumulate_composite(const rgi_gbuffer* g, const uint16_t* hdr, double* accum,
                                                  void* stream)
{
    return accumulate(g, hdr, nullptr, nullptr, true, accum, stream);
}

extern "C" int rgi_reference_error_composite(rgi_context* ctx, const rgi_gbuffer* g, const uint16_t* hdr,
                                             const double* accum, uint32_t frames, double* out, void* stream)
{
    return error(ctx, g, hdr, nullptr, nullptr, true, accum, frames, out, stream);
}
