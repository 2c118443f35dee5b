// The neural radiance cache on gfx950: multiresolution hash encoding, a
// 64-wide MLP with 5 hidden ReLU layers on the matrix cores, relative L2
// loss, Adam on FP32 master weights.
//
// Matrix-core mapping. Every layer product runs as 16 x 16 x 32 MFMAs with the
// layer's outputs as rows and the samples as columns: D (64 x 16 samples) =
// W (64 x 64) * X (64 x 16). A wave owns 64 samples, four column blocks of 16,
// so one A operand (a weight fragment) feeds four MFMAs. The output fragment
// of an MFMA puts rows 4g..4g+3 of column s in lane 16g + s; the B operand of
// the next layer wants 8 K values per lane in the same lane groups. Feeding
// the output registers straight back as the next B operand therefore fixes a
// permutation of the K index, phi(32j + 8g + q) = 16(2j + q/4) + 4g + q%4,
// which the weight copies absorb: the device copy of a layer's weights stores
// W[o][phi(k)] where the logical matrix has W[o][k]. Activations never leave
// the registers between layers. The backward pass uses the transposed copies
// the same way, and the weight gradients are plain GEMMs over the samples
// from activations and gradients stored [feature][sample].
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/restir_amd_nrc.h"
#include "device.h"
#include "internal.h"

struct rgi_nrc {
    rgi_nrc_config config;
    uint32_t res[16];
    uint32_t entries = 0; // per level
    size_t table_floats = 0;
    uint32_t steps = 0;
    // FP32 master parameters, Adam moments, gradients.
    float* table = nullptr;
    float* table_m = nullptr;
    float* table_v = nullptr;
    float* table_grad = nullptr;
    uint32_t* table_h = nullptr; // FP16 pairs, the copy the encoding reads
    float* mlp = nullptr;
    float* mlp_m = nullptr;
    float* mlp_v = nullptr;
    float* mlp_grad = nullptr;
    // Device weight copies (matrix-core fragments).
    __bf16* w_fwd = nullptr;
    __bf16* w_bwd = nullptr;
    _Float16* w_fwd_h = nullptr;
    uint8_t* w_fwd_8 = nullptr;
    float* scales = nullptr;      // [0..5] weight scales, [8..13] activation scales, [16..21] 1/(sa*sw)
    uint32_t* act_max = nullptr;  // max |a_k| of the last training batch, float bits, k = 0..5
    float* loss = nullptr;        // internal loss accumulator
    // Training scratch, grown on demand.
    __bf16* acts = nullptr;       // [6][tiles][64][256]
    __bf16* deltas = nullptr;     // [6][tiles][64][256]
    float* partial = nullptr;     // [chunks][RGI_NRC_MLP_PARAMS]
    uint32_t capacity_tiles = 0;
};

namespace rgi {
namespace nrc {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kTile = 256;       // samples per training workgroup (4 waves x 64)
constexpr int kChunkTiles = 4;   // tiles per weight-gradient workgroup
constexpr int kFwdBlocks = 42;   // fragments (64 lanes x 8) of the forward copy
constexpr int kBwdBlocks = 40;   // fragments of the transposed copy
constexpr int kFrag = 512;
constexpr uint32_t kPrime1 = 2654435761u;
constexpr uint32_t kPrime2 = 805459861u;

struct Grid {
    uint32_t res[16];
    uint32_t stride[16]; // res + 1 for a densely indexed level, 0 for a hashed one
    uint32_t mask;       // entries per level - 1
    uint32_t entries;
    float lo[3];
    float inv_extent[3];
};

// K index of slot (g, q) of fragment j in the hidden-layer order.
RGI_DEV int phi(int j, int g, int q) { return 16 * (2 * j + (q >> 2)) + 4 * g + (q & 3); }

RGI_DEV f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
RGI_DEV f32x4 mfma_f16(f16x8 a, f16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
RGI_DEV f32x4 mfma_fp8(long a, long b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0); }

RGI_DEV float half_lo(uint32_t v) { return half_to_float((uint16_t)(v & 0xffffu)); }
RGI_DEV float half_hi(uint32_t v) { return half_to_float((uint16_t)(v >> 16)); }

// Clamp to [0, 1]; NaN goes to 0, so every corner index stays in its level.
RGI_DEV float unit(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// Corner k (x + 2 y + 4 z) of level l around p: entry index and trilinear
// weight.
RGI_DEV void corner(const Grid& gp, int l, float px, float py, float pz, int k, uint32_t& idx, float& w)
{
    uint32_t r = gp.res[l];
    float rf = (float)r;
    float fx = px * rf, fy = py * rf, fz = pz * rf;
    int cx = min((int)floorf(fx), (int)r - 1), cy = min((int)floorf(fy), (int)r - 1),
        cz = min((int)floorf(fz), (int)r - 1);
    float wx = fx - (float)cx, wy = fy - (float)cy, wz = fz - (float)cz;
    uint32_t x = (uint32_t)cx + (uint32_t)(k & 1), y = (uint32_t)cy + (uint32_t)((k >> 1) & 1),
             z = (uint32_t)cz + (uint32_t)(k >> 2);
    w = (((k & 1) ? wx : 1.0f - wx) * ((k & 2) ? wy : 1.0f - wy)) * ((k & 4) ? wz : 1.0f - wz);
    uint32_t s = gp.stride[l];
    idx = s ? x + y * s + z * (s * s) : ((x ^ (y * kPrime1) ^ (z * kPrime2)) & gp.mask);
}

// Features of levels 4g .. 4g+3 (eight values, level-major).
RGI_DEV void encode4(const Grid& gp, const uint32_t* table_h, float px, float py, float pz, int g, float* e)
{
    for (int m = 0; m < 4; ++m) {
        int l = 4 * g + m;
        const uint32_t* level = table_h + (size_t)l * gp.entries;
        float a0 = 0.0f, a1 = 0.0f;
        for (int k = 0; k < 8; ++k) {
            uint32_t idx;
            float w;
            corner(gp, l, px, py, pz, k, idx, w);
            uint32_t v = level[idx];
            a0 = a0 + w * half_lo(v);
            a1 = a1 + w * half_hi(v);
        }
        e[2 * m] = a0;
        e[2 * m + 1] = a1;
    }
}

// A point's inputs: position mapped into [0, 1]^3, normal, albedo.
struct Point {
    float p[3];
    float n[3];
    float a[3];
};

RGI_DEV Point point_of(const Grid& gp, const float* rec)
{
    Point pt;
    for (int c = 0; c < 3; ++c) {
        pt.p[c] = unit((rec[c] - gp.lo[c]) * gp.inv_extent[c]);
        pt.n[c] = rec[4 + c];
        pt.a[c] = rec[8 + c];
    }
    return pt;
}

// Inputs 32 .. 39 held by lane group 0: normal * 0.5 + 0.5, albedo, 1, 0.
RGI_DEV void extras(const Point& pt, float* e)
{
    e[0] = pt.n[0] * 0.5f + 0.5f;
    e[1] = pt.n[1] * 0.5f + 0.5f;
    e[2] = pt.n[2] * 0.5f + 0.5f;
    e[3] = pt.a[0];
    e[4] = pt.a[1];
    e[5] = pt.a[2];
    e[6] = 1.0f;
    e[7] = 0.0f;
}

RGI_DEV bf16x8 pack_bf16(const float* e)
{
    bf16x8 v;
    for (int q = 0; q < 8; ++q) v[q] = (__bf16)e[q];
    return v;
}

RGI_DEV bf16x8 frag(const __bf16* base, int block, int lane)
{
    return *reinterpret_cast<const bf16x8*>(base + (size_t)block * kFrag + lane * 8);
}

RGI_DEV float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

RGI_DEV float wave_sum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    return v;
}

RGI_DEV void note_max(uint32_t* slot, float v)
{
    v = wave_max(v);
    if ((threadIdx.x & 63u) == 0u) atomicMax(slot, __float_as_uint(v));
}

// [feature][sample] store of one fragment's four rows for the sample of this
// lane.
RGI_DEV void store_rows(__bf16* plane, int t, int g, int sample, const float* v)
{
    for (int r = 0; r < 4; ++r) plane[(size_t)(16 * t + 4 * g + r) * kTile + sample] = (__bf16)v[r];
}

// ------------------------------------------------------------------ train

struct TrainArgs {
    Grid gp;
    const float* records;
    const float* targets;
    uint32_t count;
    float inv_count;
    const __bf16* w_fwd;
    const __bf16* w_bwd;
    const uint32_t* table_h;
    float* table_grad;
    __bf16* acts;   // [6][tiles][64][256]
    __bf16* deltas; // [6][tiles][64][256]
    uint32_t tiles;
    float* loss;
    uint32_t* act_max;
};

// Forward, loss and backward of 64 samples per wave: stores the activations
// and back-propagated gradients for the weight-gradient GEMMs and scatters
// the input gradient into the table gradient.
__global__ void __launch_bounds__(256) k_nrc_train(TrainArgs a)
{
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u), g = lane >> 4, s16 = lane & 15;
    const uint32_t tile = blockIdx.x;
    const size_t layer_stride = (size_t)a.tiles * 64 * kTile;
    bf16x8 X[2][4];
    Point pts[4];
    float amax = 0.0f;
    for (int b = 0; b < 4; ++b) {
        int local = wave * 64 + b * 16 + s16;
        uint32_t s = tile * kTile + (uint32_t)local;
        float rec[12] = {0};
        if (s < a.count)
            for (int c = 0; c < 12; ++c) rec[c] = a.records[(size_t)s * 12 + c];
        pts[b] = point_of(a.gp, rec);
        float e[8], x[8];
        encode4(a.gp, a.table_h, pts[b].p[0], pts[b].p[1], pts[b].p[2], g, e);
        if (g == 0) extras(pts[b], x);
        else
            for (int q = 0; q < 8; ++q) x[q] = 0.0f;
        X[0][b] = pack_bf16(e);
        X[1][b] = pack_bf16(x);
        __bf16* a0 = a.acts + (size_t)tile * 64 * kTile;
        for (int q = 0; q < 8; ++q) {
            a0[(size_t)(8 * g + q) * kTile + local] = X[0][b][q];
            a0[(size_t)(32 + 8 * g + q) * kTile + local] = X[1][b][q];
            amax = fmaxf(amax, fabsf((float)X[0][b][q]));
            amax = fmaxf(amax, fabsf((float)X[1][b][q]));
        }
    }
    note_max(a.act_max + 0, amax);

    // Hidden layers 1..5.
    uint64_t masks[5];
    for (int k = 0; k < 5; ++k) {
        f32x4 acc[4][4];
        for (int b = 0; b < 4; ++b)
            for (int t = 0; t < 4; ++t) acc[b][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int j = 0; j < 2; ++j)
            for (int t = 0; t < 4; ++t) {
                bf16x8 w = frag(a.w_fwd, k * 8 + t * 2 + j, lane);
                for (int b = 0; b < 4; ++b) acc[b][t] = mfma_bf16(w, X[j][b], acc[b][t]);
            }
        uint64_t mask = 0;
        __bf16* plane = a.acts + (size_t)(k + 1) * layer_stride + (size_t)tile * 64 * kTile;
        amax = 0.0f;
        for (int b = 0; b < 4; ++b) {
            int local = wave * 64 + b * 16 + s16;
            float v[4][4];
            for (int t = 0; t < 4; ++t)
                for (int r = 0; r < 4; ++r) {
                    float z = acc[b][t][r];
                    bool on = z > 0.0f;
                    mask |= (uint64_t)on << (b * 16 + t * 4 + r);
                    v[t][r] = on ? z : 0.0f;
                }
            for (int t = 0; t < 4; ++t) store_rows(plane, t, g, local, v[t]);
            for (int j = 0; j < 2; ++j) {
                float e[8];
                for (int r = 0; r < 4; ++r) {
                    e[r] = v[2 * j][r];
                    e[4 + r] = v[2 * j + 1][r];
                }
                X[j][b] = pack_bf16(e);
                for (int q = 0; q < 8; ++q) amax = fmaxf(amax, (float)X[j][b][q]);
            }
        }
        masks[k] = mask;
        note_max(a.act_max + k + 1, amax);
    }

    // Output layer, loss and its gradient (lane group 0 holds rows 0..3).
    f32x4 y[4];
    for (int b = 0; b < 4; ++b) y[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int j = 0; j < 2; ++j) {
        bf16x8 w = frag(a.w_fwd, 40 + j, lane);
        for (int b = 0; b < 4; ++b) y[b] = mfma_bf16(w, X[j][b], y[b]);
    }
    float loss = 0.0f;
    bf16x8 D[2][4];
    __bf16* d6 = a.deltas + (size_t)5 * layer_stride + (size_t)tile * 64 * kTile;
    for (int b = 0; b < 4; ++b) {
        int local = wave * 64 + b * 16 + s16;
        uint32_t s = tile * kTile + (uint32_t)local;
        float d[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (g == 0 && s < a.count) {
            const float* t = a.targets + (size_t)s * 4;
            float m = t[3];
            float mean = ((y[b][0] + y[b][1]) + y[b][2]) / 3.0f;
            float denom = mean * mean + 0.01f;
            float e0 = y[b][0] - t[0], e1 = y[b][1] - t[1], e2 = y[b][2] - t[2];
            loss = loss + (m * ((e0 * e0 + e1 * e1) + e2 * e2)) / denom;
            float scale = ((2.0f * m) * a.inv_count) / denom;
            d[0] = e0 * scale;
            d[1] = e1 * scale;
            d[2] = e2 * scale;
        }
        store_rows(d6, 0, g, local, d);
        float e[8] = {d[0], d[1], d[2], d[3], 0.0f, 0.0f, 0.0f, 0.0f};
        D[0][b] = pack_bf16(e);
    }
    loss = wave_sum(loss);
    if (lane == 0 && loss != 0.0f) atomicAdd(a.loss, loss * a.inv_count);

    // Back through W6 (K = the 16 output rows, slots q < 4 only), then W5..W2.
    for (int k = 5; k >= 1; --k) {
        f32x4 acc[4][4];
        for (int b = 0; b < 4; ++b)
            for (int t = 0; t < 4; ++t) acc[b][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if (k == 5) {
            for (int t = 0; t < 4; ++t) {
                bf16x8 w = frag(a.w_bwd, t, lane);
                for (int b = 0; b < 4; ++b) acc[b][t] = mfma_bf16(w, D[0][b], acc[b][t]);
            }
        } else {
            // k + 1 is the layer whose gradient D holds; block base of W_{k+1}^T.
            int base = 4 + (4 - k) * 8;
            for (int j = 0; j < 2; ++j)
                for (int t = 0; t < 4; ++t) {
                    bf16x8 w = frag(a.w_bwd, base + t * 2 + j, lane);
                    for (int b = 0; b < 4; ++b) acc[b][t] = mfma_bf16(w, D[j][b], acc[b][t]);
                }
        }
        // Gradient of layer k's pre-activation: times ReLU'(z_k).
        __bf16* plane = a.deltas + (size_t)(k - 1) * layer_stride + (size_t)tile * 64 * kTile;
        for (int b = 0; b < 4; ++b) {
            int local = wave * 64 + b * 16 + s16;
            float v[4][4];
            for (int t = 0; t < 4; ++t)
                for (int r = 0; r < 4; ++r)
                    v[t][r] = ((masks[k - 1] >> (b * 16 + t * 4 + r)) & 1u) ? acc[b][t][r] : 0.0f;
            for (int t = 0; t < 4; ++t) store_rows(plane, t, g, local, v[t]);
            for (int j = 0; j < 2; ++j) {
                float e[8];
                for (int r = 0; r < 4; ++r) {
                    e[r] = v[2 * j][r];
                    e[4 + r] = v[2 * j + 1][r];
                }
                D[j][b] = pack_bf16(e);
            }
        }
    }

    // Input gradient of the 32 encoded features: W1^T (rows 0..31) times D.
    for (int b = 0; b < 4; ++b) {
        f32x4 dx[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
        for (int j = 0; j < 2; ++j)
            for (int t = 0; t < 2; ++t) dx[t] = mfma_bf16(frag(a.w_bwd, 36 + t * 2 + j, lane), D[j][b], dx[t]);
        // Rows 16t + 4g + r: level 8t + 2g + r/2, feature r%2.
        for (int t = 0; t < 2; ++t)
            for (int h = 0; h < 2; ++h) {
                float d0 = dx[t][2 * h], d1 = dx[t][2 * h + 1];
                if (d0 == 0.0f && d1 == 0.0f) continue;
                int l = 8 * t + 2 * g + h;
                float* level = a.table_grad + (size_t)l * a.gp.entries * 2;
                for (int k = 0; k < 8; ++k) {
                    uint32_t idx;
                    float w;
                    corner(a.gp, l, pts[b].p[0], pts[b].p[1], pts[b].p[2], k, idx, w);
                    if (d0 != 0.0f) unsafeAtomicAdd(level + 2 * idx, w * d0);
                    if (d1 != 0.0f) unsafeAtomicAdd(level + 2 * idx + 1, w * d1);
                }
            }
    }
}

// Weight gradients: dW_k[o][i] = sum over samples of delta_k[o][s] a_{k-1}[i][s]
// for one chunk of kChunkTiles tiles; one workgroup per (chunk, layer), wave w
// owns output rows 16w..16w+15 (layer 6: column block w of row block 0).
__global__ void __launch_bounds__(256) k_nrc_wgrad(const __bf16* acts, const __bf16* deltas, uint32_t tiles,
                                                   float* partial)
{
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u), g = lane >> 4, s16 = lane & 15;
    const int layer = (int)blockIdx.y; // 0..5 -> W1..W6
    const uint32_t chunk = blockIdx.x;
    const size_t layer_stride = (size_t)tiles * 64 * kTile;
    const __bf16* dl = deltas + (size_t)layer * layer_stride;
    const __bf16* al = acts + (size_t)layer * layer_stride;
    const int row_tile = layer == 5 ? 0 : wave;
    const int col_lo = layer == 5 ? wave : 0, col_hi = layer == 5 ? wave + 1 : 4;
    f32x4 acc[4];
    for (int c = 0; c < 4; ++c) acc[c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tt = 0; tt < kChunkTiles; ++tt) {
        uint32_t tile = chunk * kChunkTiles + (uint32_t)tt;
        if (tile >= tiles) break;
        const __bf16* dt = dl + (size_t)tile * 64 * kTile;
        const __bf16* at = al + (size_t)tile * 64 * kTile;
        for (int kb = 0; kb < kTile / 32; ++kb) {
            int s0 = kb * 32 + 8 * g;
            bf16x8 A = *reinterpret_cast<const bf16x8*>(dt + (size_t)(16 * row_tile + s16) * kTile + s0);
            for (int c = col_lo; c < col_hi; ++c) {
                bf16x8 B = *reinterpret_cast<const bf16x8*>(at + (size_t)(16 * c + s16) * kTile + s0);
                acc[c] = mfma_bf16(A, B, acc[c]);
            }
        }
    }
    float* out = partial + (size_t)chunk * RGI_NRC_MLP_PARAMS + (size_t)layer * 4096;
    for (int c = col_lo; c < col_hi; ++c)
        for (int r = 0; r < 4; ++r) {
            int o = 16 * row_tile + 4 * g + r, i = 16 * c + s16;
            if (layer == 5 && o >= 3) continue;
            out[o * 64 + i] = acc[c][r];
        }
}

__global__ void k_nrc_reduce(const float* partial, uint32_t chunks, float* grad)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= RGI_NRC_MLP_PARAMS) return;
    float s = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) s = s + partial[(size_t)c * RGI_NRC_MLP_PARAMS + p];
    grad[p] = s;
}

// ------------------------------------------------------------------- Adam

struct AdamArgs {
    float lr, beta1, beta2, eps, c1, c2;
};

RGI_DEV float adam(const AdamArgs& h, float g, float& m, float& v, float p)
{
    m = h.beta1 * m + (1.0f - h.beta1) * g;
    v = h.beta2 * v + (1.0f - h.beta2) * (g * g);
    float mhat = m / h.c1;
    float vhat = v / h.c2;
    return p - (h.lr * mhat) / (sqrtf(vhat) + h.eps);
}

// Logical weight behind forward fragment slot (block, lane, q); 0 for padding.
RGI_DEV float fwd_weight(const float* mlp, int block, int lane, int q)
{
    int g = lane >> 4, s16 = lane & 15;
    if (block < 40) {
        int k = block / 8, t = (block % 8) / 2, j = block % 2;
        int o = 16 * t + s16;
        int col = k == 0 ? 32 * j + 8 * g + q : phi(j, g, q);
        return mlp[k * 4096 + o * 64 + col];
    }
    int j = block - 40;
    return s16 < 3 ? mlp[20480 + s16 * 64 + phi(j, g, q)] : 0.0f;
}

RGI_DEV float bwd_weight(const float* mlp, int block, int lane, int q)
{
    int g = lane >> 4, s16 = lane & 15;
    if (block < 4) { // W6^T: row = layer-5 feature, K = output row 4g + q (q < 4)
        int o = 4 * g + q;
        return (q < 4 && o < 3) ? mlp[20480 + o * 64 + 16 * block + s16] : 0.0f;
    }
    if (block < 36) { // W_k^T, k = 5..2
        int kk = (block - 4) / 8, t = ((block - 4) % 8) / 2, j = (block - 4) % 2;
        int k = 5 - kk; // 1-based layer
        return mlp[(k - 1) * 4096 + phi(j, g, q) * 64 + 16 * t + s16];
    }
    int t = (block - 36) / 2, j = (block - 36) % 2; // W1^T rows 0..31
    return mlp[phi(j, g, q) * 64 + 16 * t + s16];
}

// Layer of a forward / backward fragment (0-based W index).
RGI_DEV int fwd_layer(int block) { return block < 40 ? block / 8 : 5; }
RGI_DEV int bwd_layer(int block) { return block < 4 ? 5 : (block < 36 ? 4 - (block - 4) / 8 : 0); }

RGI_DEV uint8_t to_fp8(float x)
{
    x = fminf(fmaxf(x, -448.0f), 448.0f);
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xff);
}

// One workgroup per weight matrix: Adam on its parameters, then its
// fragments in every copy and its FP8 scales.
__global__ void __launch_bounds__(256) k_nrc_adam_mlp(float* mlp, float* m, float* v, const float* grad, AdamArgs h,
                                                      __bf16* w_fwd, __bf16* w_bwd, _Float16* w_fwd_h,
                                                      uint8_t* w_fwd_8, float* scales, uint32_t* act_max, int step)
{
    __shared__ float red[4];
    const int layer = (int)blockIdx.x;
    const int lo = layer * 4096, hi = layer == 5 ? 20480 + 192 : lo + 4096;
    float wmax = 0.0f;
    for (int p = lo + (int)threadIdx.x; p < hi; p += 256) {
        float mm = m[p], vv = v[p];
        float w = step ? adam(h, grad[p], mm, vv, mlp[p]) : mlp[p];
        if (step) {
            m[p] = mm;
            v[p] = vv;
            mlp[p] = w;
        }
        wmax = fmaxf(wmax, fabsf(w));
    }
    wmax = wave_max(wmax);
    if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = wmax;
    __syncthreads();
    wmax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    float sw = wmax > 0.0f ? 448.0f / wmax : 1.0f;
    float amax = __uint_as_float(act_max[layer]);
    float sa = amax > 0.0f ? 224.0f / amax : 1.0f;
    if (threadIdx.x == 0) {
        scales[layer] = sw;
        scales[8 + layer] = sa;
        scales[16 + layer] = 1.0f / (sa * sw);
    }
    for (int e = (int)threadIdx.x; e < kFwdBlocks * kFrag; e += 256) {
        int block = e / kFrag, lane = (e % kFrag) / 8, q = e % 8;
        if (fwd_layer(block) != layer) continue;
        float w = fwd_weight(mlp, block, lane, q);
        w_fwd[e] = (__bf16)w;
        w_fwd_h[e] = (_Float16)w;
        w_fwd_8[e] = to_fp8(w * sw);
    }
    for (int e = (int)threadIdx.x; e < kBwdBlocks * kFrag; e += 256) {
        int block = e / kFrag, lane = (e % kFrag) / 8, q = e % 8;
        if (bwd_layer(block) != layer) continue;
        w_bwd[e] = (__bf16)bwd_weight(mlp, block, lane, q);
    }
}

__global__ void k_nrc_clear_max(uint32_t* act_max, float* loss)
{
    if (threadIdx.x < 6) act_max[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *loss = 0.0f;
}

__global__ void k_nrc_copy_loss(const float* from, float* to)
{
    if (threadIdx.x == 0) *to = *from;
}

// Adam on the table: four floats (two entries) per thread; an entry pair
// whose gradient is all zero is left alone. The gradient is cleared.
__global__ void k_nrc_adam_table(float* table, float* m, float* v, float* grad, uint32_t* table_h, AdamArgs h,
                                 size_t quads)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= quads) return;
    float4 gq = reinterpret_cast<const float4*>(grad)[i];
    if (gq.x == 0.0f && gq.y == 0.0f && gq.z == 0.0f && gq.w == 0.0f) return;
    float4 pq = reinterpret_cast<float4*>(table)[i];
    float4 mq = reinterpret_cast<float4*>(m)[i];
    float4 vq = reinterpret_cast<float4*>(v)[i];
    float gs[4] = {gq.x, gq.y, gq.z, gq.w}, ps[4] = {pq.x, pq.y, pq.z, pq.w}, ms[4] = {mq.x, mq.y, mq.z, mq.w},
          vs[4] = {vq.x, vq.y, vq.z, vq.w};
    for (int c = 0; c < 4; ++c)
        if (gs[c] != 0.0f) ps[c] = adam(h, gs[c], ms[c], vs[c], ps[c]);
    reinterpret_cast<float4*>(table)[i] = make_float4(ps[0], ps[1], ps[2], ps[3]);
    reinterpret_cast<float4*>(m)[i] = make_float4(ms[0], ms[1], ms[2], ms[3]);
    reinterpret_cast<float4*>(v)[i] = make_float4(vs[0], vs[1], vs[2], vs[3]);
    reinterpret_cast<float4*>(grad)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint2 hv = make_uint2((uint32_t)float_to_half(ps[0]) | ((uint32_t)float_to_half(ps[1]) << 16),
                          (uint32_t)float_to_half(ps[2]) | ((uint32_t)float_to_half(ps[3]) << 16));
    reinterpret_cast<uint2*>(table_h)[i] = hv;
}

// Table gradient out (for the gradient entry point), then cleared.
__global__ void k_nrc_take_grad(float* grad, float* out, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (out) out[i] = grad[i];
    grad[i] = 0.0f;
}

RGI_DEV uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

RGI_DEV float uniform_pm(uint32_t seed, uint32_t stream, uint32_t i)
{
    uint32_t h = mix32(mix32(seed * 0x9e3779b9u + stream) ^ (i * 0x85ebca6bu + 0x165667b1u));
    return ((float)(h >> 8) * (1.0f / 16777216.0f)) * 2.0f - 1.0f;
}

__global__ void k_nrc_init_table(float* table, uint32_t* table_h, uint32_t seed, size_t entries_total)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= entries_total) return;
    float a = uniform_pm(seed, 1u, (uint32_t)(2 * e)) * 1e-4f, b = uniform_pm(seed, 1u, (uint32_t)(2 * e + 1)) * 1e-4f;
    table[2 * e] = a;
    table[2 * e + 1] = b;
    table_h[e] = (uint32_t)float_to_half(a) | ((uint32_t)float_to_half(b) << 16);
}

__global__ void k_nrc_init_mlp(float* mlp, uint32_t seed)
{
    int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= RGI_NRC_MLP_PARAMS) return;
    float limit = p < 20480 ? sqrtf(6.0f / 128.0f) : sqrtf(6.0f / 67.0f);
    mlp[p] = uniform_pm(seed, 2u, (uint32_t)p) * limit;
}

__global__ void k_nrc_half_table(const float* table, uint32_t* table_h, size_t entries_total)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= entries_total) return;
    table_h[e] = (uint32_t)float_to_half(table[2 * e]) | ((uint32_t)float_to_half(table[2 * e + 1]) << 16);
}

// --------------------------------------------------------------- inference

// Sources of query points: a record array, or the G-buffer's visible surfaces.
struct RecordSource {
    const float* records;
    uint32_t count;
    float* out;
    RGI_DEV bool fetch(const Grid& gp, uint32_t s, Point& pt) const
    {
        float rec[12] = {0};
        bool ok = s < count;
        if (ok)
            for (int c = 0; c < 12; ++c) rec[c] = records[(size_t)s * 12 + c];
        pt = point_of(gp, rec);
        return ok;
    }
    RGI_DEV void put(uint32_t s, const float* y) const
    {
        if (s < count) reinterpret_cast<float4*>(out)[s] = make_float4(y[0], y[1], y[2], 0.0f);
    }
};

template <int P>
struct Operand;
template <>
struct Operand<RGI_NRC_BF16> {
    typedef bf16x8 T;
    static RGI_DEV T pack(const float* e, float) { return pack_bf16(e); }
    static RGI_DEV T weights(const rgi_nrc& n, int block, int lane) { return frag(n.w_fwd, block, lane); }
    static RGI_DEV f32x4 mma(T a, T b, f32x4 c) { return mfma_bf16(a, b, c); }
};
template <>
struct Operand<RGI_NRC_FP16> {
    typedef f16x8 T;
    static RGI_DEV T pack(const float* e, float)
    {
        f16x8 v;
        for (int q = 0; q < 8; ++q) v[q] = (_Float16)e[q];
        return v;
    }
    static RGI_DEV T weights(const rgi_nrc& n, int block, int lane)
    {
        return *reinterpret_cast<const f16x8*>(n.w_fwd_h + (size_t)block * kFrag + lane * 8);
    }
    static RGI_DEV f32x4 mma(T a, T b, f32x4 c) { return mfma_f16(a, b, c); }
};
template <>
struct Operand<RGI_NRC_FP8> {
    typedef long T;
    static RGI_DEV T pack(const float* e, float scale)
    {
        uint64_t v = 0;
        for (int q = 0; q < 8; ++q) v |= (uint64_t)to_fp8(e[q] * scale) << (8 * q);
        return (long)v;
    }
    static RGI_DEV T weights(const rgi_nrc& n, int block, int lane)
    {
        return *reinterpret_cast<const long*>(n.w_fwd_8 + (size_t)block * kFrag + lane * 8);
    }
    static RGI_DEV f32x4 mma(T a, T b, f32x4 c) { return mfma_fp8(a, b, c); }
};

// The forward pass of 64 points per wave: Point pts[4] of this lane's samples
// in, y[4] (rows 0..3 in lane group 0) out.
template <int P>
RGI_DEV void forward(const rgi_nrc& n, const Grid& gp, const float* scales, const Point* pts, f32x4* y)
{
    typedef Operand<P> Op;
    const int lane = (int)(threadIdx.x & 63u), g = lane >> 4;
    typename Op::T X[2][4];
    float sa = P == RGI_NRC_FP8 ? scales[8] : 1.0f;
    for (int b = 0; b < 4; ++b) {
        float e[8], x[8];
        encode4(gp, n.table_h, pts[b].p[0], pts[b].p[1], pts[b].p[2], g, e);
        if (g == 0) extras(pts[b], x);
        else
            for (int q = 0; q < 8; ++q) x[q] = 0.0f;
        if (P != RGI_NRC_FP8) { // the BF16 / FP16 input is the rounded value
            X[0][b] = Op::pack(e, 1.0f);
            X[1][b] = Op::pack(x, 1.0f);
        } else {
            X[0][b] = Op::pack(e, sa);
            X[1][b] = Op::pack(x, sa);
        }
    }
    for (int k = 0; k < 5; ++k) {
        f32x4 acc[4][4];
        for (int b = 0; b < 4; ++b)
            for (int t = 0; t < 4; ++t) acc[b][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int j = 0; j < 2; ++j)
            for (int t = 0; t < 4; ++t) {
                typename Op::T w = Op::weights(n, k * 8 + t * 2 + j, lane);
                for (int b = 0; b < 4; ++b) acc[b][t] = Op::mma(w, X[j][b], acc[b][t]);
            }
        float deq = P == RGI_NRC_FP8 ? scales[16 + k] : 1.0f;
        float next = P == RGI_NRC_FP8 ? scales[8 + k + 1] : 1.0f;
        for (int b = 0; b < 4; ++b)
            for (int j = 0; j < 2; ++j) {
                float e[8];
                for (int r = 0; r < 4; ++r) {
                    float z0 = acc[b][2 * j][r] * deq, z1 = acc[b][2 * j + 1][r] * deq;
                    e[r] = z0 > 0.0f ? z0 : 0.0f;
                    e[4 + r] = z1 > 0.0f ? z1 : 0.0f;
                }
                X[j][b] = Op::pack(e, next);
            }
    }
    for (int b = 0; b < 4; ++b) y[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int j = 0; j < 2; ++j) {
        typename Op::T w = Op::weights(n, 40 + j, lane);
        for (int b = 0; b < 4; ++b) y[b] = Op::mma(w, X[j][b], y[b]);
    }
    if (P == RGI_NRC_FP8)
        for (int b = 0; b < 4; ++b) y[b] = y[b] * scales[16 + 5];
}

template <int P>
__global__ void __launch_bounds__(256) k_nrc_infer(rgi_nrc n, Grid gp, RecordSource src)
{
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u), g = lane >> 4, s16 = lane & 15;
    Point pts[4];
    uint32_t base = blockIdx.x * 256u + (uint32_t)wave * 64u;
    for (int b = 0; b < 4; ++b) src.fetch(gp, base + (uint32_t)(b * 16 + s16), pts[b]);
    f32x4 y[4];
    forward<P>(n, gp, n.scales, pts, y);
    if (g != 0) return;
    for (int b = 0; b < 4; ++b) {
        float v[3] = {y[b][0], y[b][1], y[b][2]};
        src.put(base + (uint32_t)(b * 16 + s16), v);
    }
}

} // namespace nrc

// ---------------------------------------------------------------- host side

namespace {

using namespace nrc;

Grid grid_of(const rgi_nrc& n)
{
    Grid gp;
    for (int l = 0; l < 16; ++l) {
        gp.res[l] = n.res[l];
        uint64_t s = (uint64_t)n.res[l] + 1;
        gp.stride[l] = s * s * s <= n.entries ? (uint32_t)s : 0u;
    }
    gp.mask = n.entries - 1;
    gp.entries = n.entries;
    for (int c = 0; c < 3; ++c) {
        gp.lo[c] = n.config.aabb_min[c];
        float ext = n.config.aabb_max[c] - n.config.aabb_min[c];
        gp.inv_extent[c] = 1.0f / ext;
    }
    return gp;
}

template <typename T>
int alloc(T** p, size_t count)
{
    if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return fail(RGI_E_NOMEM, "nrc: device allocation of %zu bytes failed", count * sizeof(T));
    }
    return RGI_OK;
}

int ensure_scratch(rgi_nrc* n, uint32_t tiles)
{
    if (tiles <= n->capacity_tiles) return RGI_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(n->acts);
    (void)hipFree(n->deltas);
    (void)hipFree(n->partial);
    n->acts = n->deltas = nullptr;
    n->partial = nullptr;
    n->capacity_tiles = 0;
    size_t per = (size_t)6 * tiles * 64 * kTile;
    uint32_t chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    int e;
    if ((e = alloc(&n->acts, per)) || (e = alloc(&n->deltas, per)) ||
        (e = alloc(&n->partial, (size_t)chunks * RGI_NRC_MLP_PARAMS)))
        return e;
    n->capacity_tiles = tiles;
    return RGI_OK;
}

AdamArgs adam_args(const rgi_nrc& n, uint32_t step)
{
    AdamArgs h;
    h.lr = n.config.learning_rate;
    h.beta1 = n.config.beta1;
    h.beta2 = n.config.beta2;
    h.eps = n.config.epsilon;
    h.c1 = (float)(1.0 - pow((double)n.config.beta1, (double)step));
    h.c2 = (float)(1.0 - pow((double)n.config.beta2, (double)step));
    return h;
}

// Forward, backward and the MLP gradient into n->mlp_grad (the table gradient
// accumulates into n->table_grad).
int gradients(rgi_nrc* n, const float* records, const float* targets, uint32_t count, hipStream_t stream)
{
    uint32_t tiles = (count + kTile - 1) / kTile;
    int e = ensure_scratch(n, tiles);
    if (e) return e;
    hipLaunchKernelGGL(k_nrc_clear_max, dim3(1), dim3(64), 0, stream, n->act_max, n->loss);
    TrainArgs a;
    a.gp = grid_of(*n);
    a.records = records;
    a.targets = targets;
    a.count = count;
    a.inv_count = 1.0f / (float)count;
    a.w_fwd = n->w_fwd;
    a.w_bwd = n->w_bwd;
    a.table_h = n->table_h;
    a.table_grad = n->table_grad;
    a.acts = n->acts;
    a.deltas = n->deltas;
    a.tiles = tiles;
    a.loss = n->loss;
    a.act_max = n->act_max;
    hipLaunchKernelGGL(k_nrc_train, dim3(tiles), dim3(256), 0, stream, a);
    if ((e = check_launch("nrc train"))) return e;
    uint32_t chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    hipLaunchKernelGGL(k_nrc_wgrad, dim3(chunks, 6), dim3(256), 0, stream, n->acts, n->deltas, tiles, n->partial);
    if ((e = check_launch("nrc weight gradients"))) return e;
    hipLaunchKernelGGL(k_nrc_reduce, dim3((RGI_NRC_MLP_PARAMS + 255) / 256), dim3(256), 0, stream, n->partial, chunks,
                       n->mlp_grad);
    return check_launch("nrc gradient reduction");
}

int refresh_copies(rgi_nrc* n, int step, hipStream_t stream)
{
    AdamArgs h = adam_args(*n, step ? n->steps : 1);
    hipLaunchKernelGGL(k_nrc_adam_mlp, dim3(6), dim3(256), 0, stream, n->mlp, n->mlp_m, n->mlp_v, n->mlp_grad, h,
                       n->w_fwd, n->w_bwd, n->w_fwd_h, n->w_fwd_8, n->scales, n->act_max, step);
    return check_launch("nrc adam (mlp)");
}

} // namespace
} // namespace rgi

using namespace rgi;

extern "C" {

void rgi_nrc_config_default(rgi_nrc_config* c)
{
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->log2_table = 19;
    c->base_resolution = 16;
    c->max_resolution = 512;
    c->seed = 1;
    c->learning_rate = 0.005f;
    c->beta1 = 0.9f;
    c->beta2 = 0.99f;
    c->epsilon = 1e-8f;
    for (int i = 0; i < 3; ++i) {
        c->aabb_min[i] = -1.0f;
        c->aabb_max[i] = 1.0f;
    }
}

int rgi_nrc_level_resolution(const rgi_nrc_config* c, uint32_t level, uint32_t* res)
{
    if (!c || !res || level >= RGI_NRC_LEVELS) return fail(RGI_E_INVALID, "nrc_level_resolution: bad argument");
    if (c->base_resolution < 2 || c->max_resolution < c->base_resolution || c->max_resolution > 65536)
        return fail(RGI_E_INVALID, "nrc: resolutions must satisfy 2 <= base <= max <= 65536");
    double b = exp((log((double)c->max_resolution) - log((double)c->base_resolution)) / (RGI_NRC_LEVELS - 1));
    double r = floor((double)c->base_resolution * pow(b, (double)level) + 1e-9);
    if (r > (double)c->max_resolution) r = (double)c->max_resolution;
    *res = (uint32_t)r;
    return RGI_OK;
}

int rgi_nrc_create(rgi_nrc** out, const rgi_nrc_config* config)
{
    if (!out || !config) return fail(RGI_E_INVALID, "nrc_create: out and config are required");
    *out = nullptr;
    if (config->log2_table < 10 || config->log2_table > 22)
        return fail(RGI_E_INVALID, "nrc_create: log2_table must be in 10..22");
    for (int c = 0; c < 3; ++c)
        if (!(config->aabb_max[c] > config->aabb_min[c])) return fail(RGI_E_INVALID, "nrc_create: empty bounds");
    if (!(config->learning_rate > 0.0f) || !(config->beta1 >= 0.0f && config->beta1 < 1.0f) ||
        !(config->beta2 >= 0.0f && config->beta2 < 1.0f) || !(config->epsilon > 0.0f))
        return fail(RGI_E_INVALID, "nrc_create: Adam hyperparameters out of range");
    rgi_nrc* n = new rgi_nrc();
    n->config = *config;
    for (uint32_t l = 0; l < 16; ++l) {
        int e = rgi_nrc_level_resolution(config, l, &n->res[l]);
        if (e) {
            delete n;
            return e;
        }
    }
    n->entries = 1u << config->log2_table;
    n->table_floats = (size_t)RGI_NRC_LEVELS * n->entries * 2;
    int e;
    if ((e = alloc(&n->table, n->table_floats)) || (e = alloc(&n->table_m, n->table_floats)) ||
        (e = alloc(&n->table_v, n->table_floats)) || (e = alloc(&n->table_grad, n->table_floats)) ||
        (e = alloc(&n->table_h, n->table_floats / 2)) || (e = alloc(&n->mlp, RGI_NRC_MLP_PARAMS)) ||
        (e = alloc(&n->mlp_m, RGI_NRC_MLP_PARAMS)) || (e = alloc(&n->mlp_v, RGI_NRC_MLP_PARAMS)) ||
        (e = alloc(&n->mlp_grad, RGI_NRC_MLP_PARAMS)) || (e = alloc(&n->w_fwd, (size_t)kFwdBlocks * kFrag)) ||
        (e = alloc(&n->w_bwd, (size_t)kBwdBlocks * kFrag)) || (e = alloc(&n->w_fwd_h, (size_t)kFwdBlocks * kFrag)) ||
        (e = alloc(&n->w_fwd_8, (size_t)kFwdBlocks * kFrag)) || (e = alloc(&n->scales, 32)) ||
        (e = alloc(&n->act_max, 8)) || (e = alloc(&n->loss, 1))) {
        rgi_nrc_destroy(n);
        return e;
    }
    (void)hipMemset(n->table_m, 0, n->table_floats * 4);
    (void)hipMemset(n->table_v, 0, n->table_floats * 4);
    (void)hipMemset(n->table_grad, 0, n->table_floats * 4);
    (void)hipMemset(n->mlp_m, 0, RGI_NRC_MLP_PARAMS * 4);
    (void)hipMemset(n->mlp_v, 0, RGI_NRC_MLP_PARAMS * 4);
    (void)hipMemset(n->mlp_grad, 0, RGI_NRC_MLP_PARAMS * 4);
    (void)hipMemset(n->act_max, 0, 32);
    (void)hipMemset(n->scales, 0, 128);
    size_t entries_total = n->table_floats / 2;
    hipLaunchKernelGGL(k_nrc_init_table, dim3((unsigned)((entries_total + 255) / 256)), dim3(256), 0, 0, n->table,
                       n->table_h, config->seed, entries_total);
    hipLaunchKernelGGL(k_nrc_init_mlp, dim3((RGI_NRC_MLP_PARAMS + 255) / 256), dim3(256), 0, 0, n->mlp, config->seed);
    if ((e = refresh_copies(n, 0, 0)) || (e = check_launch("nrc init"))) {
        rgi_nrc_destroy(n);
        return e;
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        rgi_nrc_destroy(n);
        return fail(RGI_E_DEVICE, "nrc_create: initialisation failed");
    }
    *out = n;
    return RGI_OK;
}

void rgi_nrc_destroy(rgi_nrc* n)
{
    if (!n) return;
    (void)hipDeviceSynchronize();
    void* ptrs[] = {n->table, n->table_m, n->table_v, n->table_grad, n->table_h, n->mlp, n->mlp_m, n->mlp_v,
                    n->mlp_grad, n->w_fwd, n->w_bwd, n->w_fwd_h, n->w_fwd_8, n->scales, n->act_max, n->loss,
                    n->acts, n->deltas, n->partial};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete n;
}

int rgi_nrc_train(rgi_nrc* n, const float* records, const float* targets, uint32_t count, float* loss, void* stream)
{
    if (!n || !records || !targets || count == 0) return fail(RGI_E_INVALID, "nrc_train: bad argument");
    hipStream_t s = as_stream(stream);
    int e = gradients(n, records, targets, count, s);
    if (e) return e;
    n->steps += 1;
    if ((e = refresh_copies(n, 1, s))) return e;
    AdamArgs h = adam_args(*n, n->steps);
    size_t quads = n->table_floats / 4;
    hipLaunchKernelGGL(k_nrc_adam_table, dim3((unsigned)((quads + 255) / 256)), dim3(256), 0, s, n->table, n->table_m,
                       n->table_v, n->table_grad, n->table_h, h, quads);
    if ((e = check_launch("nrc adam (table)"))) return e;
    if (loss) hipLaunchKernelGGL(k_nrc_copy_loss, dim3(1), dim3(64), 0, s, n->loss, loss);
    return check_launch("nrc train");
}

int rgi_nrc_gradients(rgi_nrc* n, const float* records, const float* targets, uint32_t count, float* mlp_grad,
                      float* table_grad, float* loss, void* stream)
{
    if (!n || !records || !targets || count == 0 || !mlp_grad)
        return fail(RGI_E_INVALID, "nrc_gradients: bad argument");
    hipStream_t s = as_stream(stream);
    int e = gradients(n, records, targets, count, s);
    if (e) return e;
    if (hipMemcpyAsync(mlp_grad, n->mlp_grad, RGI_NRC_MLP_PARAMS * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(RGI_E_DEVICE, "nrc_gradients: copy failed");
    hipLaunchKernelGGL(k_nrc_take_grad, dim3((unsigned)((n->table_floats + 255) / 256)), dim3(256), 0, s,
                       n->table_grad, table_grad, n->table_floats);
    if ((e = check_launch("nrc gradients"))) return e;
    if (loss) hipLaunchKernelGGL(k_nrc_copy_loss, dim3(1), dim3(64), 0, s, n->loss, loss);
    return check_launch("nrc gradients");
}

int rgi_nrc_infer(rgi_nrc* n, const float* records, uint32_t count, int precision, float* out, void* stream)
{
    if (!n || !records || !out) return fail(RGI_E_INVALID, "nrc_infer: bad argument");
    if (count == 0) return RGI_OK;
    hipStream_t s = as_stream(stream);
    RecordSource src{records, count, out};
    Grid gp = grid_of(*n);
    dim3 blocks((count + 255) / 256);
    switch (precision) {
    case RGI_NRC_BF16: hipLaunchKernelGGL(k_nrc_infer<RGI_NRC_BF16>, blocks, dim3(256), 0, s, *n, gp, src); break;
    case RGI_NRC_FP16: hipLaunchKernelGGL(k_nrc_infer<RGI_NRC_FP16>, blocks, dim3(256), 0, s, *n, gp, src); break;
    case RGI_NRC_FP8: hipLaunchKernelGGL(k_nrc_infer<RGI_NRC_FP8>, blocks, dim3(256), 0, s, *n, gp, src); break;
    default: return fail(RGI_E_INVALID, "nrc_infer: unknown precision %d", precision);
    }
    return check_launch("nrc infer");
}

int rgi_nrc_get_params(const rgi_nrc* n, float* mlp, float* table, void* stream)
{
    if (!n) return fail(RGI_E_INVALID, "nrc_get_params: nrc is NULL");
    hipStream_t s = as_stream(stream);
    if (mlp && hipMemcpyAsync(mlp, n->mlp, RGI_NRC_MLP_PARAMS * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(RGI_E_DEVICE, "nrc_get_params: copy failed");
    if (table && hipMemcpyAsync(table, n->table, n->table_floats * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(RGI_E_DEVICE, "nrc_get_params: copy failed");
    return RGI_OK;
}

int rgi_nrc_set_params(rgi_nrc* n, const float* mlp, const float* table, void* stream)
{
    if (!n) return fail(RGI_E_INVALID, "nrc_set_params: nrc is NULL");
    hipStream_t s = as_stream(stream);
    if (mlp && hipMemcpyAsync(n->mlp, mlp, RGI_NRC_MLP_PARAMS * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(RGI_E_DEVICE, "nrc_set_params: copy failed");
    if (table) {
        if (hipMemcpyAsync(n->table, table, n->table_floats * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return fail(RGI_E_DEVICE, "nrc_set_params: copy failed");
        size_t entries_total = n->table_floats / 2;
        hipLaunchKernelGGL(k_nrc_half_table, dim3((unsigned)((entries_total + 255) / 256)), dim3(256), 0, s, n->table,
                           n->table_h, entries_total);
    }
    (void)hipMemsetAsync(n->table_m, 0, n->table_floats * 4, s);
    (void)hipMemsetAsync(n->table_v, 0, n->table_floats * 4, s);
    (void)hipMemsetAsync(n->mlp_m, 0, RGI_NRC_MLP_PARAMS * 4, s);
    (void)hipMemsetAsync(n->mlp_v, 0, RGI_NRC_MLP_PARAMS * 4, s);
    n->steps = 0;
    return refresh_copies(n, 0, s);
}

int rgi_nrc_steps(const rgi_nrc* n, uint32_t* steps)
{
    if (!n || !steps) return fail(RGI_E_INVALID, "nrc_steps: bad argument");
    *steps = n->steps;
    return RGI_OK;
}

} // extern "C"
