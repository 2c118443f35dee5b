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
#include "batch.h"
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
    uint32_t* levels = nullptr;   // res[16], stride[16] (res + 1 for a densely indexed level, else 0)
    // Moving average of the parameters and its device copies: what inference
    // reads when config.ema_decay > 0 (training reads the parameters).
    float* ema_table = nullptr;
    uint32_t* ema_table_h = nullptr;
    float* ema_mlp = nullptr;
    __bf16* ema_w_fwd = nullptr;
    __bf16* ema_w_bwd = nullptr;
    _Float16* ema_w_fwd_h = nullptr;
    uint8_t* ema_w_fwd_8 = nullptr;
    float* ema_scales = nullptr;
    float* loss = nullptr;        // internal loss accumulator
    // Training scratch, grown on demand.
    __bf16* acts = nullptr;       // [6][tiles][64][256]
    __bf16* deltas = nullptr;     // [6][tiles][64][256]
    float* partial = nullptr;     // [chunks][RGI_NRC_MLP_PARAMS]
    uint32_t capacity_tiles = 0;
    // Gather scratch, grown on demand: surface pixel counts per 1024-pixel
    // block, their exclusive offsets (the total last) and the surface list.
    uint32_t* surface_counts = nullptr;
    uint32_t* surface_offsets = nullptr;
    uint32_t* surface_list = nullptr;
    uint32_t capacity_pixels = 0;
};

namespace rgi {
namespace nrc {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kTrainB = 2;              // 16-sample column blocks per wave in training
// The ReLU masks hold 16 bits per column block in 64 bits, so B <= 4 fits;
// only B = 2 has run on the GPU against the oracle's gradients, and the
// training kernels are built for that value alone.
static_assert(kTrainB == 2, "k_nrc_train is validated at two 16-sample column blocks per wave");
constexpr int kTile = 64 * kTrainB;     // samples per training workgroup (4 waves)
constexpr int kChunk = 1024;            // samples per weight-gradient workgroup
constexpr int kChunkTiles = kChunk / kTile;
constexpr int kInferB = 1;              // column blocks per wave in inference
constexpr int kFwdBlocks = 42;          // fragments (64 lanes x 8) of the forward copy
constexpr int kBwdBlocks = 40;          // fragments of the transposed copy
constexpr int kFrag = 512;
constexpr uint32_t kPrime1 = 2654435761u;
constexpr uint32_t kPrime2 = 805459861u;

struct Grid {
    const uint32_t* levels; // device: res[16], then stride[16] (res + 1 when dense, else 0)
    uint32_t mask;          // entries per level - 1
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

// The four levels a lane encodes: lane group g owns levels 4g .. 4g+3, which
// are input features 8g .. 8g+7 (slot q of the first K fragment is feature
// 8g + q).
struct LaneLevels {
    float rf[4];
    int rmax[4];
    uint32_t stride[4];
    uint32_t offset[4]; // first entry of the level
};

RGI_DEV LaneLevels lane_levels(const Grid& gp, int g)
{
    LaneLevels L;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        uint32_t r = gp.levels[4 * g + m];
        L.rf[m] = (float)r;
        L.rmax[m] = (int)r - 1;
        L.stride[m] = gp.levels[16 + 4 * g + m];
        L.offset[m] = (uint32_t)(4 * g + m) * gp.entries;
    }
    return L;
}

// The cell of level m around p and the fractional position in it.
RGI_DEV void level_cell(const LaneLevels& L, int m, const float* p, uint32_t* c, float* w)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float f = p[a] * L.rf[m];
        int fl = min((int)floorf(f), L.rmax[m]);
        c[a] = (uint32_t)fl;
        w[a] = f - (float)fl;
    }
}

// Entry (within its level) and trilinear weight of corner k = x + 2y + 4z.
RGI_DEV void corner_of(uint32_t stride, uint32_t mask, const uint32_t* c, const float* w, int k, uint32_t& idx,
                       float& weight)
{
    uint32_t x = c[0] + (uint32_t)(k & 1), y = c[1] + (uint32_t)((k >> 1) & 1), z = c[2] + (uint32_t)(k >> 2);
    weight = (((k & 1) ? w[0] : 1.0f - w[0]) * ((k & 2) ? w[1] : 1.0f - w[1])) * ((k & 4) ? w[2] : 1.0f - w[2]);
    idx = stride ? x + y * stride + z * (stride * stride) : ((x ^ (y * kPrime1) ^ (z * kPrime2)) & mask);
}

// Features of this lane's four levels (eight values, level-major): FP32 sums
// of FP16 table values in corner order.
RGI_DEV void encode_lane(const LaneLevels& L, uint32_t mask, const uint32_t* table_h, const float* p, float* e)
{
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        uint32_t c[3];
        float w[3];
        level_cell(L, m, p, c, w);
        uint32_t v[8];
        float wt[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint32_t idx;
            corner_of(L.stride[m], mask, c, w, k, idx, wt[k]);
            v[k] = table_h[L.offset[m] + idx];
        }
        float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a0 = a0 + wt[k] * half_lo(v[k]);
            a1 = a1 + wt[k] * half_hi(v[k]);
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

// x clamped to [lo, hi], NaN to lo.
RGI_DEV float bounded(float x, float lo, float hi) { return x > lo ? (x < hi ? x : hi) : lo; }

// Every input is finite: a NaN anywhere would reach the weight gradients
// through products with zero back-propagated gradients.
RGI_DEV Point point_of(const Grid& gp, const float* rec)
{
    Point pt;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        pt.p[c] = unit((rec[c] - gp.lo[c]) * gp.inv_extent[c]);
        pt.n[c] = bounded(rec[4 + c], -1.0f, 1.0f);
        pt.a[c] = bounded(rec[8 + c], 0.0f, 1.0f);
    }
    return pt;
}

RGI_DEV Point load_point(const Grid& gp, const float* records, uint32_t s, uint32_t count)
{
    float rec[12];
    if (s < count) {
        const float4* r = reinterpret_cast<const float4*>(records) + (size_t)s * 3;
        float4 a = r[0], b = r[1], c = r[2];
        rec[0] = a.x; rec[1] = a.y; rec[2] = a.z; rec[3] = a.w;
        rec[4] = b.x; rec[5] = b.y; rec[6] = b.z; rec[7] = b.w;
        rec[8] = c.x; rec[9] = c.y; rec[10] = c.z; rec[11] = c.w;
    } else {
#pragma unroll
        for (int i = 0; i < 12; ++i) rec[i] = 0.0f;
    }
    return point_of(gp, rec);
}

// Inputs 32 .. 39, held by lane group 0: normal * 0.5 + 0.5, albedo, 1, 0.
RGI_DEV void extras(const Point& pt, int g, float* e)
{
    if (g == 0) {
        e[0] = pt.n[0] * 0.5f + 0.5f;
        e[1] = pt.n[1] * 0.5f + 0.5f;
        e[2] = pt.n[2] * 0.5f + 0.5f;
        e[3] = pt.a[0];
        e[4] = pt.a[1];
        e[5] = pt.a[2];
        e[6] = 1.0f;
        e[7] = 0.0f;
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) e[q] = 0.0f;
    }
}

RGI_DEV bf16x8 pack_bf16(const float* e)
{
    bf16x8 v;
#pragma unroll
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

// [feature][sample] store of one fragment's four rows for this lane's sample.
RGI_DEV void store_rows(__bf16* plane, int t, int g, int sample, const float* v)
{
#pragma unroll
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
    __bf16* acts;   // [6][tiles][64][kTile]
    __bf16* deltas; // [6][tiles][64][kTile]
    uint32_t tiles;
    float* loss;
    uint32_t* act_max;
};

// Forward, loss and backward of 16 * kTrainB samples per wave: stores the
// activations and back-propagated gradients for the weight-gradient GEMMs
// and scatters the input gradient into the table gradient.
RGI_DEV void nrc_train_body(const TrainArgs& a, const uint32_t tile)
{
    constexpr int B = kTrainB;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u), g = lane >> 4, s16 = lane & 15;
    const size_t layer_stride = (size_t)a.tiles * 64 * kTile;
    const size_t tile_base = (size_t)tile * 64 * kTile;
    const LaneLevels L = lane_levels(a.gp, g);
    bf16x8 X[2][B];
    Point pts[B];
    float amax = 0.0f;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int local = wave * 16 * B + b * 16 + s16;
        pts[b] = load_point(a.gp, a.records, tile * kTile + (uint32_t)local, a.count);
        float e[8], x[8];
        encode_lane(L, a.gp.mask, a.table_h, pts[b].p, e);
        extras(pts[b], g, x);
        X[0][b] = pack_bf16(e);
        X[1][b] = pack_bf16(x);
        __bf16* a0 = a.acts + tile_base;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            a0[(size_t)(8 * g + q) * kTile + local] = X[0][b][q];
            a0[(size_t)(32 + 8 * g + q) * kTile + local] = X[1][b][q];
            amax = fmaxf(amax, fmaxf(fabsf((float)X[0][b][q]), fabsf((float)X[1][b][q])));
        }
    }
    note_max(a.act_max + 0, amax);

    // Hidden layers 1..5.
    uint64_t masks[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        f32x4 acc[B][4];
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[b][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bf16x8 w = frag(a.w_fwd, k * 8 + t * 2 + j, lane);
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b][t] = mfma_bf16(w, X[j][b], acc[b][t]);
            }
        uint64_t mask = 0;
        __bf16* plane = a.acts + (size_t)(k + 1) * layer_stride + tile_base;
        amax = 0.0f;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int local = wave * 16 * B + b * 16 + s16;
            float v[4][4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float z = acc[b][t][r];
                    bool on = z > 0.0f;
                    mask |= (uint64_t)on << (b * 16 + t * 4 + r);
                    v[t][r] = on ? z : 0.0f;
                }
#pragma unroll
            for (int t = 0; t < 4; ++t) store_rows(plane, t, g, local, v[t]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float e[8];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    e[r] = v[2 * j][r];
                    e[4 + r] = v[2 * j + 1][r];
                }
                X[j][b] = pack_bf16(e);
#pragma unroll
                for (int q = 0; q < 8; ++q) amax = fmaxf(amax, (float)X[j][b][q]);
            }
        }
        masks[k] = mask;
        note_max(a.act_max + k + 1, amax);
    }

    // Output layer, loss and its gradient (lane group 0 holds rows 0..3).
    f32x4 y[B];
#pragma unroll
    for (int b = 0; b < B; ++b) y[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        bf16x8 w = frag(a.w_fwd, 40 + j, lane);
#pragma unroll
        for (int b = 0; b < B; ++b) y[b] = mfma_bf16(w, X[j][b], y[b]);
    }
    float loss = 0.0f;
    bf16x8 D[2][B];
    __bf16* d6 = a.deltas + (size_t)5 * layer_stride + tile_base;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int local = wave * 16 * B + b * 16 + s16;
        const uint32_t s = tile * kTile + (uint32_t)local;
        float d[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (g == 0 && s < a.count) {
            float4 t = reinterpret_cast<const float4*>(a.targets)[s];
            if (t.w > 0.0f) {
                float mean = ((y[b][0] + y[b][1]) + y[b][2]) / 3.0f;
                float denom = mean * mean + 0.01f;
                float e0 = y[b][0] - t.x, e1 = y[b][1] - t.y, e2 = y[b][2] - t.z;
                loss = loss + (t.w * ((e0 * e0 + e1 * e1) + e2 * e2)) / denom;
                float scale = ((2.0f * t.w) * a.inv_count) / denom;
                d[0] = e0 * scale;
                d[1] = e1 * scale;
                d[2] = e2 * scale;
            }
        }
        store_rows(d6, 0, g, local, d);
        float e[8] = {d[0], d[1], d[2], d[3], 0.0f, 0.0f, 0.0f, 0.0f};
        D[0][b] = pack_bf16(e);
        D[1][b] = pack_bf16(e); // unused for the output layer (K = 16 rows)
    }
    loss = wave_sum(loss);
    if (lane == 0 && loss != 0.0f) atomicAdd(a.loss, loss * a.inv_count);

    // Back through W6 (K = the 16 output rows, slots q < 4 only), then W5..W2:
    // after step k, D holds the gradient of hidden layer k's pre-activation.
#pragma unroll
    for (int k = 5; k >= 1; --k) {
        f32x4 acc[B][4];
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[b][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if (k == 5) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bf16x8 w = frag(a.w_bwd, t, lane);
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b][t] = mfma_bf16(w, D[0][b], acc[b][t]);
            }
        } else {
            const int base = 4 + (4 - k) * 8; // fragments of W_{k+1}^T
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    bf16x8 w = frag(a.w_bwd, base + t * 2 + j, lane);
#pragma unroll
                    for (int b = 0; b < B; ++b) acc[b][t] = mfma_bf16(w, D[j][b], acc[b][t]);
                }
        }
        __bf16* plane = a.deltas + (size_t)(k - 1) * layer_stride + tile_base;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int local = wave * 16 * B + b * 16 + s16;
            float v[4][4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    v[t][r] = ((masks[k - 1] >> (b * 16 + t * 4 + r)) & 1ull) ? acc[b][t][r] : 0.0f;
#pragma unroll
            for (int t = 0; t < 4; ++t) store_rows(plane, t, g, local, v[t]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float e[8];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    e[r] = v[2 * j][r];
                    e[4 + r] = v[2 * j + 1][r];
                }
                D[j][b] = pack_bf16(e);
            }
        }
    }

    // Input gradient of the encoded features, W1^T times D. The transposed
    // copy orders its rows so that output row 16t + 4g + r is input feature
    // 8g + 4t + r: this lane receives the gradients of its own four levels
    // (level 4g + 2t + r/2, feature r%2).
#pragma unroll
    for (int b = 0; b < B; ++b) {
        f32x4 dx[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 2; ++t) dx[t] = mfma_bf16(frag(a.w_bwd, 36 + t * 2 + j, lane), D[j][b], dx[t]);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            float d0 = dx[m >> 1][2 * (m & 1)], d1 = dx[m >> 1][2 * (m & 1) + 1];
            if (d0 == 0.0f && d1 == 0.0f) continue;
            uint32_t c[3];
            float w[3];
            level_cell(L, m, pts[b].p, c, w);
            float* level = a.table_grad + (size_t)L.offset[m] * 2;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                uint32_t idx;
                float wt;
                corner_of(L.stride[m], a.gp.mask, c, w, k, idx, wt);
                if (d0 != 0.0f) unsafeAtomicAdd(level + 2 * idx, wt * d0);
                if (d1 != 0.0f) unsafeAtomicAdd(level + 2 * idx + 1, wt * d1);
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_nrc_train(TrainArgs a) { nrc_train_body(a, blockIdx.x); }

// Weight gradients: dW_k[o][i] = sum over samples of delta_k[o][s] a_{k-1}[i][s]
// over one chunk of kChunk samples; one workgroup per (chunk, layer). For
// W1..W5, wave w owns output rows 16w..16w+15; for W6, column block w of
// rows 0..15.
RGI_DEV void nrc_wgrad_body(const __bf16* acts, const __bf16* deltas, uint32_t tiles, float* partial,
                            const uint32_t chunk, const int layer /* 0..5 -> W1..W6 */)
{
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u), g = lane >> 4, s16 = lane & 15;
    const size_t layer_stride = (size_t)tiles * 64 * kTile;
    const __bf16* dl = deltas + (size_t)layer * layer_stride;
    const __bf16* al = acts + (size_t)layer * layer_stride;
    float* out = partial + (size_t)chunk * RGI_NRC_MLP_PARAMS + (size_t)layer * 4096;
    if (layer < 5) {
        f32x4 acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int tt = 0; tt < kChunkTiles; ++tt) {
            uint32_t tile = chunk * kChunkTiles + (uint32_t)tt;
            if (tile >= tiles) break;
            const __bf16* dt = dl + (size_t)tile * 64 * kTile;
            const __bf16* at = al + (size_t)tile * 64 * kTile;
#pragma unroll
            for (int kb = 0; kb < kTile / 32; ++kb) {
                int s0 = kb * 32 + 8 * g;
                bf16x8 A = *reinterpret_cast<const bf16x8*>(dt + (size_t)(16 * wave + s16) * kTile + s0);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    bf16x8 Bv = *reinterpret_cast<const bf16x8*>(at + (size_t)(16 * c + s16) * kTile + s0);
                    acc[c] = mfma_bf16(A, Bv, acc[c]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[(16 * wave + 4 * g + r) * 64 + 16 * c + s16] = acc[c][r];
    } else {
        f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int tt = 0; tt < kChunkTiles; ++tt) {
            uint32_t tile = chunk * kChunkTiles + (uint32_t)tt;
            if (tile >= tiles) break;
            const __bf16* dt = dl + (size_t)tile * 64 * kTile;
            const __bf16* at = al + (size_t)tile * 64 * kTile;
#pragma unroll
            for (int kb = 0; kb < kTile / 32; ++kb) {
                int s0 = kb * 32 + 8 * g;
                bf16x8 A = *reinterpret_cast<const bf16x8*>(dt + (size_t)s16 * kTile + s0);
                bf16x8 Bv = *reinterpret_cast<const bf16x8*>(at + (size_t)(16 * wave + s16) * kTile + s0);
                acc = mfma_bf16(A, Bv, acc);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int o = 4 * g + r;
            if (o < 3) out[o * 64 + 16 * wave + s16] = acc[r];
        }
    }
}

__global__ void __launch_bounds__(256) k_nrc_wgrad(const __bf16* acts, const __bf16* deltas, uint32_t tiles,
                                                   float* partial)
{
    nrc_wgrad_body(acts, deltas, tiles, partial, blockIdx.x, (int)blockIdx.y);
}

// Sum of the chunks' partial gradients, in chunk order.
RGI_DEV void nrc_reduce_body(const float* partial, uint32_t chunks, float* grad, uint32_t block)
{
    uint32_t p = block * blockDim.x + threadIdx.x;
    if (p >= RGI_NRC_MLP_PARAMS) return;
    float s = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) s = s + partial[(size_t)c * RGI_NRC_MLP_PARAMS + p];
    grad[p] = s;
}

__global__ void k_nrc_reduce(const float* partial, uint32_t chunks, float* grad)
{
    nrc_reduce_body(partial, chunks, grad, blockIdx.x);
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

// The same for the transposed copy: W6^T (K = output row 4g + q, q < 4),
// W5^T .. W2^T, then W1^T with row 16t + i holding input feature
// 8(i / 4) + 4t + i % 4.
RGI_DEV float bwd_weight(const float* mlp, int block, int lane, int q)
{
    int g = lane >> 4, s16 = lane & 15;
    if (block < 4) {
        int o = 4 * g + q;
        return (q < 4 && o < 3) ? mlp[20480 + o * 64 + 16 * block + s16] : 0.0f;
    }
    if (block < 36) {
        int kk = (block - 4) / 8, t = ((block - 4) % 8) / 2, j = (block - 4) % 2;
        int k = 5 - kk; // 1-based layer
        return mlp[(k - 1) * 4096 + phi(j, g, q) * 64 + 16 * t + s16];
    }
    int t = (block - 36) / 2, j = (block - 36) % 2;
    int feature = 8 * (s16 >> 2) + 4 * t + (s16 & 3);
    return mlp[phi(j, g, q) * 64 + feature];
}

RGI_DEV int fwd_layer(int block) { return block < 40 ? block / 8 : 5; }
RGI_DEV int bwd_layer(int block) { return block < 4 ? 5 : (block < 36 ? 4 - (block - 4) / 8 : 0); }

RGI_DEV uint8_t to_fp8(float x)
{
    x = fminf(fmaxf(x, -448.0f), 448.0f);
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xff);
}

// One workgroup per weight matrix: Adam on its parameters (when `step`), its
// largest magnitude, and its FP8 weight and activation scales.
RGI_DEV void nrc_adam_mlp_body(float* mlp, float* m, float* v, const float* grad, const AdamArgs& h, float* scales,
                               const uint32_t* act_max, int step, const int layer, float* red)
{
    const int lo = layer * 4096, hi = layer == 5 ? 20480 + 192 : lo + 4096;
    float wmax = 0.0f;
    for (int p = lo + (int)threadIdx.x; p < hi; p += 256) {
        float w = mlp[p];
        if (step) {
            float mm = m[p], vv = v[p];
            w = adam(h, grad[p], mm, vv, w);
            m[p] = mm;
            v[p] = vv;
            mlp[p] = w;
        }
        wmax = fmaxf(wmax, fabsf(w));
    }
    wmax = wave_max(wmax);
    if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = wmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        wmax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        float sw = wmax > 0.0f ? 448.0f / wmax : 1.0f;
        float amax = __uint_as_float(act_max[layer]);
        float sa = amax > 0.0f ? 224.0f / amax : 1.0f;
        scales[layer] = sw;
        scales[8 + layer] = sa;
        scales[16 + layer] = 1.0f / (sa * sw);
    }
}

__global__ void __launch_bounds__(256) k_nrc_adam_mlp(float* mlp, float* m, float* v, const float* grad, AdamArgs h,
                                                      float* scales, const uint32_t* act_max, int step)
{
    __shared__ float red[4];
    nrc_adam_mlp_body(mlp, m, v, grad, h, scales, act_max, step, (int)blockIdx.x, red);
}

// Every fragment slot of the forward (BF16, FP16, FP8) and transposed (BF16)
// copies, one per thread.
RGI_DEV void nrc_copies_body(const float* mlp, const float* scales, __bf16* w_fwd, __bf16* w_bwd, _Float16* w_fwd_h,
                             uint8_t* w_fwd_8, uint32_t block)
{
    int e = (int)(block * blockDim.x + threadIdx.x);
    if (e < kFwdBlocks * kFrag) {
        int block = e / kFrag, lane = (e % kFrag) / 8, q = e % 8;
        float w = fwd_weight(mlp, block, lane, q);
        w_fwd[e] = (__bf16)w;
        w_fwd_h[e] = (_Float16)w;
        w_fwd_8[e] = to_fp8(w * scales[fwd_layer(block)]);
        return;
    }
    e -= kFwdBlocks * kFrag;
    if (e < kBwdBlocks * kFrag) {
        int block = e / kFrag, lane = (e % kFrag) / 8, q = e % 8;
        w_bwd[e] = (__bf16)bwd_weight(mlp, block, lane, q);
    }
}

__global__ void k_nrc_copies(const float* mlp, const float* scales, __bf16* w_fwd, __bf16* w_bwd, _Float16* w_fwd_h,
                             uint8_t* w_fwd_8)
{
    nrc_copies_body(mlp, scales, w_fwd, w_bwd, w_fwd_h, w_fwd_8, blockIdx.x);
}

// The parameters' moving average after a step: e = d e + (1 - d) w, one
// workgroup per weight matrix, with the FP8 weight scale of the averaged
// matrix (the activation scales are the trained ones).
RGI_DEV void nrc_ema_mlp_body(const float* mlp, float* ema, float d, float one_minus_d, const float* scales,
                              float* ema_scales, const int layer, float* red)
{
    const int lo = layer * 4096, hi = layer == 5 ? 20480 + 192 : lo + 4096;
    float wmax = 0.0f;
    for (int p = lo + (int)threadIdx.x; p < hi; p += 256) {
        float e = d * ema[p] + one_minus_d * mlp[p];
        ema[p] = e;
        wmax = fmaxf(wmax, fabsf(e));
    }
    wmax = wave_max(wmax);
    if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = wmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        wmax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        float sw = wmax > 0.0f ? 448.0f / wmax : 1.0f;
        float sa = scales[8 + layer];
        ema_scales[layer] = sw;
        ema_scales[8 + layer] = sa;
        ema_scales[16 + layer] = 1.0f / (sa * sw);
    }
}

__global__ void __launch_bounds__(256) k_nrc_ema_mlp(const float* mlp, float* ema, float d, float one_minus_d,
                                                     const float* scales, float* ema_scales)
{
    __shared__ float red[4];
    nrc_ema_mlp_body(mlp, ema, d, one_minus_d, scales, ema_scales, (int)blockIdx.x, red);
}

// The table's moving average, four values per thread, and its FP16 copy.
RGI_DEV void nrc_ema_table_body(const float* table, float* ema, uint32_t* ema_h, float d, float one_minus_d,
                                size_t quads, uint32_t block)
{
    size_t i = (size_t)block * blockDim.x + threadIdx.x;
    if (i >= quads) return;
    float4 w = reinterpret_cast<const float4*>(table)[i];
    float4 e = reinterpret_cast<float4*>(ema)[i];
    e = make_float4(d * e.x + one_minus_d * w.x, d * e.y + one_minus_d * w.y, d * e.z + one_minus_d * w.z,
                    d * e.w + one_minus_d * w.w);
    reinterpret_cast<float4*>(ema)[i] = e;
    reinterpret_cast<uint2*>(ema_h)[i] = make_uint2((uint32_t)float_to_half(e.x) | ((uint32_t)float_to_half(e.y) << 16),
                                                    (uint32_t)float_to_half(e.z) | ((uint32_t)float_to_half(e.w) << 16));
}

__global__ void k_nrc_ema_table(const float* table, float* ema, uint32_t* ema_h, float d, float one_minus_d,
                                size_t quads)
{
    nrc_ema_table_body(table, ema, ema_h, d, one_minus_d, quads, blockIdx.x);
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
RGI_DEV void nrc_adam_table_body(float* table, float* m, float* v, float* grad, uint32_t* table_h, const AdamArgs& h,
                                 size_t quads, uint32_t block)
{
    size_t i = (size_t)block * blockDim.x + threadIdx.x;
    if (i >= quads) return;
    float4 gq = reinterpret_cast<const float4*>(grad)[i];
    if (gq.x == 0.0f && gq.y == 0.0f && gq.z == 0.0f && gq.w == 0.0f) return;
    float4 pq = reinterpret_cast<float4*>(table)[i];
    float4 mq = reinterpret_cast<float4*>(m)[i];
    float4 vq = reinterpret_cast<float4*>(v)[i];
    float gs[4] = {gq.x, gq.y, gq.z, gq.w}, ps[4] = {pq.x, pq.y, pq.z, pq.w}, ms[4] = {mq.x, mq.y, mq.z, mq.w},
          vs[4] = {vq.x, vq.y, vq.z, vq.w};
#pragma unroll
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

__global__ void k_nrc_adam_table(float* table, float* m, float* v, float* grad, uint32_t* table_h, AdamArgs h,
                                 size_t quads)
{
    nrc_adam_table_body(table, m, v, grad, table_h, h, quads, blockIdx.x);
}

// ------------------------------------------------------- batched training

// One cache's training step in a batch of caches: the arguments of every
// kernel of rgi_nrc_train's chain. Each kernel of the chain runs once for
// the whole batch, each cache taking its own slice of the workgroups, so the
// chain's many small launches fill the chip together.
struct NrcStep {
    TrainArgs a;
    uint32_t active;  // the cache trains in this batch
    uint32_t chunks;
    uint32_t ema;     // a moving average is kept
    float d, om;      // its decay this step
    float* partial;
    float* mlp;
    float* mlp_m;
    float* mlp_v;
    float* mlp_grad;
    AdamArgs h;
    float* scales;
    __bf16* w_fwd;
    __bf16* w_bwd;
    _Float16* w_fwd_h;
    uint8_t* w_fwd_8;
    float* table;
    float* table_m;
    float* table_v;
    float* table_grad;
    uint32_t* table_h;
    size_t quads;
    float* ema_mlp;
    float* ema_scales;
    __bf16* ema_w_fwd;
    __bf16* ema_w_bwd;
    _Float16* ema_w_fwd_h;
    uint8_t* ema_w_fwd_8;
    float* ema_table;
    uint32_t* ema_table_h;
    float* loss_out;  // where the step's loss goes, or NULL
};
static_assert(sizeof(NrcStep) <= kNrcStepBytes, "a cache's step must fit its slot of a batch's table");

// Step i of a table: the steps lie kNrcStepBytes apart (not sizeof(NrcStep)).
RGI_DEV const NrcStep& step_at(const uint8_t* steps, uint32_t i)
{
    return *reinterpret_cast<const NrcStep*>(steps + (size_t)i * kNrcStepBytes);
}

#define RGI_NRC_STEP(steps, first, count)                                                                              \
    uint32_t b;                                                                                                        \
    const NrcStep& st = step_at(steps, batch_view(first, count, blockIdx.x, b))

__global__ void k_nrc_clear_max_batch(const uint8_t* __restrict__ steps)
{
    const NrcStep& st = step_at(steps, blockIdx.x);
    if (!st.active) return;
    if (threadIdx.x < 6) st.a.act_max[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *st.a.loss = 0.0f;
}

__global__ void __launch_bounds__(256) k_nrc_train_batch(const uint8_t* __restrict__ steps,
                                                         const uint32_t* __restrict__ first, uint32_t count)
{
    RGI_NRC_STEP(steps, first, count);
    nrc_train_body(st.a, b);
}

__global__ void __launch_bounds__(256) k_nrc_wgrad_batch(const uint8_t* __restrict__ steps,
                                                         const uint32_t* __restrict__ first, uint32_t count)
{
    RGI_NRC_STEP(steps, first, count);
    nrc_wgrad_body(st.a.acts, st.a.deltas, st.a.tiles, st.partial, b % st.chunks, (int)(b / st.chunks));
}

__global__ void k_nrc_reduce_batch(const uint8_t* __restrict__ steps, const uint32_t* __restrict__ first,
                                   uint32_t count)
{
    RGI_NRC_STEP(steps, first, count);
    nrc_reduce_body(st.partial, st.chunks, st.mlp_grad, b);
}

__global__ void __launch_bounds__(256) k_nrc_adam_mlp_batch(const uint8_t* __restrict__ steps,
                                                            const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ float red[4];
    RGI_NRC_STEP(steps, first, count);
    nrc_adam_mlp_body(st.mlp, st.mlp_m, st.mlp_v, st.mlp_grad, st.h, st.scales, st.a.act_max, 1, (int)b, red);
}

// The parameters' copies (ema = 0) or the moving average's (ema = 1).
template <int kEma>
__global__ void k_nrc_copies_batch(const uint8_t* __restrict__ steps, const uint32_t* __restrict__ first,
                                   uint32_t count)
{
    RGI_NRC_STEP(steps, first, count);
    if (kEma) nrc_copies_body(st.ema_mlp, st.ema_scales, st.ema_w_fwd, st.ema_w_bwd, st.ema_w_fwd_h, st.ema_w_fwd_8, b);
    else nrc_copies_body(st.mlp, st.scales, st.w_fwd, st.w_bwd, st.w_fwd_h, st.w_fwd_8, b);
}

__global__ void k_nrc_adam_table_batch(const uint8_t* __restrict__ steps, const uint32_t* __restrict__ first,
                                       uint32_t count)
{
    RGI_NRC_STEP(steps, first, count);
    nrc_adam_table_body(st.table, st.table_m, st.table_v, st.table_grad, st.table_h, st.h, st.quads, b);
}

__global__ void __launch_bounds__(256) k_nrc_ema_mlp_batch(const uint8_t* __restrict__ steps,
                                                           const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ float red[4];
    RGI_NRC_STEP(steps, first, count);
    nrc_ema_mlp_body(st.mlp, st.ema_mlp, st.d, st.om, st.scales, st.ema_scales, (int)b, red);
}

__global__ void k_nrc_ema_table_batch(const uint8_t* __restrict__ steps, const uint32_t* __restrict__ first,
                                      uint32_t count)
{
    RGI_NRC_STEP(steps, first, count);
    nrc_ema_table_body(st.table, st.ema_table, st.ema_table_h, st.d, st.om, st.quads, b);
}

__global__ void k_nrc_copy_loss_batch(const uint8_t* __restrict__ steps)
{
    const NrcStep& st = step_at(steps, blockIdx.x);
    if (st.active && st.loss_out && threadIdx.x == 0) *st.loss_out = *st.a.loss;
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
#pragma unroll
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
#pragma unroll
        for (int q = 0; q < 8; ++q) v |= (uint64_t)to_fp8(e[q] * scale) << (8 * q);
        return (long)v;
    }
    static RGI_DEV T weights(const rgi_nrc& n, int block, int lane)
    {
        return *reinterpret_cast<const long*>(n.w_fwd_8 + (size_t)block * kFrag + lane * 8);
    }
    static RGI_DEV f32x4 mma(T a, T b, f32x4 c) { return mfma_fp8(a, b, c); }
};

// The forward pass of 16 * B points per wave: this lane's points in, y[b]
// (rows 0..3 in lane group 0) out.
template <int P, int B>
RGI_DEV void forward(const rgi_nrc& n, const Grid& gp, const LaneLevels& L, const Point* pts, f32x4* y)
{
    typedef Operand<P> Op;
    const int lane = (int)(threadIdx.x & 63u), g = lane >> 4;
    typename Op::T X[2][B];
    const float sa = P == RGI_NRC_FP8 ? n.scales[8] : 1.0f;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        float e[8], x[8];
        encode_lane(L, gp.mask, n.table_h, pts[b].p, e);
        extras(pts[b], g, x);
        X[0][b] = Op::pack(e, sa);
        X[1][b] = Op::pack(x, sa);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        f32x4 acc[B][4];
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[b][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                typename Op::T w = Op::weights(n, k * 8 + t * 2 + j, lane);
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b][t] = Op::mma(w, X[j][b], acc[b][t]);
            }
        const float deq = P == RGI_NRC_FP8 ? n.scales[16 + k] : 1.0f;
        const float next = P == RGI_NRC_FP8 ? n.scales[8 + k + 1] : 1.0f;
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                float e[8];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float z0 = acc[b][2 * j][r] * deq, z1 = acc[b][2 * j + 1][r] * deq;
                    e[r] = z0 > 0.0f ? z0 : 0.0f;
                    e[4 + r] = z1 > 0.0f ? z1 : 0.0f;
                }
                X[j][b] = Op::pack(e, next);
            }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) y[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        typename Op::T w = Op::weights(n, 40 + j, lane);
#pragma unroll
        for (int b = 0; b < B; ++b) y[b] = Op::mma(w, X[j][b], y[b]);
    }
    if (P == RGI_NRC_FP8) {
        const float deq = n.scales[16 + 5];
#pragma unroll
        for (int b = 0; b < B; ++b) y[b] = y[b] * deq;
    }
}

template <int P>
__global__ void __launch_bounds__(256) k_nrc_infer(rgi_nrc n, Grid gp, const float* records, uint32_t count, float* out)
{
    constexpr int B = kInferB;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u), g = lane >> 4, s16 = lane & 15;
    const LaneLevels L = lane_levels(gp, g);
    const uint32_t base = blockIdx.x * (64u * B) + (uint32_t)wave * (16u * B);
    Point pts[B];
#pragma unroll
    for (int b = 0; b < B; ++b) pts[b] = load_point(gp, records, base + (uint32_t)(b * 16 + s16), count);
    f32x4 y[B];
    forward<P, B>(n, gp, L, pts, y);
    if (g != 0) return;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        uint32_t s = base + (uint32_t)(b * 16 + s16);
        if (s < count) reinterpret_cast<float4*>(out)[s] = make_float4(y[b][0], y[b][1], y[b][2], 0.0f);
    }
}

// A visible surface's inputs from the G-buffer: the reconstructed position,
// the decoded surface normal and the diffuse albedo, albedo * (1 - metallic),
// as the GI passes shade it. False for the sky.
RGI_DEV bool gbuffer_point(const Grid& gp, const rgi_frame& f, const rgi_gbuffer& g, F2 texel, int x, int y,
                           Point& pt)
{
    uint32_t t = (uint32_t)y * (uint32_t)g.width + (uint32_t)x;
    float depth = g.depth[t];
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    F3 pos = reconstruct_position(f, uv, depth);
    F4 nm = unorm8(g.normal[t]);
    F3 n = normalize(F3{nm.x * 2.0f - 1.0f, nm.y * 2.0f - 1.0f, nm.z * 2.0f - 1.0f});
    F4 al = unorm8(g.albedo[t]);
    bool surface = depth < kSkyDepth;
    float rec[12] = {pos.x, pos.y, pos.z, 0.0f, n.x, n.y, n.z, 0.0f, al.x * (1.0f - nm.w), al.y * (1.0f - nm.w),
                     al.z * (1.0f - nm.w), 0.0f};
    if (!surface)
        for (int i = 0; i < 12; ++i) rec[i] = i == 6 ? 1.0f : 0.0f;
    pt = point_of(gp, rec);
    return surface;
}

// Every pixel of the G-buffer: a wave takes an 8 x 2B pixel block
// (neighbouring pixels share the encoding's cells), rows 2b and 2b + 1 of
// the block in column block b; a workgroup 16 x 4B pixels. Writes the
// cache's radiance, clamped to >= 0, plus the G-buffer emission when
// `emission` is set, as RGBA16F (alpha 1); sky pixels get 0.
template <int P>
RGI_DEV void infer_gbuffer_body(const rgi_nrc& n, const Grid& gp, const rgi_frame& f, const rgi_gbuffer& gb,
                                F2 texel, const uint16_t* emission, uint16_t* out, const uint32_t bx, const uint32_t by)
{
    constexpr int B = kInferB;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u), g = lane >> 4, s16 = lane & 15;
    const LaneLevels L = lane_levels(gp, g);
    const int x = (int)(bx * 16u) + (wave & 1) * 8 + (s16 & 7);
    Point pts[B];
    bool surface[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        int y = (int)(by * (4u * B)) + (wave >> 1) * (2 * B) + 2 * b + (s16 >> 3);
        surface[b] = gbuffer_point(gp, f, gb, texel, min(x, gb.width - 1), min(y, gb.height - 1), pts[b]);
    }
    // A wave whose pixels are all sky writes zeros and leaves: the forward
    // pass is wave-local.
    bool any_surface = false;
#pragma unroll
    for (int b = 0; b < B; ++b) any_surface = any_surface || surface[b];
    if (__ballot(any_surface) == 0ull) {
        if (g != 0 || x >= gb.width) return;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            int y = (int)(by * (4u * B)) + (wave >> 1) * (2 * B) + 2 * b + (s16 >> 3);
            if (y < gb.height) store_h4(out, (uint32_t)y * (uint32_t)gb.width + (uint32_t)x, F4{0.0f, 0.0f, 0.0f, 0.0f});
        }
        return;
    }
    f32x4 yv[B];
    forward<P, B>(n, gp, L, pts, yv);
    if (g != 0 || x >= gb.width) return;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        int y = (int)(by * (4u * B)) + (wave >> 1) * (2 * B) + 2 * b + (s16 >> 3);
        if (y >= gb.height) continue;
        uint32_t o = (uint32_t)y * (uint32_t)gb.width + (uint32_t)x;
        F4 r = F4{0.0f, 0.0f, 0.0f, 0.0f};
        if (surface[b]) {
            r = F4{fmaxf(yv[b][0], 0.0f), fmaxf(yv[b][1], 0.0f), fmaxf(yv[b][2], 0.0f), 1.0f};
            if (emission) {
                F4 e = load_h4(emission, o);
                r = F4{r.x + e.x, r.y + e.y, r.z + e.z, 1.0f};
            }
        }
        store_h4(out, o, r);
    }
}

template <int P>
__global__ void __launch_bounds__(256) k_nrc_infer_gbuffer(rgi_nrc n, Grid gp, rgi_frame f, rgi_gbuffer gb,
                                                           F2 texel, const uint16_t* emission, uint16_t* out)
{
    infer_gbuffer_body<P>(n, gp, f, gb, texel, emission, out, blockIdx.x, blockIdx.y);
}

// One query of a batch of G-buffer queries (each its own cache, frame,
// G-buffer and output plane): the batched launch gives query q the
// workgroups [first[q], first[q + 1]), which cover its grid of bx columns.
struct NrcQuery {
    rgi_nrc n;       // the cache's inference view (its moving average's copies)
    Grid gp;
    rgi_frame f;
    rgi_gbuffer gb;
    F2 texel;
    const uint16_t* emission;
    uint16_t* out;
    uint32_t bx;
    int precision;
};
static_assert(sizeof(NrcQuery) <= kNrcQueryBytes, "a query must fit its slot of a batch's table");

template <int P>
__global__ void __launch_bounds__(256) k_nrc_infer_gbuffer_batch(const uint8_t* __restrict__ queries,
                                                                 const uint32_t* __restrict__ first, uint32_t count)
{
    uint32_t b;
    const NrcQuery& q =
        *reinterpret_cast<const NrcQuery*>(queries + (size_t)batch_view(first, count, blockIdx.x, b) * kNrcQueryBytes);
    if (q.precision != P) return; // the whole workgroup: the pass has no barrier
    infer_gbuffer_body<P>(q.n, q.gp, q.f, q.gb, q.texel, q.emission, q.out, b % q.bx, b / q.bx);
}

// The surface pixels of a G-buffer, in row-major order, as a list: a
// 256-thread block owns 1024 consecutive pixels (four chunks of 256), counts
// them, one workgroup turns the counts into offsets, and each block writes
// its surface pixels' indices at its offset.
constexpr uint32_t kSurfaceBlock = 1024;

RGI_DEV void surface_count_body(const float* depth, uint32_t npix, uint32_t* counts, uint32_t block,
                                uint32_t* per_wave)
{
    uint32_t base = block * kSurfaceBlock, c = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        uint32_t p = base + k * 256u + threadIdx.x;
        c += (uint32_t)__popcll(__ballot(p < npix && depth[p] < kSkyDepth));
    }
    if ((threadIdx.x & 63u) == 0) per_wave[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[block] = ((per_wave[0] + per_wave[1]) + per_wave[2]) + per_wave[3];
}

__global__ void __launch_bounds__(256) k_nrc_surface_count(const float* depth, uint32_t npix, uint32_t* counts)
{
    __shared__ uint32_t per_wave[4];
    surface_count_body(depth, npix, counts, blockIdx.x, per_wave);
}

RGI_DEV void surface_scan_body(const uint32_t* counts, uint32_t blocks, uint32_t* offsets, uint32_t* part)
{
    uint32_t per = (blocks + 1023u) / 1024u, lo = threadIdx.x * per, hi = min(lo + per, blocks), s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += counts[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t step = 1; step < 1024u; step <<= 1) {
        uint32_t v = threadIdx.x >= step ? part[threadIdx.x - step] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t i = lo; i < hi; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    if (threadIdx.x == 1023u) offsets[blocks] = part[1023];
}

__global__ void __launch_bounds__(1024) k_nrc_surface_scan(const uint32_t* counts, uint32_t blocks, uint32_t* offsets)
{
    __shared__ uint32_t part[1024];
    surface_scan_body(counts, blocks, offsets, part);
}

RGI_DEV void surface_list_body(const float* depth, uint32_t npix, const uint32_t* offsets, uint32_t* list,
                               uint32_t block, uint32_t (*per_wave)[4])
{
    const uint32_t wave = threadIdx.x >> 6, base = block * kSurfaceBlock;
    bool surface[4];
    uint64_t mask[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        uint32_t p = base + k * 256u + threadIdx.x;
        surface[k] = p < npix && depth[p] < kSkyDepth;
        mask[k] = __ballot(surface[k]);
        if ((threadIdx.x & 63u) == 0) per_wave[k][wave] = (uint32_t)__popcll(mask[k]);
    }
    __syncthreads();
    uint32_t run = offsets[block];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        uint32_t before = run;
        for (uint32_t w = 0; w < wave; ++w) before += per_wave[k][w];
        if (surface[k]) list[before + rank_in(mask[k])] = base + k * 256u + threadIdx.x;
        run += ((per_wave[k][0] + per_wave[k][1]) + per_wave[k][2]) + per_wave[k][3];
    }
}

__global__ void __launch_bounds__(256) k_nrc_surface_list(const float* depth, uint32_t npix, const uint32_t* offsets,
                                                          uint32_t* list)
{
    __shared__ uint32_t per_wave[4][4];
    surface_list_body(depth, npix, offsets, list, blockIdx.x, per_wave);
}

// Training records of one frame: `batch` records stratified over the Ns
// surface pixels in row-major order (record i from surface pixel floor((i Ns
// + u_i) / batch), u_i in [0, Ns) from `seed`), each with its surface inputs
// (as the G-buffer query computes them) and the target radiance di_diffuse +
// albedo * color (either term optional); weight 0 for every record when the
// G-buffer has no surface pixel.
//
// With the GI reservoirs (gi_sample, gi_normal: GI-size planes, the frame's
// temporal reservoirs), the record of every `vertex_every`-th i (from i = 0)
// is the path vertex that ReSTIR GI chose for that pixel instead: its
// reservoir's sample point, when the sample is a surface seen on screen
// (evaluateSample's visibility test, restir.sh:454-461: the point projects
// onto the screen and the depth buffer there reconstructs it within 0.1 scene
// units), taken at the pixel where it lies; a pixel whose sample is hidden,
// the sky or absent keeps its own record.
RGI_DEV void gather_body(const rgi_frame& f, const rgi_gbuffer& g, F2 texel, const uint16_t* di_diffuse,
                         const uint16_t* color, const float* gi_sample, const float* gi_normal, uint32_t vertex_every,
                         const uint32_t* surface_total, const uint32_t* surface_list, uint32_t batch, uint32_t seed,
                         float* records, float* targets, uint32_t block)
{
    uint32_t i = block * blockDim.x + threadIdx.x;
    if (i >= batch) return;
    uint64_t ns = *surface_total;
    uint32_t h = pcg(i + pcg(seed));
    uint32_t p = 0;
    if (ns > 0) {
        uint64_t offset = ((uint64_t)(h >> 8) * ns) >> 24;
        p = surface_list[(uint32_t)(((uint64_t)i * ns + offset) / batch)];
    }
    int x = (int)(p % (uint32_t)g.width), y = (int)(p / (uint32_t)g.width);
    if (gi_sample && ns > 0 && i % vertex_every == 0u) {
        // The pixel's GI reservoir (GI-size planes, point-sampled at its uv).
        F2 puv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
        uint32_t gt = texel_index(puv, f.gi_width, f.gi_height);
        F4 sample = load_f4(gi_sample, gt);
        float id = load_f4(gi_normal, gt).w;
        F3 sp = xyz(sample);
        F2 suv;
        float sview;
        if (sample.w > 0.0f && !(id > kSkySampleId * 0.5f) && project_to_screen(f, sp, suv, sview)) {
            uint32_t st = texel_index(suv, g.width, g.height);
            float sdepth = g.depth[st];
            if (sdepth < kSkyDepth && length(reconstruct_position(f, suv, sdepth) - sp) <= 0.1f * f.scene_scale) {
                x = (int)(st % (uint32_t)g.width);
                y = (int)(st / (uint32_t)g.width);
            }
        }
    }
    uint32_t t = (uint32_t)y * (uint32_t)g.width + (uint32_t)x;
    float depth = g.depth[t];
    F2 uv = F2{((float)x + 0.5f) * texel.x, ((float)y + 0.5f) * texel.y};
    F3 pos = reconstruct_position(f, uv, depth);
    F4 nm = unorm8(g.normal[t]);
    F3 n = normalize(F3{nm.x * 2.0f - 1.0f, nm.y * 2.0f - 1.0f, nm.z * 2.0f - 1.0f});
    F4 al = unorm8(g.albedo[t]);
    F3 albedo = F3{al.x, al.y, al.z} * (1.0f - nm.w);
    bool surface = depth < kSkyDepth;
    F3 target = splat3(0.0f);
    if (di_diffuse) target = target + xyz(load_h4(di_diffuse, t));
    // The GI colour is a GI-size plane, point-sampled at the pixel's uv.
    if (color) target = target + xyz(load_h4(color, texel_index(uv, f.gi_width, f.gi_height))) * albedo;
    float4* r = reinterpret_cast<float4*>(records) + (size_t)i * 3;
    r[0] = surface ? make_float4(pos.x, pos.y, pos.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r[1] = surface ? make_float4(n.x, n.y, n.z, 0.0f) : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    r[2] = surface ? make_float4(albedo.x, albedo.y, albedo.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    reinterpret_cast<float4*>(targets)[i] = make_float4(target.x, target.y, target.z, surface ? 1.0f : 0.0f);
}

__global__ void k_nrc_gather(rgi_frame f, rgi_gbuffer g, F2 texel, const uint16_t* di_diffuse, const uint16_t* color,
                             const float* gi_sample, const float* gi_normal, uint32_t vertex_every,
                             const uint32_t* surface_total, const uint32_t* surface_list, uint32_t batch, uint32_t seed,
                             float* records, float* targets)
{
    gather_body(f, g, texel, di_diffuse, color, gi_sample, gi_normal, vertex_every, surface_total, surface_list, batch,
                seed, records, targets, blockIdx.x);
}

// One cache's training-record gather in a batch: nrc_gather's arguments and
// the cache's surface-list scratch. The batched launches give gather q the
// workgroups [first[q], first[q + 1]) of each slice pass (the surface count
// and list, the gather); the scan runs one workgroup per gather.
struct NrcGather {
    rgi_frame f;
    rgi_gbuffer g;
    F2 texel;
    const uint16_t* di_diffuse;
    const uint16_t* color;
    const float* gi_sample;
    const float* gi_normal;
    uint32_t vertex_every, batch, seed, npix, blocks, active;
    uint32_t* counts;
    uint32_t* offsets;
    uint32_t* list;
    float* records;
    float* targets;
};
static_assert(sizeof(NrcGather) <= kNrcGatherBytes, "a gather must fit its slot of a batch's table");

RGI_DEV const NrcGather& gather_at(const uint8_t* gathers, uint32_t i)
{
    return *reinterpret_cast<const NrcGather*>(gathers + (size_t)i * kNrcGatherBytes);
}

__global__ void __launch_bounds__(256) k_nrc_surface_count_batch(const uint8_t* __restrict__ gathers,
                                                                 const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ uint32_t per_wave[4];
    uint32_t b;
    const NrcGather& q = gather_at(gathers, batch_view(first, count, blockIdx.x, b));
    surface_count_body(q.g.depth, q.npix, q.counts, b, per_wave);
}

__global__ void __launch_bounds__(1024) k_nrc_surface_scan_batch(const uint8_t* __restrict__ gathers)
{
    __shared__ uint32_t part[1024];
    const NrcGather& q = gather_at(gathers, blockIdx.x);
    if (!q.active) return;
    surface_scan_body(q.counts, q.blocks, q.offsets, part);
}

__global__ void __launch_bounds__(256) k_nrc_surface_list_batch(const uint8_t* __restrict__ gathers,
                                                                const uint32_t* __restrict__ first, uint32_t count)
{
    __shared__ uint32_t per_wave[4][4];
    uint32_t b;
    const NrcGather& q = gather_at(gathers, batch_view(first, count, blockIdx.x, b));
    surface_list_body(q.g.depth, q.npix, q.offsets, q.list, b, per_wave);
}

__global__ void __launch_bounds__(256) k_nrc_gather_batch(const uint8_t* __restrict__ gathers,
                                                          const uint32_t* __restrict__ first, uint32_t count)
{
    uint32_t b;
    const NrcGather& q = gather_at(gathers, batch_view(first, count, blockIdx.x, b));
    gather_body(q.f, q.g, q.texel, q.di_diffuse, q.color, q.gi_sample, q.gi_normal, q.vertex_every,
                q.offsets + q.blocks, q.list, q.batch, q.seed, q.records, q.targets, b);
}

} // namespace nrc

// ---------------------------------------------------------------- host side

namespace {

using namespace nrc;
Grid grid_of(const rgi_nrc& n);
rgi_nrc inference_view(const rgi_nrc& n);
} // namespace

int nrc_infer_gbuffer(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* emission, int precision,
                      uint16_t* out, hipStream_t stream)
{
    nrc::Grid gp = grid_of(*n);
    const rgi_nrc view = inference_view(*n);
    F2 texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
    dim3 blocks((uint32_t)(g.width + 15) / 16u, (uint32_t)(g.height + 4 * nrc::kInferB - 1) / (4u * nrc::kInferB));
    switch (precision) {
    case RGI_NRC_BF16:
        hipLaunchKernelGGL(nrc::k_nrc_infer_gbuffer<RGI_NRC_BF16>, blocks, dim3(256), 0, stream, view, gp, f, g, texel,
                           emission, out);
        break;
    case RGI_NRC_FP16:
        hipLaunchKernelGGL(nrc::k_nrc_infer_gbuffer<RGI_NRC_FP16>, blocks, dim3(256), 0, stream, view, gp, f, g, texel,
                           emission, out);
        break;
    case RGI_NRC_FP8:
        hipLaunchKernelGGL(nrc::k_nrc_infer_gbuffer<RGI_NRC_FP8>, blocks, dim3(256), 0, stream, view, gp, f, g, texel,
                           emission, out);
        break;
    default: return fail(RGI_E_INVALID, "nrc: unknown precision %d", precision);
    }
    return check_launch("nrc infer (G-buffer)");
}

int nrc_query_plan(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* emission, int precision,
                   uint16_t* out, void* slot, uint32_t* blocks)
{
    if (precision < RGI_NRC_BF16 || precision > RGI_NRC_FP8) return fail(RGI_E_INVALID, "nrc: unknown precision %d", precision);
    nrc::NrcQuery q;
    memset(&q, 0, sizeof(q));
    q.n = inference_view(*n);
    q.gp = grid_of(*n);
    q.f = f;
    q.gb = g;
    q.texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
    q.emission = emission;
    q.out = out;
    q.bx = (uint32_t)(g.width + 15) / 16u;
    q.precision = precision;
    memcpy(slot, &q, sizeof(q));
    *blocks = q.bx * ((uint32_t)(g.height + 4 * nrc::kInferB - 1) / (4u * nrc::kInferB));
    return RGI_OK;
}

int nrc_query_batch(const void* queries, const uint32_t* first, uint32_t count, uint32_t blocks,
                    const bool* precisions, hipStream_t stream)
{
    const uint8_t* q = static_cast<const uint8_t*>(queries);
    if (blocks == 0u) return RGI_OK;
    if (precisions[RGI_NRC_BF16])
        hipLaunchKernelGGL(nrc::k_nrc_infer_gbuffer_batch<RGI_NRC_BF16>, dim3(blocks), dim3(256), 0, stream, q, first, count);
    if (precisions[RGI_NRC_FP16])
        hipLaunchKernelGGL(nrc::k_nrc_infer_gbuffer_batch<RGI_NRC_FP16>, dim3(blocks), dim3(256), 0, stream, q, first, count);
    if (precisions[RGI_NRC_FP8])
        hipLaunchKernelGGL(nrc::k_nrc_infer_gbuffer_batch<RGI_NRC_FP8>, dim3(blocks), dim3(256), 0, stream, q, first, count);
    return check_launch("nrc infer (G-buffer, batch)");
}

// The cache's surface-list scratch for npix pixels (grown after a device
// sync).
static int ensure_surface_list(rgi_nrc* n, uint32_t npix)
{
    const uint32_t blocks = (npix + nrc::kSurfaceBlock - 1) / nrc::kSurfaceBlock;
    if (npix > n->capacity_pixels) {
        (void)hipDeviceSynchronize();
        (void)hipFree(n->surface_counts);
        (void)hipFree(n->surface_offsets);
        (void)hipFree(n->surface_list);
        n->surface_counts = n->surface_offsets = n->surface_list = nullptr;
        n->capacity_pixels = 0;
        uint32_t cap_blocks = blocks;
        if (hipMalloc((void**)&n->surface_counts, (size_t)cap_blocks * 4) != hipSuccess ||
            hipMalloc((void**)&n->surface_offsets, (size_t)(cap_blocks + 1) * 4) != hipSuccess ||
            hipMalloc((void**)&n->surface_list, (size_t)npix * 4) != hipSuccess)
            return fail(RGI_E_NOMEM, "nrc gather: surface list of %u pixels could not be allocated", npix);
        n->capacity_pixels = npix;
    }
    return RGI_OK;
}

int nrc_gather_plan(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* di_diffuse,
                    const uint16_t* color, const float* gi_sample, const float* gi_normal, uint32_t vertex_every,
                    uint32_t batch, uint32_t seed, float* records, float* targets, void* slot, uint32_t* blocks)
{
    const uint32_t npix = (uint32_t)g.width * (uint32_t)g.height;
    int e = ensure_surface_list(n, npix);
    if (e != RGI_OK) return e;
    nrc::NrcGather q;
    memset(&q, 0, sizeof(q));
    q.f = f;
    q.g = g;
    q.texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
    q.di_diffuse = di_diffuse;
    q.color = color;
    q.gi_sample = gi_sample;
    q.gi_normal = gi_normal;
    q.vertex_every = vertex_every;
    q.batch = batch;
    q.seed = seed;
    q.npix = npix;
    q.blocks = (npix + nrc::kSurfaceBlock - 1) / nrc::kSurfaceBlock;
    q.active = 1u;
    q.counts = n->surface_counts;
    q.offsets = n->surface_offsets;
    q.list = n->surface_list;
    q.records = records;
    q.targets = targets;
    memcpy(slot, &q, sizeof(q));
    blocks[0] = q.blocks;
    blocks[1] = (batch + 255u) / 256u;
    return RGI_OK;
}

int nrc_gather_batch(const void* gathers, const uint32_t* first, uint32_t count, const uint32_t* blocks,
                     hipStream_t stream)
{
    if (blocks[0] == 0u) return RGI_OK;
    const uint8_t* q = static_cast<const uint8_t*>(gathers);
    const uint32_t* surface_first = first;
    const uint32_t* gather_first = first + (count + 1);
    hipLaunchKernelGGL(nrc::k_nrc_surface_count_batch, dim3(blocks[0]), dim3(256), 0, stream, q, surface_first, count);
    hipLaunchKernelGGL(nrc::k_nrc_surface_scan_batch, dim3(count), dim3(1024), 0, stream, q);
    hipLaunchKernelGGL(nrc::k_nrc_surface_list_batch, dim3(blocks[0]), dim3(256), 0, stream, q, surface_first, count);
    hipLaunchKernelGGL(nrc::k_nrc_gather_batch, dim3(blocks[1]), dim3(256), 0, stream, q, gather_first, count);
    return check_launch("nrc gather (batch)");
}

int nrc_gather(rgi_nrc* n, const rgi_frame& f, const rgi_gbuffer& g, const uint16_t* di_diffuse,
               const uint16_t* color, const float* gi_sample, const float* gi_normal, uint32_t vertex_every,
               uint32_t batch, uint32_t seed, float* records, float* targets, hipStream_t stream)
{
    uint32_t npix = (uint32_t)g.width * (uint32_t)g.height;
    uint32_t blocks = (npix + nrc::kSurfaceBlock - 1) / nrc::kSurfaceBlock;
    int e = ensure_surface_list(n, npix);
    if (e != RGI_OK) return e;
    hipLaunchKernelGGL(nrc::k_nrc_surface_count, dim3(blocks), dim3(256), 0, stream, g.depth, npix, n->surface_counts);
    hipLaunchKernelGGL(nrc::k_nrc_surface_scan, dim3(1), dim3(1024), 0, stream, n->surface_counts, blocks,
                       n->surface_offsets);
    hipLaunchKernelGGL(nrc::k_nrc_surface_list, dim3(blocks), dim3(256), 0, stream, g.depth, npix, n->surface_offsets,
                       n->surface_list);
    F2 texel = F2{1.0f / (float)g.width, 1.0f / (float)g.height};
    hipLaunchKernelGGL(nrc::k_nrc_gather, dim3((batch + 255) / 256), dim3(256), 0, stream, f, g, texel, di_diffuse,
                       color, gi_sample, gi_normal, vertex_every, n->surface_offsets + blocks, n->surface_list, batch,
                       seed, records, targets);
    return check_launch("nrc gather");
}

namespace {

Grid grid_of(const rgi_nrc& n)
{
    Grid gp;
    gp.levels = n.levels;
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
                       n->scales, n->act_max, step);
    int e = check_launch("nrc adam (mlp)");
    if (e) return e;
    hipLaunchKernelGGL(k_nrc_copies, dim3(((kFwdBlocks + kBwdBlocks) * kFrag + 255) / 256), dim3(256), 0, stream,
                       n->mlp, n->scales, n->w_fwd, n->w_bwd, n->w_fwd_h, n->w_fwd_8);
    return check_launch("nrc weight copies");
}

// The moving average's decay at step t: the configured decay, held below
// (1 + t) / (10 + t) over the first steps so that the initial parameters
// fade out of the average quickly.
float ema_decay_at(const rgi_nrc& n, uint32_t t)
{
    float warm = (float)(1.0 + (double)t) / (float)(10.0 + (double)t);
    return fminf(n.config.ema_decay, warm);
}

// One step of the moving average (after the parameters' Adam step) and the
// inference copies of the averaged weights.
int refresh_average(rgi_nrc* n, hipStream_t stream)
{
    if (!(n->config.ema_decay > 0.0f)) return RGI_OK;
    float d = ema_decay_at(*n, n->steps), om = 1.0f - d;
    hipLaunchKernelGGL(k_nrc_ema_mlp, dim3(6), dim3(256), 0, stream, n->mlp, n->ema_mlp, d, om, n->scales,
                       n->ema_scales);
    hipLaunchKernelGGL(k_nrc_copies, dim3(((kFwdBlocks + kBwdBlocks) * kFrag + 255) / 256), dim3(256), 0, stream,
                       n->ema_mlp, n->ema_scales, n->ema_w_fwd, n->ema_w_bwd, n->ema_w_fwd_h, n->ema_w_fwd_8);
    size_t quads = n->table_floats / 4;
    hipLaunchKernelGGL(k_nrc_ema_table, dim3((unsigned)((quads + 255) / 256)), dim3(256), 0, stream, n->table,
                       n->ema_table, n->ema_table_h, d, om, quads);
    return check_launch("nrc moving average");
}

// The average restarts at the parameters (after creation or set_params):
// copies of the parameters and of their device copies, bit for bit.
int reset_average(rgi_nrc* n, hipStream_t stream)
{
    if (!(n->config.ema_decay > 0.0f)) return RGI_OK;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    bool ok = hipMemcpyAsync(n->ema_table, n->table, n->table_floats * 4, k, stream) == hipSuccess &&
              hipMemcpyAsync(n->ema_table_h, n->table_h, n->table_floats * 2, k, stream) == hipSuccess &&
              hipMemcpyAsync(n->ema_mlp, n->mlp, RGI_NRC_MLP_PARAMS * 4, k, stream) == hipSuccess &&
              hipMemcpyAsync(n->ema_w_fwd, n->w_fwd, (size_t)kFwdBlocks * kFrag * 2, k, stream) == hipSuccess &&
              hipMemcpyAsync(n->ema_w_bwd, n->w_bwd, (size_t)kBwdBlocks * kFrag * 2, k, stream) == hipSuccess &&
              hipMemcpyAsync(n->ema_w_fwd_h, n->w_fwd_h, (size_t)kFwdBlocks * kFrag * 2, k, stream) == hipSuccess &&
              hipMemcpyAsync(n->ema_w_fwd_8, n->w_fwd_8, (size_t)kFwdBlocks * kFrag, k, stream) == hipSuccess &&
              hipMemcpyAsync(n->ema_scales, n->scales, 32 * 4, k, stream) == hipSuccess;
    return ok ? RGI_OK : fail(RGI_E_DEVICE, "nrc: moving average reset failed");
}

// The cache as inference sees it: the averaged weights and table when the
// average is on.
rgi_nrc inference_view(const rgi_nrc& n)
{
    rgi_nrc v = n;
    if (n.config.ema_decay > 0.0f) {
        v.table_h = n.ema_table_h;
        v.w_fwd = n.ema_w_fwd;
        v.w_fwd_h = n.ema_w_fwd_h;
        v.w_fwd_8 = n.ema_w_fwd_8;
        v.scales = n.ema_scales;
    }
    return v;
}

} // namespace

int nrc_step_plan(rgi_nrc* n, const float* records, const float* targets, uint32_t count, float* loss, void* slot,
                  uint32_t* blocks)
{
    if (!n || !records || !targets || count == 0) return fail(RGI_E_INVALID, "nrc step: bad argument");
    const uint32_t tiles = (count + kTile - 1) / kTile;
    int e = ensure_scratch(n, tiles);
    if (e) return e;
    NrcStep st;
    memset(&st, 0, sizeof(st));
    st.a.gp = grid_of(*n);
    st.a.records = records;
    st.a.targets = targets;
    st.a.count = count;
    st.a.inv_count = 1.0f / (float)count;
    st.a.w_fwd = n->w_fwd;
    st.a.w_bwd = n->w_bwd;
    st.a.table_h = n->table_h;
    st.a.table_grad = n->table_grad;
    st.a.acts = n->acts;
    st.a.deltas = n->deltas;
    st.a.tiles = tiles;
    st.a.loss = n->loss;
    st.a.act_max = n->act_max;
    st.active = 1u;
    st.chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    // The step about to be taken: rgi_nrc_train's step count after its
    // increment.
    const uint32_t step = n->steps + 1;
    st.ema = n->config.ema_decay > 0.0f ? 1u : 0u;
    st.d = st.ema ? ema_decay_at(*n, step) : 0.0f;
    st.om = 1.0f - st.d;
    st.partial = n->partial;
    st.mlp = n->mlp;
    st.mlp_m = n->mlp_m;
    st.mlp_v = n->mlp_v;
    st.mlp_grad = n->mlp_grad;
    st.h = adam_args(*n, step);
    st.scales = n->scales;
    st.w_fwd = n->w_fwd;
    st.w_bwd = n->w_bwd;
    st.w_fwd_h = n->w_fwd_h;
    st.w_fwd_8 = n->w_fwd_8;
    st.table = n->table;
    st.table_m = n->table_m;
    st.table_v = n->table_v;
    st.table_grad = n->table_grad;
    st.table_h = n->table_h;
    st.quads = n->table_floats / 4;
    st.ema_mlp = n->ema_mlp;
    st.ema_scales = n->ema_scales;
    st.ema_w_fwd = n->ema_w_fwd;
    st.ema_w_bwd = n->ema_w_bwd;
    st.ema_w_fwd_h = n->ema_w_fwd_h;
    st.ema_w_fwd_8 = n->ema_w_fwd_8;
    st.ema_table = n->ema_table;
    st.ema_table_h = n->ema_table_h;
    st.loss_out = loss;
    memcpy(slot, &st, sizeof(st));
    const uint32_t copies = ((kFwdBlocks + kBwdBlocks) * kFrag + 255) / 256,
                   quads = (uint32_t)((st.quads + 255) / 256);
    blocks[kNrcTrain] = tiles;
    blocks[kNrcWgrad] = st.chunks * 6u;
    blocks[kNrcReduce] = (RGI_NRC_MLP_PARAMS + 255) / 256;
    blocks[kNrcAdamMlp] = 6u;
    blocks[kNrcCopies] = copies;
    blocks[kNrcAdamTable] = quads;
    blocks[kNrcEmaMlp] = st.ema ? 6u : 0u;
    blocks[kNrcEmaCopies] = st.ema ? copies : 0u;
    blocks[kNrcEmaTable] = st.ema ? quads : 0u;
    return RGI_OK;
}

void nrc_step_taken(rgi_nrc* n) { n->steps += 1; }

int nrc_train_batch(const void* steps, const uint32_t* first, uint32_t count, const uint32_t* blocks,
                    hipStream_t stream)
{
    const uint8_t* st = static_cast<const uint8_t*>(steps);
    auto at = [&](int p) { return first + (size_t)p * (count + 1); };
    hipLaunchKernelGGL(k_nrc_clear_max_batch, dim3(count), dim3(64), 0, stream, st);
    hipLaunchKernelGGL(k_nrc_train_batch, dim3(blocks[kNrcTrain]), dim3(256), 0, stream, st, at(kNrcTrain), count);
    hipLaunchKernelGGL(k_nrc_wgrad_batch, dim3(blocks[kNrcWgrad]), dim3(256), 0, stream, st, at(kNrcWgrad), count);
    hipLaunchKernelGGL(k_nrc_reduce_batch, dim3(blocks[kNrcReduce]), dim3(256), 0, stream, st, at(kNrcReduce), count);
    hipLaunchKernelGGL(k_nrc_adam_mlp_batch, dim3(blocks[kNrcAdamMlp]), dim3(256), 0, stream, st, at(kNrcAdamMlp),
                       count);
    hipLaunchKernelGGL(k_nrc_copies_batch<0>, dim3(blocks[kNrcCopies]), dim3(256), 0, stream, st, at(kNrcCopies), count);
    hipLaunchKernelGGL(k_nrc_adam_table_batch, dim3(blocks[kNrcAdamTable]), dim3(256), 0, stream, st, at(kNrcAdamTable),
                       count);
    if (blocks[kNrcEmaMlp] > 0u) {
        hipLaunchKernelGGL(k_nrc_ema_mlp_batch, dim3(blocks[kNrcEmaMlp]), dim3(256), 0, stream, st, at(kNrcEmaMlp),
                           count);
        hipLaunchKernelGGL(k_nrc_copies_batch<1>, dim3(blocks[kNrcEmaCopies]), dim3(256), 0, stream, st,
                           at(kNrcEmaCopies), count);
        hipLaunchKernelGGL(k_nrc_ema_table_batch, dim3(blocks[kNrcEmaTable]), dim3(256), 0, stream, st,
                           at(kNrcEmaTable), count);
    }
    hipLaunchKernelGGL(k_nrc_copy_loss_batch, dim3(count), dim3(64), 0, stream, st);
    return check_launch("nrc train (batch)");
}
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
    c->ema_decay = 0.95f;
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
    if (!(config->ema_decay >= 0.0f && config->ema_decay < 1.0f))
        return fail(RGI_E_INVALID, "nrc_create: ema_decay must be in [0, 1)");
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
        (e = alloc(&n->act_max, 8)) || (e = alloc(&n->loss, 1)) || (e = alloc(&n->levels, 32))) {
        rgi_nrc_destroy(n);
        return e;
    }
    if (config->ema_decay > 0.0f &&
        ((e = alloc(&n->ema_table, n->table_floats)) || (e = alloc(&n->ema_table_h, n->table_floats / 2)) ||
         (e = alloc(&n->ema_mlp, RGI_NRC_MLP_PARAMS)) || (e = alloc(&n->ema_w_fwd, (size_t)kFwdBlocks * kFrag)) ||
         (e = alloc(&n->ema_w_bwd, (size_t)kBwdBlocks * kFrag)) ||
         (e = alloc(&n->ema_w_fwd_h, (size_t)kFwdBlocks * kFrag)) ||
         (e = alloc(&n->ema_w_fwd_8, (size_t)kFwdBlocks * kFrag)) || (e = alloc(&n->ema_scales, 32)))) {
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
    uint32_t levels[32];
    for (int l = 0; l < 16; ++l) {
        uint64_t st = (uint64_t)n->res[l] + 1;
        levels[l] = n->res[l];
        levels[16 + l] = st * st * st <= n->entries ? (uint32_t)st : 0u;
    }
    if (hipMemcpy(n->levels, levels, sizeof(levels), hipMemcpyHostToDevice) != hipSuccess) {
        rgi_nrc_destroy(n);
        return fail(RGI_E_DEVICE, "nrc_create: level table upload failed");
    }
    (void)hipMemset(n->scales, 0, 128);
    size_t entries_total = n->table_floats / 2;
    hipLaunchKernelGGL(k_nrc_init_table, dim3((unsigned)((entries_total + 255) / 256)), dim3(256), 0, 0, n->table,
                       n->table_h, config->seed, entries_total);
    hipLaunchKernelGGL(k_nrc_init_mlp, dim3((RGI_NRC_MLP_PARAMS + 255) / 256), dim3(256), 0, 0, n->mlp, config->seed);
    if ((e = refresh_copies(n, 0, 0)) || (e = check_launch("nrc init")) || (e = reset_average(n, 0))) {
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
                    n->mlp_grad, n->w_fwd, n->w_bwd, n->w_fwd_h, n->w_fwd_8, n->scales, n->act_max, n->loss, n->levels,
                    n->acts, n->deltas, n->partial, n->surface_counts, n->surface_offsets, n->surface_list,
                    n->ema_table, n->ema_table_h, n->ema_mlp, n->ema_w_fwd, n->ema_w_bwd, n->ema_w_fwd_h,
                    n->ema_w_fwd_8, n->ema_scales};
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
    if ((e = refresh_average(n, s))) return e;
    if (loss) hipLaunchKernelGGL(k_nrc_copy_loss, dim3(1), dim3(64), 0, s, n->loss, loss);
    return check_launch("nrc train");
}

int rgi_nrc_train_batch(rgi_nrc* const* caches, const float* const* records, const float* const* targets,
                        const uint32_t* counts, float* const* losses, uint32_t count, void* stream)
{
    if (!caches || !records || !targets || !counts || count == 0 || count > 1024)
        return fail(RGI_E_INVALID, "nrc_train_batch: bad argument (count %u, at most 1024)", count);
    for (uint32_t i = 0; i < count; ++i) {
        if (!caches[i] || !records[i] || !targets[i] || counts[i] == 0)
            return fail(RGI_E_INVALID, "nrc_train_batch: cache %u has a NULL argument or no records", i);
        for (uint32_t k = 0; k < i; ++k)
            if (caches[k] == caches[i]) return fail(RGI_E_INVALID, "nrc_train_batch: cache %u appears twice", i);
    }
    hipStream_t s = as_stream(stream);
    const size_t steps_bytes = (size_t)count * kNrcStepBytes;
    std::vector<uint8_t> host(steps_bytes + (size_t)kNrcPasses * (count + 1) * sizeof(uint32_t), 0);
    uint32_t* first = reinterpret_cast<uint32_t*>(host.data() + steps_bytes);
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t blocks[kNrcPasses] = {};
        int e = nrc_step_plan(caches[i], records[i], targets[i], counts[i], losses ? losses[i] : nullptr,
                              host.data() + (size_t)i * kNrcStepBytes, blocks);
        if (e) return e;
        for (int p = 0; p < kNrcPasses; ++p) first[p * (count + 1) + i + 1] = first[p * (count + 1) + i] + blocks[p];
    }
    uint32_t totals[kNrcPasses];
    for (int p = 0; p < kNrcPasses; ++p) totals[p] = first[p * (count + 1) + count];
    void* table = nullptr;
    if (hipMalloc(&table, host.size()) != hipSuccess) return fail(RGI_E_NOMEM, "nrc_train_batch: table allocation");
    int e = hipMemcpy(table, host.data(), host.size(), hipMemcpyHostToDevice) == hipSuccess
                ? nrc_train_batch(table, reinterpret_cast<const uint32_t*>(static_cast<uint8_t*>(table) + steps_bytes),
                                  count, totals, s)
                : fail(RGI_E_DEVICE, "nrc_train_batch: table upload");
    const bool ran = hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(table);
    if (e) return e;
    if (!ran) return fail(RGI_E_DEVICE, "nrc_train_batch: the steps failed");
    for (uint32_t i = 0; i < count; ++i) nrc_step_taken(caches[i]);
    return RGI_OK;
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
    Grid gp = grid_of(*n);
    const rgi_nrc view = inference_view(*n);
    dim3 blocks((count + 64 * kInferB - 1) / (64 * kInferB));
    switch (precision) {
    case RGI_NRC_BF16:
        hipLaunchKernelGGL(k_nrc_infer<RGI_NRC_BF16>, blocks, dim3(256), 0, s, view, gp, records, count, out);
        break;
    case RGI_NRC_FP16:
        hipLaunchKernelGGL(k_nrc_infer<RGI_NRC_FP16>, blocks, dim3(256), 0, s, view, gp, records, count, out);
        break;
    case RGI_NRC_FP8:
        hipLaunchKernelGGL(k_nrc_infer<RGI_NRC_FP8>, blocks, dim3(256), 0, s, view, gp, records, count, out);
        break;
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
    int e = refresh_copies(n, 0, s);
    return e ? e : reset_average(n, s);
}

int rgi_nrc_query_gbuffer(rgi_nrc* n, const rgi_frame* f, const rgi_gbuffer* g, const uint16_t* emission,
                          int precision, uint16_t* out, void* stream)
{
    if (!n || !f || !g || !out) return fail(RGI_E_INVALID, "nrc_query_gbuffer: NULL argument");
    if (g->width <= 0 || g->height <= 0 || !g->depth || !g->normal || !g->albedo)
        return fail(RGI_E_INVALID, "nrc_query_gbuffer: depth, normal and albedo planes are required");
    return nrc_infer_gbuffer(n, *f, *g, emission, precision, out, as_stream(stream));
}

int rgi_nrc_gather_gbuffer(rgi_nrc* n, const rgi_frame* f, const rgi_gbuffer* g, const uint16_t* di_diffuse,
                           const uint16_t* color, uint32_t batch, uint32_t seed, float* records, float* targets,
                           void* stream)
{
    if (!n || !f || !g || !records || !targets || batch == 0)
        return fail(RGI_E_INVALID, "nrc_gather_gbuffer: bad argument");
    if (g->width <= 0 || g->height <= 0 || !g->depth || !g->normal || !g->albedo)
        return fail(RGI_E_INVALID, "nrc_gather_gbuffer: depth, normal and albedo planes are required");
    return nrc_gather(n, *f, *g, di_diffuse, color, nullptr, nullptr, 1u, batch, seed, records, targets,
                      as_stream(stream));
}

int rgi_nrc_gather_vertices(rgi_nrc* n, const rgi_frame* f, const rgi_gbuffer* g, const rgi_gi_reservoirs* reservoirs,
                            const uint16_t* di_diffuse, const uint16_t* color, uint32_t batch, uint32_t seed,
                            float* records, float* targets, void* stream)
{
    if (!n || !f || !g || !reservoirs || !records || !targets || batch == 0)
        return fail(RGI_E_INVALID, "nrc_gather_vertices: bad argument");
    if (g->width <= 0 || g->height <= 0 || !g->depth || !g->normal || !g->albedo)
        return fail(RGI_E_INVALID, "nrc_gather_vertices: depth, normal and albedo planes are required");
    if (!reservoirs->sample || !reservoirs->normal) return fail(RGI_E_INVALID, "nrc_gather_vertices: reservoir planes are required");
    if (f->gi_width <= 0 || f->gi_height <= 0 || f->gi_width > g->width || f->gi_height > g->height)
        return fail(RGI_E_INVALID, "nrc_gather_vertices: GI size %d x %d out of range", f->gi_width, f->gi_height);
    return nrc_gather(n, *f, *g, di_diffuse, color, reservoirs->sample, reservoirs->normal, 1u, batch, seed, records,
                      targets, as_stream(stream));
}

int rgi_nrc_get_average(const rgi_nrc* n, float* mlp, float* table, void* stream)
{
    if (!n) return fail(RGI_E_INVALID, "nrc_get_average: nrc is NULL");
    if (!(n->config.ema_decay > 0.0f)) return rgi_nrc_get_params(n, mlp, table, stream);
    hipStream_t s = as_stream(stream);
    if (mlp && hipMemcpyAsync(mlp, n->ema_mlp, RGI_NRC_MLP_PARAMS * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(RGI_E_DEVICE, "nrc_get_average: copy failed");
    if (table && hipMemcpyAsync(table, n->ema_table, n->table_floats * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(RGI_E_DEVICE, "nrc_get_average: copy failed");
    return RGI_OK;
}

int rgi_nrc_steps(const rgi_nrc* n, uint32_t* steps)
{
    if (!n || !steps) return fail(RGI_E_INVALID, "nrc_steps: bad argument");
    *steps = n->steps;
    return RGI_OK;
}

} // extern "C"
