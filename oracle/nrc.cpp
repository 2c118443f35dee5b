// Host oracle of the neural radiance cache (include/restir_amd_nrc.h) — TEST
// INFRASTRUCTURE ONLY: the checker of the HIP training and inference kernels.
//
// The reference has no counterpart to follow; this restates the published
// methods the header names: the multiresolution hash encoding of Mueller et
// al. 2022 (trilinear interpolation of per-level feature tables, dense
// indexing where a level fits, the spatial hash of Teschner et al. 2003
// elsewhere), a ReLU MLP without biases, the relative L2 loss of Mueller et
// al. 2021, back-propagation, and Adam (Kingma and Ba 2015) with bias
// correction.
//
// Two modes. `emulate` = 1 rounds where the GPU rounds (FP16 table values,
// BF16 weights, inputs, activations and back-propagated gradients) and sums
// in double, so it differs from the GPU only by the order of FP32
// accumulation inside the matrix cores and the atomics. `emulate` = 0 rounds
// nowhere and is the exact function of the FP32 parameters that the finite
// difference checks differentiate.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "oracle.h"

namespace {

const int kLevels = 16;
const int kWidth = 64;
const int kParams = 5 * 64 * 64 + 3 * 64;

float bf16_round(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

float half_round(float x)
{
    uint16_t h = orc_half_bits(x);
    uint32_t sign = (uint32_t)(h >> 15) << 31, exp = (h >> 10) & 0x1f, man = h & 0x3ff, bits;
    if (exp == 0) {
        float v = ldexpf((float)man, -24);
        return sign ? -v : v;
    }
    if (exp == 31) bits = sign | 0x7f800000u | (man << 13);
    else bits = sign | ((exp + 112) << 23) | (man << 13);
    float r;
    memcpy(&r, &bits, 4);
    return r;
}

struct Model {
    uint32_t res[16];
    uint32_t stride[16];
    uint32_t entries;
    float lo[3], inv_extent[3];
};

Model model_of(const uint32_t* res, uint32_t log2_table, const float* lo, const float* hi)
{
    Model m;
    m.entries = 1u << log2_table;
    for (int l = 0; l < kLevels; ++l) {
        m.res[l] = res[l];
        uint64_t s = (uint64_t)res[l] + 1;
        m.stride[l] = s * s * s <= m.entries ? (uint32_t)s : 0u;
    }
    for (int c = 0; c < 3; ++c) {
        m.lo[c] = lo[c];
        m.inv_extent[c] = 1.0f / (hi[c] - lo[c]);
    }
    return m;
}

float unit(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }
float bounded(float x, float lo, float hi) { return x > lo ? (x < hi ? x : hi) : lo; }

// Entry and trilinear weight of corner k (x + 2y + 4z) of level l, in FP32
// exactly as the kernel computes them.
void corner(const Model& m, int l, const float* p, int k, uint32_t& idx, float& w)
{
    uint32_t r = m.res[l];
    float rf = (float)r;
    float f[3], wt[3];
    int c[3];
    for (int a = 0; a < 3; ++a) {
        f[a] = p[a] * rf;
        int fl = (int)floorf(f[a]);
        c[a] = fl < (int)r - 1 ? fl : (int)r - 1;
        wt[a] = f[a] - (float)c[a];
    }
    uint32_t x = (uint32_t)c[0] + (uint32_t)(k & 1), y = (uint32_t)c[1] + (uint32_t)((k >> 1) & 1),
             z = (uint32_t)c[2] + (uint32_t)(k >> 2);
    w = (((k & 1) ? wt[0] : 1.0f - wt[0]) * ((k & 2) ? wt[1] : 1.0f - wt[1])) * ((k & 4) ? wt[2] : 1.0f - wt[2]);
    uint32_t s = m.stride[l];
    idx = s ? x + y * s + z * (s * s) : ((x ^ (y * 2654435761u) ^ (z * 805459861u)) & (m.entries - 1));
}

struct Sample {
    float p[3];
    double x[64];      // MLP input
    double z[5][64];   // pre-activations of the hidden layers
    double a[6][64];   // a[0] = input, a[k] = ReLU(z[k-1]) (rounded when emulating)
    double y[3];
};

double wt(const float* mlp, int layer, int o, int i, int emulate)
{
    float w = layer < 5 ? mlp[layer * 4096 + o * 64 + i] : mlp[20480 + o * 64 + i];
    return emulate ? (double)bf16_round(w) : (double)w;
}

double round_act(double v, int emulate) { return emulate ? (double)bf16_round((float)v) : v; }

void forward(const Model& m, const float* mlp, const float* table, const float* rec, int emulate, Sample& s)
{
    for (int c = 0; c < 3; ++c) s.p[c] = unit((rec[c] - m.lo[c]) * m.inv_extent[c]);
    for (int l = 0; l < kLevels; ++l) {
        const float* level = table + (size_t)l * m.entries * 2;
        if (emulate) { // FP32 sums of FP16 values in corner order, as the kernel
            float a0 = 0.0f, a1 = 0.0f;
            for (int k = 0; k < 8; ++k) {
                uint32_t idx;
                float w;
                corner(m, l, s.p, k, idx, w);
                a0 = a0 + w * half_round(level[2 * idx]);
                a1 = a1 + w * half_round(level[2 * idx + 1]);
            }
            s.x[2 * l] = bf16_round(a0);
            s.x[2 * l + 1] = bf16_round(a1);
        } else {
            double a0 = 0.0, a1 = 0.0;
            for (int k = 0; k < 8; ++k) {
                uint32_t idx;
                float w;
                corner(m, l, s.p, k, idx, w);
                a0 += (double)w * level[2 * idx];
                a1 += (double)w * level[2 * idx + 1];
            }
            s.x[2 * l] = a0;
            s.x[2 * l + 1] = a1;
        }
    }
    float nn[3], al[3];
    for (int c = 0; c < 3; ++c) {
        nn[c] = bounded(rec[4 + c], -1.0f, 1.0f);
        al[c] = bounded(rec[8 + c], 0.0f, 1.0f);
    }
    float ex[8] = {nn[0] * 0.5f + 0.5f, nn[1] * 0.5f + 0.5f, nn[2] * 0.5f + 0.5f, al[0], al[1], al[2], 1.0f, 0.0f};
    for (int q = 0; q < 32; ++q) s.x[32 + q] = q < 8 ? (emulate ? bf16_round(ex[q]) : ex[q]) : 0.0;
    for (int i = 0; i < 64; ++i) s.a[0][i] = s.x[i];
    for (int k = 0; k < 5; ++k)
        for (int o = 0; o < 64; ++o) {
            double z = 0.0;
            for (int i = 0; i < 64; ++i) z += wt(mlp, k, o, i, emulate) * s.a[k][i];
            s.z[k][o] = emulate ? (double)(float)z : z;
            s.a[k + 1][o] = s.z[k][o] > 0.0 ? round_act(s.z[k][o], emulate) : 0.0;
        }
    for (int o = 0; o < 3; ++o) {
        double y = 0.0;
        for (int i = 0; i < 64; ++i) y += wt(mlp, 5, o, i, emulate) * s.a[5][i];
        s.y[o] = y;
    }
}

} // namespace

extern "C" {

void orc_nrc_forward(const uint32_t* res, uint32_t log2_table, const float* lo, const float* hi, const float* mlp,
                     const float* table, const float* records, uint32_t count, int emulate, double* out)
{
    Model m = model_of(res, log2_table, lo, hi);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        Sample s;
        forward(m, mlp, table, records + (size_t)i * 12, emulate, s);
        for (int c = 0; c < 3; ++c) out[i * 3 + c] = s.y[c];
    }
}

// Loss = (1 / count) sum_i w_i sum_c (y_c - t_c)^2 / (mean_c(y)^2 + 0.01),
// the mean treated as a constant; gradients of every MLP weight (logical
// layout) and every table value. Serial in sample order. With `denominators`
// not NULL, sample i's denominator is written there, or, when `frozen` is
// set, read from there: the loss at perturbed parameters with the
// denominators held fixed is the function the gradients differentiate.
void orc_nrc_gradients(const uint32_t* res, uint32_t log2_table, const float* lo, const float* hi, const float* mlp,
                       const float* table, const float* records, const float* targets, uint32_t count, int emulate,
                       double* loss, double* mlp_grad, double* table_grad, double* denominators, int frozen)
{
    Model m = model_of(res, log2_table, lo, hi);
    memset(mlp_grad, 0, sizeof(double) * kParams);
    memset(table_grad, 0, sizeof(double) * (size_t)kLevels * m.entries * 2);
    double total = 0.0;
    double inv_count = emulate ? (double)(1.0f / (float)count) : 1.0 / (double)count;
    Sample s;
    for (uint32_t n = 0; n < count; ++n) {
        forward(m, mlp, table, records + (size_t)n * 12, emulate, s);
        const float* t = targets + (size_t)n * 4;
        double weight = t[3];
        double mean = (s.y[0] + s.y[1] + s.y[2]) / 3.0;
        double denom = mean * mean + 0.01;
        if (denominators) {
            if (frozen) denom = denominators[n];
            else denominators[n] = denom;
        }
        double e[3], d[16] = {0};
        for (int c = 0; c < 3; ++c) e[c] = s.y[c] - (double)t[c];
        total += weight * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) / denom;
        for (int c = 0; c < 3; ++c) d[c] = round_act(e[c] * (2.0 * weight * inv_count / denom), emulate);
        // Output layer.
        for (int o = 0; o < 3; ++o)
            for (int i = 0; i < 64; ++i) mlp_grad[20480 + o * 64 + i] += d[o] * s.a[5][i];
        double delta[64], next[64];
        for (int i = 0; i < 64; ++i) {
            double g = 0.0;
            for (int o = 0; o < 3; ++o) g += wt(mlp, 5, o, i, emulate) * d[o];
            delta[i] = s.z[4][i] > 0.0 ? round_act(g, emulate) : 0.0;
        }
        for (int k = 4; k >= 0; --k) { // delta holds dL/dz of hidden layer k (W index k)
            for (int o = 0; o < 64; ++o)
                for (int i = 0; i < 64; ++i) mlp_grad[k * 4096 + o * 64 + i] += delta[o] * s.a[k][i];
            for (int i = 0; i < 64; ++i) {
                double g = 0.0;
                for (int o = 0; o < 64; ++o) g += wt(mlp, k, o, i, emulate) * delta[o];
                next[i] = k > 0 ? (s.z[k - 1][i] > 0.0 ? round_act(g, emulate) : 0.0) : g;
            }
            memcpy(delta, next, sizeof(delta));
        }
        // delta = dL/dx: scatter the encoded part onto the table.
        for (int l = 0; l < kLevels; ++l)
            for (int k = 0; k < 8; ++k) {
                uint32_t idx;
                float w;
                corner(m, l, s.p, k, idx, w);
                double* level = table_grad + (size_t)l * m.entries * 2;
                level[2 * idx] += (double)w * delta[2 * l];
                level[2 * idx + 1] += (double)w * delta[2 * l + 1];
            }
    }
    *loss = total * inv_count;
}

// One Adam step in FP32, in the kernels' operation order; entries whose
// gradient is exactly zero are skipped when skip_zero is set.
void orc_nrc_adam(float* p, float* m, float* v, const float* g, uint32_t count, float lr, float beta1, float beta2,
                  float eps, float c1, float c2, int skip_zero)
{
    for (uint32_t i = 0; i < count; ++i) {
        if (skip_zero && g[i] == 0.0f) continue;
        m[i] = beta1 * m[i] + (1.0f - beta1) * g[i];
        v[i] = beta2 * v[i] + (1.0f - beta2) * (g[i] * g[i]);
        float mhat = m[i] / c1;
        float vhat = v[i] / c2;
        p[i] = p[i] - (lr * mhat) / (sqrtf(vhat) + eps);
    }
}

// Entry index and weight of one corner, for the encoding's pins.
void orc_nrc_corner(const uint32_t* res, uint32_t log2_table, int level, const float* p01, int k, uint32_t* idx,
                    float* w)
{
    float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    Model m = model_of(res, log2_table, lo, hi);
    corner(m, level, p01, k, *idx, *w);
}

float orc_bf16_round(float x) { return bf16_round(x); }

} // extern "C"
