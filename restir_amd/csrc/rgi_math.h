// Transcendentals and conversions whose results must be identical on the
// host and on gfx950. The reference calls its GPU vendor's sin, cos and
// fp16 stores (restir.sh:201-212, cs_di_spatial.sc:75-76, RGBA16F images);
// those differ between vendors, so this project defines them once:
//
//   rgi_sincos   Cody-Waite reduction by pi/2 (three-part split of pi/2,
//                each part short enough that k * part is exact for the
//                |k| < 2^11 this is used with), then Taylor polynomials of
//                sin (to r^11) and cos (to r^12) on |r| <= pi/4, evaluated by
//                Horner's rule. Error within 2 ulp of the true value on
//                [-2^10, 2^10].
//   rgi_half_bits  float -> binary16, round to nearest even, with subnormals,
//                overflow to infinity and NaN kept quiet: IEEE 754 conversion.
//   rgi_expf     exp for the denoiser's edge-stopping weights (svgf.sh,
//                fs_svgf_atrous.sc): Cody-Waite reduction by ln 2 (13-bit
//                high part, exact times |k| <= 150), a degree-7 Taylor
//                polynomial of e^r on |r| <= ln(2)/2, and scaling by 2^k
//                through the exponent bits, with one rounding into the
//                subnormal range. Within 2 ulp of the true value.
//   rgi_exp2f    exp2 for specularAlbedo (brdf.sh:73): k = round(x), then
//                e^((x - k) ln 2) by the polynomial of rgi_expf (x - k is
//                exact) and the same scaling. Within 2 ulp.
//   rgi_pow64    pow(x, 64) (kNormalPower, svgf.sh:14) as six squarings.
//
// Every function is plain float arithmetic in a fixed order, so it rounds
// the same wherever it is compiled with FP contraction off.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define RGI_HD __host__ __device__ inline
#else
#define RGI_HD inline
#endif

// pi/2 = RGI_PIO2_A + RGI_PIO2_B + RGI_PIO2_C (A and B have 12 significant bits).
#define RGI_PIO2_A 0x1.92p+0f
#define RGI_PIO2_B 0x1.fb4p-12f
#define RGI_PIO2_C 0x1.4442d2p-24f
#define RGI_TWO_OVER_PI 0x1.45f306p-1f

RGI_HD void rgi_sincos(float x, float* out_sin, float* out_cos)
{
    float k = __builtin_floorf(x * RGI_TWO_OVER_PI + 0.5f);
    float r = x - k * RGI_PIO2_A;
    r = r - k * RGI_PIO2_B;
    r = r - k * RGI_PIO2_C;
    float r2 = r * r;
    // 1/n! rounded to float, n = 3 .. 12.
    float s = -0x1.ae6456p-26f;             // -1/11!
    s = s * r2 + 0x1.71de3ap-19f;           // +1/9!
    s = s * r2 - 0x1.a01a02p-13f;           // -1/7!
    s = s * r2 + 0x1.111112p-7f;            // +1/5!
    s = s * r2 - 0x1.555556p-3f;            // -1/3!
    s = r + (r * r2) * s;
    float c = 0x1.1eed8ep-29f;              // +1/12!
    c = c * r2 - 0x1.27e4fcp-22f;           // -1/10!
    c = c * r2 + 0x1.a01a02p-16f;           // +1/8!
    c = c * r2 - 0x1.6c16c2p-10f;           // -1/6!
    c = c * r2 + 0x1.555556p-5f;            // +1/4!
    c = c * r2 - 0.5f;                      // -1/2!
    c = 1.0f + r2 * c;
    int quadrant = (int)k & 3;
    float sv = s, cv = c;
    if (quadrant == 1) { sv = c; cv = -s; }
    else if (quadrant == 2) { sv = -s; cv = -c; }
    else if (quadrant == 3) { sv = -c; cv = s; }
    *out_sin = sv;
    *out_cos = cv;
}

RGI_HD float rgi_sin(float x)
{
    float s, c;
    rgi_sincos(x, &s, &c);
    return s;
}

RGI_HD float rgi_cos(float x)
{
    float s, c;
    rgi_sincos(x, &s, &c);
    return c;
}

RGI_HD uint32_t rgi_float_bits(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
}

RGI_HD float rgi_bits_float(uint32_t u)
{
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// float -> binary16 bits, round to nearest, ties to even.
RGI_HD uint16_t rgi_half_bits(float value)
{
    uint32_t u = rgi_float_bits(value);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t exponent = (u >> 23) & 0xffu;
    uint32_t mantissa = u & 0x7fffffu;
    if (exponent == 0xffu) {
        // Infinity, or a quiet NaN carrying the top mantissa bits.
        return (uint16_t)(sign | 0x7c00u | (mantissa ? (0x200u | (mantissa >> 13)) : 0u));
    }
    int e = (int)exponent - 127 + 15;
    if (e >= 31) {
        return (uint16_t)(sign | 0x7c00u);
    }
    if (e <= 0) {
        // Subnormal (or zero) in binary16: shift the full significand.
        if (e < -10) {
            return (uint16_t)sign;
        }
        uint32_t significand = mantissa | 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half = significand >> shift;
        uint32_t remainder = significand & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1u);
        if (remainder > halfway || (remainder == halfway && (half & 1u))) {
            half += 1u;
        }
        return (uint16_t)(sign | half);
    }
    uint32_t half = ((uint32_t)e << 10) | (mantissa >> 13);
    uint32_t remainder = mantissa & 0x1fffu;
    if (remainder > 0x1000u || (remainder == 0x1000u && (half & 1u))) {
        half += 1u; // may carry into the exponent, up to infinity: still correct
    }
    return (uint16_t)(sign | half);
}

RGI_HD float rgi_half_float(uint16_t bits)
{
    uint32_t sign = (uint32_t)(bits & 0x8000u) << 16;
    uint32_t exponent = (bits >> 10) & 0x1fu;
    uint32_t mantissa = bits & 0x3ffu;
    if (exponent == 0u) {
        if (mantissa == 0u) {
            return rgi_bits_float(sign);
        }
        // Subnormal: value = mantissa * 2^-24, exact in float.
        float v = (float)mantissa * 0x1p-24f;
        return sign ? -v : v;
    }
    if (exponent == 31u) {
        return rgi_bits_float(sign | 0x7f800000u | (mantissa << 13));
    }
    return rgi_bits_float(sign | ((exponent + 112u) << 23) | (mantissa << 13));
}

// ln 2 = RGI_LN2_HI + RGI_LN2_LO, RGI_LN2_HI with 13 significant bits.
#define RGI_LN2_HI 0x1.62ep-1f
#define RGI_LN2_LO 0x1.0bfbe8p-15f
#define RGI_LOG2E 0x1.715476p+0f
#define RGI_LN2 0x1.62e430p-1f

// e^r times 2^k, for |r| <= ln(2)/2 + a little and k in [-150, 128].
RGI_HD float rgi_exp_scaled(float r, float kf)
{
    // 1/n! rounded to float, n = 2 .. 7.
    float q = 0x1.a01a02p-13f;             // 1/7!
    q = q * r + 0x1.6c16c2p-10f;           // 1/6!
    q = q * r + 0x1.111112p-7f;            // 1/5!
    q = q * r + 0x1.555556p-5f;            // 1/4!
    q = q * r + 0x1.555556p-3f;            // 1/3!
    q = q * r + 0.5f;                      // 1/2!
    float p = 1.0f + (r + (r * r) * q);
    int k = (int)kf;
    if (k > 127) {
        p = p * 2.0f;
        k -= 1;
    }
    if (k < -149) {
        p = p * 0.5f;
        k += 1;
    }
    uint32_t bits = k >= -126 ? (uint32_t)(k + 127) << 23 : 1u << (uint32_t)(k + 149);
    float scale;
    memcpy(&scale, &bits, 4);
    return p * scale;
}

RGI_HD float rgi_expf(float x)
{
    if (x != x) return x;
    if (x > 0x1.62e42ep+6f) return __builtin_inff();
    if (x < -104.0f) return 0.0f;
    float kf = __builtin_floorf(x * RGI_LOG2E + 0.5f);
    float r = (x - kf * RGI_LN2_HI) - kf * RGI_LN2_LO;
    return rgi_exp_scaled(r, kf);
}

RGI_HD float rgi_exp2f(float x)
{
    if (x != x) return x;
    if (x >= 128.0f) return __builtin_inff();
    if (x < -150.0f) return 0.0f;
    float kf = __builtin_floorf(x + 0.5f);
    return rgi_exp_scaled((x - kf) * RGI_LN2, kf);
}

RGI_HD float rgi_pow64(float x)
{
    float x2 = x * x;
    float x4 = x2 * x2;
    float x8 = x4 * x4;
    float x16 = x8 * x8;
    float x32 = x16 * x16;
    return x32 * x32;
}
