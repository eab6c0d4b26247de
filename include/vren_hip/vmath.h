/*
 * vren_hip/vmath.h -- the float arithmetic that the bit-exact stages of the clustered-shading path are
 * defined by.
 *
 * The reference evaluates these stages in GLSL (find_unique_clusters.comp:48-76,
 * clustered_shading.glsl:7-110, assign_lights.comp:77-102, discretize_point_light_positions.comp:31-39,
 * point_light_position_to_view_space.comp:30), where `log`, `pow`, `normalize`, `length` and the mat4*vec4
 * product are whatever the Vulkan driver emits. Their integer outputs (cluster keys, Morton codes, light
 * lists) can only be reproduced bit for bit if every consumer evaluates the same sequence of IEEE-754
 * binary32 operations. This header *is* that sequence: the HIP kernels and the CPU oracle both include it,
 * both are compiled with floating-point contraction disabled, and both rely only on correctly rounded
 * +, -, *, / and sqrt.
 *
 * vren_logf: argument reduction x = 2^e * m with m in [sqrt(1/2), sqrt(2)), then the odd series
 *   ln(m) = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ...),  s = (m - 1) / (m + 1),  |s| <= 0.1716,
 * truncated after s^11 (the first omitted term is below 2^-34 relative). The series coefficients are the
 * exact Taylor coefficients 1/(2k+1); ln 2 is split into a 16-bit head (so e * head is exact) and a tail.
 * Measured error against double-precision log is asserted (<= 2 ulp) by the oracle tests.
 *
 * vren_expf: x = n ln2 + r with |r| <= ln2 / 2 (Cody-Waite split of ln 2), exp(r) by its Taylor series to r^10
 * (exact coefficients 1 / k!), scaled by 2^n in two steps. Used by the sliced cluster-key mode and the synthetic
 * log-uniform generator.
 *
 * vren_powi: a^k for an unsigned integer k by binary exponentiation (right-to-left square-and-multiply),
 * which is all the path needs (the depth-slice exponent of clustered_shading.glsl:221 is the integer
 * slice index).
 */
#ifndef VREN_HIP_VMATH_H_
#define VREN_HIP_VMATH_H_

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define VREN_HD __host__ __device__ inline
#else
#define VREN_HD static inline
#endif

#define VREN_F32_INF (__builtin_inff())

VREN_HD uint32_t vren_f32_bits(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, sizeof(u));
    return u;
}

VREN_HD float vren_bits_f32(uint32_t u)
{
    float x;
    __builtin_memcpy(&x, &u, sizeof(x));
    return x;
}

VREN_HD float vren_sqrtf(float x)
{
    return __builtin_sqrtf(x);
}

VREN_HD float vren_floorf(float x)
{
    return __builtin_floorf(x);
}

/* IEEE binary16 -> binary32, exact (RGBA16F G-buffer normals, gbuffer.cpp:7-13). */
VREN_HD float vren_half_to_float(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1Fu;
    uint32_t man = h & 0x3FFu;
    if (exp == 0x1Fu) {
        return vren_bits_f32(sign | 0x7F800000u | (man << 13));
    }
    if (exp == 0) {
        if (man == 0) {
            return vren_bits_f32(sign);
        }
        /* subnormal half: value = man * 2^-24, exact in binary32 */
        float v = (float)man * 5.9604644775390625e-08f;
        return sign ? -v : v;
    }
    return vren_bits_f32(sign | ((exp + 112u) << 23) | (man << 13));
}

VREN_HD float vren_logf(float x)
{
    uint32_t u = vren_f32_bits(x);
    if (x != x) {
        return x;
    }
    if (x < 0.0f) {
        return vren_bits_f32(0x7FC00000u);
    }
    if (x == 0.0f) {
        return -VREN_F32_INF;
    }
    if (u == 0x7F800000u) {
        return x;
    }
    int32_t e = 0;
    if ((u >> 23) == 0) {
        /* subnormal: scale by 2^25 (exact) */
        x = x * 33554432.0f;
        u = vren_f32_bits(x);
        e = -25;
    }
    e += (int32_t)((u >> 23) & 0xFFu) - 127;
    float m = vren_bits_f32((u & 0x007FFFFFu) | 0x3F800000u); /* [1, 2) */
    if (m > 1.41421353816986083984375f) {                     /* float nearest sqrt(2) */
        m = m * 0.5f;                                          /* exact */
        e += 1;
    }
    float f = m - 1.0f; /* exact (Sterbenz) */
    float s = f / (2.0f + f);
    float z = s * s;
    float p = 0.0909090909090909090909f;                       /* 1/11 */
    p = p * z + 0.1111111111111111111111f;                     /* 1/9 */
    p = p * z + 0.1428571428571428571429f;                     /* 1/7 */
    p = p * z + 0.2f;                                          /* 1/5 */
    p = p * z + 0.3333333333333333333333f;                     /* 1/3 */
    float r = s * z * p;                                       /* s^3/3 + s^5/5 + ... */
    float lnm = 2.0f * s + 2.0f * r;
    float fe = (float)e;
    const float ln2_hi = 0.693145751953125f;     /* ln 2 rounded to 16 significant bits */
    const float ln2_lo = 1.428606765330187e-06f; /* ln 2 - ln2_hi */
    return fe * ln2_hi + (lnm + fe * ln2_lo);
}

/* exp(x) for |x| < 80: x = n ln2 + r, |r| <= ln2/2, exp(r) by its Taylor series to r^10 / 10! */
VREN_HD float vren_expf(float x)
{
    const float inv_ln2 = 1.44269504088896340736f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.428606765330187e-06f;
    float n = vren_floorf(x * inv_ln2 + 0.5f);
    float r = (x - n * ln2_hi) - n * ln2_lo;
    float p = 2.7557319223985890653e-07f;      /* 1/10! */
    p = p * r + 2.7557319223985890653e-06f;    /* 1/9! */
    p = p * r + 2.4801587301587301587e-05f;    /* 1/8! */
    p = p * r + 1.9841269841269841270e-04f;    /* 1/7! */
    p = p * r + 1.3888888888888888889e-03f;    /* 1/6! */
    p = p * r + 8.3333333333333333333e-03f;    /* 1/5! */
    p = p * r + 4.1666666666666666667e-02f;    /* 1/4! */
    p = p * r + 1.6666666666666666667e-01f;    /* 1/3! */
    p = p * r + 0.5f;                          /* 1/2! */
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int32_t ni = (int32_t)n;
    /* scale by 2^n in two steps so that 2^n itself never overflows */
    int32_t n1 = ni / 2;
    int32_t n2 = ni - n1;
    float s1 = vren_bits_f32((uint32_t)(n1 + 127) << 23);
    float s2 = vren_bits_f32((uint32_t)(n2 + 127) << 23);
    return (p * s1) * s2;
}

VREN_HD float vren_powi(float a, uint32_t k)
{
    float result = 1.0f;
    float base = a;
    while (k != 0u) {
        if (k & 1u) {
            result = result * base;
        }
        k >>= 1;
        if (k != 0u) {
            base = base * base;
        }
    }
    return result;
}

VREN_HD float vren_minf(float a, float b)
{
    /* GLSL min(x, y) = y < x ? y : x */
    return b < a ? b : a;
}

VREN_HD float vren_maxf(float a, float b)
{
    /* GLSL max(x, y) = x < y ? y : x */
    return a < b ? b : a;
}

/* float -> uint32 as the path defines it: NaN and negatives give 0, values >= 2^32 saturate. */
VREN_HD uint32_t vren_f2u_sat(float x)
{
    if (!(x > 0.0f)) {
        return 0u;
    }
    if (x >= 4294967296.0f) {
        return 0xFFFFFFFFu;
    }
    return (uint32_t)x;
}

/* vren_f2u_sat(vren_floorf((u + 1) / 2 * 3)), the normal code's cell index along one face axis
 * (clustered_shading.glsl:37-38), for u < 1.5 or NaN, as three comparisons: the float chain is non-decreasing in
 * u, so the index reaches c exactly when u reaches the smallest float that does (found by bisection over float
 * bit patterns; the CPU tests compare both forms on every float below 1.5). A face coordinate is a component
 * divided by the largest one, so |u| <= 1 up to rounding. */
VREN_HD uint32_t vren_third_index(float u)
{
    return (uint32_t)(u >= -0x1.555554p-2f) + (uint32_t)(u >= 0x1.555556p-2f) + (uint32_t)(u >= 0x1.fffffep-1f);
}

VREN_HD float vren_dot3(const float a[3], const float b[3])
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* normalize(v) = v / sqrt(dot(v, v)), component-wise division */
VREN_HD void vren_normalize3(const float v[3], float out[3])
{
    float len = vren_sqrtf(vren_dot3(v, v));
    out[0] = v[0] / len;
    out[1] = v[1] / len;
    out[2] = v[2] / len;
}

/* r = M * v for a column-major mat4 (m[c * 4 + r], glm/GLSL layout), summed left to right */
VREN_HD void vren_mat4_mul_vec4(const float m[16], const float v[4], float r[4])
{
    for (int i = 0; i < 4; i++) {
        r[i] = ((m[0 * 4 + i] * v[0] + m[1 * 4 + i] * v[1]) + m[2 * 4 + i] * v[2]) + m[3 * 4 + i] * v[3];
    }
}

#endif /* VREN_HIP_VMATH_H_ */
