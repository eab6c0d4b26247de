/*
 * vren_hip/sampler.h -- texel selection of the material textures that shade samples (shade.comp:89-96).
 *
 * The reference samples its bindless RGBA8 textures (texture_manager.hpp:32, tinygltf_parser.cpp:134-157) with
 * the texture's own sampler from a compute shader, where the implicit level of detail is the base level: lambda = 0
 * selects the magnification filter, and with a single mip level that filter alone decides the result. The
 * coordinate transformation and wrapping follow Vulkan 1.3 (sections 16.6-16.7 of the specification) for an
 * unnormalized coordinate u = s * size:
 *   NEAREST: i = floor(u), u = s * size rounded once;
 *   LINEAR:  i0 = floor(v), i1 = i0 + 1, alpha = v - floor(v), v = s * size - 0.5 rounded once (a fused
 *            multiply-add, so every translation unit computes the same v whatever its contraction mode);
 * then every integer coordinate is wrapped: REPEAT i mod size; MIRRORED_REPEAT (size - 1) - mirror((i mod 2 size)
 * - size), mirror(a) = a >= 0 ? a : -(1 + a); CLAMP_TO_EDGE clamp(i, 0, size - 1). The wrap is evaluated in float
 * (vren_fmod_floor: exact for coordinates below 2^24 in magnitude), which the GPU does in a few instructions where
 * a 64-bit integer remainder takes a long software sequence.
 * A driver's sampler converts u to fixed point (8 bits of sub-texel precision on current GPUs) before it takes
 * the floor and the fraction; that rounding is the driver's and no test can pin it without one. Here u, floor and
 * the fraction are the float operations above, shared by the kernels and the CPU oracle so that both pick the same
 * texels and the same weights. A non-finite coordinate selects texel 0 with weight 0.
 */
#ifndef VREN_HIP_SAMPLER_H_
#define VREN_HIP_SAMPLER_H_

#include "capi.h" /* VREN_FILTER_*, VREN_ADDRESS_* */
#include "vmath.h"

/* x mod n for an integer-valued x and a period n, in float arithmetic: r = x - floor(x / n) n (one fused
 * multiply-add), brought into [0, n) by one correction either way. Exact for |x| < 2^24 (every step is an integer
 * below 2^24); beyond, the same float steps define it on both sides. */
VREN_HD float vren_fmod_floor(float x, float n)
{
    const float q = vren_floorf(x / n);
    float r = __builtin_fmaf(-q, n, x);
    if (r < 0.0f) r += n;
    if (r >= n) r -= n;
    return r;
}

/* the integer coordinate of texel floor(v) (and of its right neighbour) wrapped by `mode` for a texture side of
 * `size` texels: REPEAT floor(v) mod size; MIRRORED_REPEAT t = floor(v) mod 2 size, t < size ? t : 2 size - 1 - t;
 * CLAMP_TO_EDGE clamp(floor(v), 0, size - 1). All in float and 32-bit integers (no 64-bit division). */
VREN_HD void vren_wrap_pair(float v, uint32_t size, uint32_t mode, uint32_t* a, uint32_t* b)
{
    const float n = (float)size;
    if (mode == VREN_ADDRESS_CLAMP_TO_EDGE) {
        const float f = vren_floorf(v);
        const float f0 = f < 0.0f ? 0.0f : (f > n - 1.0f ? n - 1.0f : f);
        const float f1 = f + 1.0f < 0.0f ? 0.0f : (f + 1.0f > n - 1.0f ? n - 1.0f : f + 1.0f);
        *a = (uint32_t)f0;
        *b = (uint32_t)f1;
        return;
    }
    const uint32_t period = mode == VREN_ADDRESS_MIRRORED_REPEAT ? 2u * size : size;
    uint32_t t0 = (uint32_t)vren_fmod_floor(vren_floorf(v), (float)period);
    if (t0 >= period) t0 = period - 1u; /* never taken for exact arithmetic; keeps the index in range */
    const uint32_t t1 = t0 + 1u == period ? 0u : t0 + 1u;
    if (mode == VREN_ADDRESS_MIRRORED_REPEAT) {
        *a = t0 < size ? t0 : 2u * size - 1u - t0;
        *b = t1 < size ? t1 : 2u * size - 1u - t1;
        return;
    }
    *a = t0;
    *b = t1;
}

/* texel coordinates along one axis: *i0 (and *i1, *alpha for LINEAR; *i1 = *i0 and *alpha = 0 for NEAREST). Texture
 * sides up to 2^24 texels. */
VREN_HD void vren_texel_axis(float s, uint32_t size, uint32_t filter, uint32_t mode, uint32_t* i0, uint32_t* i1,
                             float* alpha)
{
    const float u = s * (float)size;
    if (!__builtin_isfinite(u)) {
        *i0 = 0u;
        *i1 = 0u;
        *alpha = 0.0f;
        return;
    }
    if (filter == VREN_FILTER_LINEAR) {
        const float v = __builtin_fmaf(s, (float)size, -0.5f);
        *alpha = v - vren_floorf(v);
        vren_wrap_pair(v, size, mode, i0, i1);
        return;
    }
    uint32_t unused;
    vren_wrap_pair(u, size, mode, i0, &unused);
    *i1 = *i0;
    *alpha = 0.0f;
}

#endif /* VREN_HIP_SAMPLER_H_ */
