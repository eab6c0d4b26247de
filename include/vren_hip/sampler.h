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
 * - size), mirror(a) = a >= 0 ? a : -(1 + a); CLAMP_TO_EDGE clamp(i, 0, size - 1).
 * A driver's sampler converts u to fixed point (8 bits of sub-texel precision on current GPUs) before it takes
 * the floor and the fraction; that rounding is the driver's and no test can pin it without one. Here u, floor and
 * the fraction are the float operations above, shared by the kernels and the CPU oracle so that both pick the same
 * texels and the same weights. A non-finite coordinate selects texel 0 with weight 0.
 */
#ifndef VREN_HIP_SAMPLER_H_
#define VREN_HIP_SAMPLER_H_

#include "capi.h" /* VREN_FILTER_*, VREN_ADDRESS_* */
#include "vmath.h"

/* floor(x) as an integer, saturated to +-2^40 (far beyond any texture size times a usable repeat count) */
VREN_HD int64_t vren_floor_i64(float x)
{
    const float f = vren_floorf(x);
    if (!(f > -1099511627776.0f)) return f != f ? 0 : -1099511627776ll;
    if (f > 1099511627776.0f) return 1099511627776ll;
    return (int64_t)f;
}

VREN_HD uint32_t vren_wrap_texel(int64_t i, uint32_t size, uint32_t mode)
{
    const int64_t n = (int64_t)size;
    if (mode == VREN_ADDRESS_CLAMP_TO_EDGE) return (uint32_t)(i < 0 ? 0 : (i >= n ? n - 1 : i));
    if (mode == VREN_ADDRESS_MIRRORED_REPEAT) {
        int64_t t = i % (2 * n);
        if (t < 0) t += 2 * n;
        const int64_t a = t - n;
        const int64_t m = a >= 0 ? a : -(1 + a);
        return (uint32_t)(n - 1 - m);
    }
    int64_t r = i % n;
    if (r < 0) r += n;
    return (uint32_t)r;
}

/* texel coordinates along one axis: *i0 (and *i1, *alpha for LINEAR; *i1 = *i0 and *alpha = 0 for NEAREST) */
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
        const float fl = vren_floorf(v);
        const int64_t k = vren_floor_i64(v);
        *alpha = v - fl;
        *i0 = vren_wrap_texel(k, size, mode);
        *i1 = vren_wrap_texel(k + 1, size, mode);
        return;
    }
    *i0 = vren_wrap_texel(vren_floor_i64(u), size, mode);
    *i1 = *i0;
    *alpha = 0.0f;
}

#endif /* VREN_HIP_SAMPLER_H_ */
