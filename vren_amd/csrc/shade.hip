// shade.hip -- shade (clustered_shading.cpp:709-811, shade.comp:69-141, pbr.glsl:6-112) with resolved per-pixel
// materials, for gfx950.
//
// One thread per pixel, one 256-thread workgroup per 32x8 strip of a tile, flattened over every frame of a batch. The
// reference evaluates everything in fp32, including inverse(P) and inverse(V) per invocation; that formula loses
// ~1e-4 relative accuracy wherever a light sits close to the surface or a low roughness makes the GGX
// denominator cancel. This kernel keeps the reference's light model (constants PI = 3.14, EPSILON = 0.001,
// a = roughness, k = roughness, +0.0001 in the specular denominator, ambient 0.03 * albedo, gamma 1/2.2) and
// reorganises the arithmetic so the fp32 result stays within ~1e-6 of an FP64 evaluation:
//   - the pixel's world position is reconstructed in double (per pixel), d = p - light and |d| are evaluated in
//     double per light, and the attenuation through (r - |d|) / r, where it cancels near the light's radius;
//   - the GGX denominator n.h^2 (a^2 - 1) + 1 is evaluated as (1 - |n|^2) + |n x h|^2 / |h|^2 + a^2 (n.h)^2 / |h|^2
//     with the unnormalised half vector h = v - l, a sum of non-negative terms (no cancellation near n.h = 1);
//   - contributions that are exactly zero in the reference formula (attenuation 0 or n.l <= 0) are skipped.
// The background test (view depth >= far - 0.001) uses the double reconstruction and is exactly the oracle's.
#include "chain.hpp"
#include "../../include/vren_hip/sampler.h"

namespace {

constexpr float kPi = 3.14f;
constexpr float kInvPi = 1.0f / 3.14f;

// reciprocal on the transcendental unit (v_rcp_f32, 1 ulp) in place of an IEEE division (~10 instructions); every
// quotient of the forward shade goes through it, which keeps it within ~1e-6 of the FP64 evaluation
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// f16 -> f32 of a normal component on the conversion unit: exact for every finite value and for +-inf; a NaN
// component gives a NaN radiance either way (only its payload could differ from the oracle's software conversion)
__device__ __forceinline__ float normal_h2f(uint32_t h)
{
    const uint16_t b = (uint16_t)h;
    _Float16 x;
    __builtin_memcpy(&x, &b, 2);
    return (float)x;
}

struct PixelIn {
    double P[3];
    bool shaded;
};

__device__ __forceinline__ PixelIn reconstruct(const FrameDev& fd, uint32_t gx, uint32_t gy, float depth)
{
    PixelIn o;
    // view depth and the background test exactly as the oracle evaluates them (shade.comp:75-86): the projection's
    // inverse has zero x/y coefficients in its z and w rows, so r[2] and r[3] do not depend on the pixel's x/y.
    // The NDC x/y (v[0], v[1]) are formed only on the paths that read them.
    const double dz = (double)depth;
    auto ndc = [&](double v[4]) {
        v[0] = (double)gx * fd.inv_W;
        v[1] = (double)gy * fd.inv_H;
        v[2] = dz;
        v[3] = 1.0;
        v[1] = 1.0 - v[1];
        v[0] = v[0] * 2.0 - 1.0;
        v[1] = v[1] * 2.0 - 1.0;
    };
    double r[4];
    if (fd.proj_xy_sparse) {
        // one multiply-add each: the NDC transform folded into the row's two nonzero coefficients (the x/y of the
        // world position differ from the unfolded order by an ulp of double)
        r[0] = fd.rec_xy[0] * (double)gx + fd.rec_xy[1];
        r[1] = fd.rec_xy[2] * (double)gy + fd.rec_xy[3];
    } else {
        double v[4];
        ndc(v);
        for (int i = 0; i < 2; i++)
            r[i] = ((fd.inv_proj_d[0 * 4 + i] * v[0] + fd.inv_proj_d[1 * 4 + i] * v[1]) +
                    fd.inv_proj_d[2 * 4 + i] * v[2]) + fd.inv_proj_d[3 * 4 + i] * v[3];
    }
    double w, z;
    {
        // the view depth and the background test round exactly as the oracle's (no fused multiply-add here; this
        // file is built with -ffp-contract=fast-honor-pragmas)
#pragma clang fp contract(off)
        if (fd.proj_zw_xy_zero) {
            // ((0 x + 0 y) + m2 depth) + m3 is m2 depth + m3 exactly (up to the sign of a zero)
            r[2] = fd.inv_proj_d[2 * 4 + 2] * dz + fd.inv_proj_d[3 * 4 + 2];
            r[3] = fd.inv_proj_d[2 * 4 + 3] * dz + fd.inv_proj_d[3 * 4 + 3];
        } else {
            double v[4];
            ndc(v);
            for (int i = 2; i < 4; i++)
                r[i] = ((fd.inv_proj_d[0 * 4 + i] * v[0] + fd.inv_proj_d[1 * 4 + i] * v[1]) +
                        fd.inv_proj_d[2 * 4 + i] * v[2]) + fd.inv_proj_d[3 * 4 + i] * v[3];
        }
        w = r[3];
    }
    // 1 / w: hardware reciprocal and two Newton steps (within ~1 ulp, double)
    double iw = __builtin_amdgcn_rcp(w);
    iw = iw + iw * (1.0 - w * iw);
    iw = iw + iw * (1.0 - w * iw);
    // the view depth r[2] / w: the product r[2] * iw is within 2 ulp of the correctly rounded quotient the oracle
    // takes, so it decides the background test whenever it lies more than ~45 ulp from the threshold; otherwise
    // (and for a non-finite product, whose test is false) the test takes the IEEE division itself
    const double bg = (double)fd.bg_threshold;
    z = r[2] * iw;
    const bool near = !(__builtin_fabs(z - bg) > 1e-14 * __builtin_fabs(bg));
    if (__builtin_amdgcn_ballot_w64(near)) { // a wave-uniform branch: the division is skipped by almost every wave
#pragma clang fp contract(off)
        const double zd = r[2] / w;
        z = near ? zd : z;
    }
    o.shaded = z < bg;
    const double q[4] = {r[0] * iw, r[1] * iw, z, 1.0};
    for (int i = 0; i < 3; i++)
        o.P[i] = ((fd.inv_view_d[0 * 4 + i] * q[0] + fd.inv_view_d[1 * 4 + i] * q[1]) + fd.inv_view_d[2 * 4 + i] * q[2]) +
                 fd.inv_view_d[3 * 4 + i] * q[3];
    return o;
}

struct Surface {
    float V[3], N[3];
    float albedo[3];
    float rough, metal;
    float nn;      // 1 - |N|^2
    float a2;      // roughness^2
    float F0[3];
    float NdotV, ggxV;
};

__device__ __forceinline__ float dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// pbr_apply_light (pbr.glsl:42-73) for a light direction L (from the light towards the surface) and radiance
__device__ __forceinline__ void apply_light(const Surface& s, const float L[3], const float rad[3], float acc[3])
{
    const float mL[3] = {-L[0], -L[1], -L[2]};
    const float NdotL = fmaxf(dot3(s.N, mL), 0.0f);
    if (NdotL <= 0.0f) return;
    const float h[3] = {s.V[0] + mL[0], s.V[1] + mL[1], s.V[2] + mL[2]};
    const float hh = dot3(h, h);
    const float Nh = dot3(s.N, h);
    const float irh = hh > 0.0f ? __builtin_amdgcn_rsqf(hh) : 0.0f; // 1 / |h|
    float inner = 1.0f;
    if (Nh > 0.0f && hh > 0.0f) {
        const float c[3] = {s.N[1] * h[2] - s.N[2] * h[1], s.N[2] * h[0] - s.N[0] * h[2], s.N[0] * h[1] - s.N[1] * h[0]};
        const float inv_hh = irh * irh;
        inner = s.nn + dot3(c, c) * inv_hh + s.a2 * (Nh * Nh) * inv_hh;
    }
    const float k = s.rough;
    const float HdotV = fmaxf(dot3(h, s.V) * irh, 0.0f);
    float om = 1.0f - HdotV;
    float om2 = om * om;
    const float f5 = om2 * om2 * om;
    // spec = NDF G / (4 n.v n.l + 0.0001) with NDF = a^2 / (PI inner^2) and G = ggxV n.l / (n.l (1 - k) + k):
    // the three quotients share one reciprocal
    const float spec = (s.a2 * s.ggxV * NdotL) *
                       frcp(kPi * inner * inner * (NdotL * (1.0f - k) + k) * (4.0f * s.NdotV * NdotL + 0.0001f));
    const float km = 1.0f - s.metal;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float F = s.F0[c] + (1.0f - s.F0[c]) * f5;
        const float kD = (1.0f - F) * km;
        acc[c] += fmaxf((kD * s.albedo[c] * kInvPi + F * spec) * rad[c] * NdotL, 0.0f);
    }
}

// per-pixel constants of the forward shade, arranged for shade_light
struct ShadeSurface {
    float V[3], N[3];
    float GA[3];  // (1 - F0) albedo (1 - metallic) / PI: the diffuse term of a light is (1 - f5) GA
    float G[3];   // 1 - F0
    float F0[3];
    float k;      // roughness, the Schlick-GGX k (pbr.glsl:26-40)
    float nn;     // 1 - |N|^2
    float a2;     // roughness^2
    float a2ggxV; // roughness^2 n.v / (n.v (1 - k) + k)
    float NdotV4; // 4 n.v
};

// pbr_apply_light (pbr.glsl:42-73) for a light direction L (from the light towards the surface), the light's
// colour and its attenuation: acc += max(((1 - F) A + F spec) colour att n.l, 0) per channel, with 1 - F taken as
// (1 - F0)(1 - f5), which it equals (F = F0 + (1 - F0) f5), without the cancellation of 1 - F. The weight
// w = att n.l is positive (both factors are tested), so max(x w, 0) = max(x, 0) w and w leaves the clamp.
__device__ __forceinline__ void shade_light(const ShadeSurface& s, const float L[3], const float col[3], float att,
                                            float acc[3])
{
    const float mL[3] = {-L[0], -L[1], -L[2]};
    const float NdotL = fmaxf(dot3(s.N, mL), 0.0f);
    if (NdotL <= 0.0f) return;
    const float h[3] = {s.V[0] + mL[0], s.V[1] + mL[1], s.V[2] + mL[2]};
    const float hh = dot3(h, h);
    const float Nh = dot3(s.N, h);
    const float irh = hh > 0.0f ? __builtin_amdgcn_rsqf(hh) : 0.0f; // 1 / |h|
    float inner = 1.0f;
    if (Nh > 0.0f && hh > 0.0f) {
        const float c[3] = {s.N[1] * h[2] - s.N[2] * h[1], s.N[2] * h[0] - s.N[0] * h[2], s.N[0] * h[1] - s.N[1] * h[0]};
        const float inv_hh = irh * irh;
        inner = s.nn + dot3(c, c) * inv_hh + s.a2 * (Nh * Nh) * inv_hh;
    }
    const float HdotV = fmaxf(dot3(h, s.V) * irh, 0.0f);
    const float om = 1.0f - HdotV;
    const float om2 = om * om;
    const float f5 = om2 * om2 * om;
    const float om5 = 1.0f - f5;
    // spec = NDF G / (4 n.v n.l + 0.0001) with NDF = a^2 / (PI inner^2) and G = ggxV n.l / (n.l (1 - k) + k):
    // the three quotients share one reciprocal
    const float spec = (s.a2ggxV * NdotL) *
                       frcp(kPi * inner * inner * (NdotL * (1.0f - s.k) + s.k) * (s.NdotV4 * NdotL + 0.0001f));
    const float w = att * NdotL;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float F = s.F0[c] + s.G[c] * f5;
        acc[c] += fmaxf((s.GA[c] * om5 + F * spec) * col[c], 0.0f) * w;
    }
}

constexpr uint32_t kStage = 512; // lights staged in LDS per workgroup of the adjoint

// A light as the forward shade's loop reads it from LDS: the position and r^2 already in double (the loop's
// distance test and the cancelling difference r^2 - |d|^2 are taken in double), the colour and radius, 1 / r.
// A light past the table or with r <= 0 (or NaN) gets rr = -1, which no squared distance is below.
struct alignas(16) StagedLight {
    double x, y, z, rr;
    float4 col; // rgb, radius
    float inv_r, pad0, pad1, pad2;
};
constexpr uint32_t kFwdStage = 256; // lights staged per forward workgroup (16 KB; tiles stage ~6 on average)
constexpr int kShadePPT = 1;     // default pixels per thread of the forward shade (2 is a tuner choice)

// The pixel's list entries [off, end): light records from the LDS stage (STAGED, entry e at e - start, the
// position's w holding 1 / radius) or from global memory; SMOOTH selects the windowed attenuation.
__device__ __forceinline__ StagedLight stage_light(const FrameDev& fd, uint32_t li)
{
    StagedLight sl;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f), c = p;
    if (li < fd.L) {
        const vf4 cv = *((gvf4*)fd.lights + li), pv = *((gvf4*)fd.light_pos + li);
        c = make_float4(cv.x, cv.y, cv.z, cv.w);
        p = make_float4(pv.x, pv.y, pv.z, pv.w);
    }
    sl.x = (double)p.x;
    sl.y = (double)p.y;
    sl.z = (double)p.z;
    sl.rr = c.w > 0.0f ? (double)c.w * (double)c.w : -1.0;
    sl.col = c;
    sl.inv_r = frcp(c.w);
    sl.pad0 = sl.pad1 = sl.pad2 = 0.0f;
    return sl;
}

template <bool STAGED, bool SMOOTH>
__device__ __forceinline__ void light_loop(const FrameDev& fd, const PixelIn& px, const ShadeSurface& s, uint32_t off,
                                           uint32_t end, uint32_t start, const StagedLight* s_light, float acc[3])
{
    for (uint32_t e = off; e < end; e++) {
        StagedLight sl;
        if constexpr (STAGED) {
            sl = s_light[e - start];
        } else {
            const uint32_t li = *((guint*)fd.indices + e);
            if (li >= fd.L) continue;
            sl = stage_light(fd, li);
        }
        // d = p - light (pbr.glsl:87) and |d| in double; the attenuation is evaluated through (r - |d|) / r,
        // which is where 1 - (|d|/r)^1.5 (or 1 - (|d|/r)^4) cancels near the light's radius
        const double d[3] = {px.P[0] - sl.x, px.P[1] - sl.y, px.P[2] - sl.z};
        const double dist2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        const float r = sl.col.w;
        const double rr = sl.rr;
        if (!(dist2 > 0.0) || !(dist2 < rr)) continue;
        const float d2f = (float)dist2;
        const float inv = __builtin_amdgcn_rsqf(d2f); // 1 / |d|
        const float s0 = d2f * inv;                     // |d|
        // 1 - x = (r - |d|) / r = (r^2 - |d|^2) / ((r + |d|) r): the difference that cancels near the radius is
        // taken in double, the rest needs only float accuracy
        const float num = (float)(rr - dist2);
        float att;
        if constexpr (SMOOTH) {
            // (1 - x^4)^2 = ((1 - x)(1 + x)(1 + x^2))^2
            const float omx = num * frcp((r + s0) * r);
            const float x = 1.0f - omx;
            const float w = omx * (1.0f + x) * (1.0f + x * x);
            att = w * w;
        } else {
            // 1 - x^1.5 = (1 - x)(1 + x + x^2) / (1 + x^1.5), with x = |d| / r
            const float x = s0 * sl.inv_r;
            att = num * (1.0f + x + x * x) * frcp((r + s0) * r * (1.0f + x * __builtin_amdgcn_sqrtf(x)));
        }
        if (!(att > 0.0f)) continue;
        const float L[3] = {(float)d[0] * inv, (float)d[1] * inv, (float)d[2] * inv};
        const float col[3] = {sl.col.x, sl.col.y, sl.col.z};
        shade_light(s, L, col, att, acc);
    }
}

// The per-pixel constants of the forward shade and the radiance accumulator's start, the ambient term
// (shade.comp:133), to which the lights are added.
__device__ __forceinline__ ShadeSurface shade_surface(const FrameDev& fd, const PixelIn& px, uint2 nb, float4 ar,
                                                      float metal, float acc[3])
{
    ShadeSurface s;
    {
        const float V[3] = {(float)(fd.eye[0] - px.P[0]), (float)(fd.eye[1] - px.P[1]), (float)(fd.eye[2] - px.P[2])};
        const float il = __builtin_amdgcn_rsqf(dot3(V, V)); // |V| is scene-sized, never subnormal
        for (int i = 0; i < 3; i++) s.V[i] = V[i] * il;
    }
    s.N[0] = normal_h2f(nb.x & 0xFFFFu);
    s.N[1] = normal_h2f(nb.x >> 16);
    s.N[2] = normal_h2f(nb.y & 0xFFFFu);
    s.nn = (float)(1.0 - ((double)s.N[0] * s.N[0] + (double)s.N[1] * s.N[1] + (double)s.N[2] * s.N[2]));
    const float rough = ar.w, km = 1.0f - metal;
    s.k = rough;
    s.a2 = rough * rough;
    const float albedo[3] = {ar.x, ar.y, ar.z};
    for (int c = 0; c < 3; c++) {
        s.F0[c] = 0.04f * km + albedo[c] * metal;
        s.G[c] = 1.0f - s.F0[c];
        s.GA[c] = s.G[c] * (albedo[c] * km * kInvPi);
    }
    const float NdotV = fmaxf(dot3(s.N, s.V), 0.0f);
    s.a2ggxV = s.a2 * (NdotV * frcp(NdotV * (1.0f - rough) + rough));
    s.NdotV4 = 4.0f * NdotV;
    for (int c = 0; c < 3; c++) acc[c] = 0.03f * albedo[c];
    return s;
}

// The pixel's list entries [off, end) from the LDS stage (STAGED, entry e at e - start) or from global memory; one
// loop per (staging, attenuation) pair: both are uniform over the workgroup, so no entry pays for the other branch's
// registers or its test.
template <bool STAGED>
__device__ __forceinline__ void shade_lights(const FrameDev& fd, const PixelIn& px, const ShadeSurface& s,
                                             uint32_t off, uint32_t end, uint32_t start, const StagedLight* s_light,
                                             float acc[3])
{
    if (fd.attenuation == VREN_ATTEN_SMOOTH) light_loop<STAGED, true>(fd, px, s, off, end, start, s_light, acc);
    else light_loop<STAGED, false>(fd, px, s, off, end, start, s_light, acc);
}

// The directional lights, the gamma and the pixel's store.
__device__ __forceinline__ void shade_finish(const FrameDev& fd, const ShadeSurface& s, float acc[3], uint64_t p)
{
    for (uint32_t i = 0; i < fd.n_dir; i++) {
        const float4* D = reinterpret_cast<const float4*>(fd.dir_lights) + 2 * i;
        const float4 dir = D[0], col = D[1];
        const float L[3] = {dir.x, dir.y, dir.z};
        const float c3[3] = {col.x, col.y, col.z};
        shade_light(s, L, c3, 1.0f, acc);
    }
    // gamma (shade.comp:135-137): c^(1/2.2) = 2^(log2(c) / 2.2) on the transcendental unit (v_log_f32 /
    // v_exp_f32, ~1 ulp each: ~1e-7 relative on the result); c >= 0, and c = 0 gives 2^-inf = 0
    const float gexp = 1.0f / 2.2f;
    float4 o;
    o.x = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(acc[0]) * gexp);
    o.y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(acc[1]) * gexp);
    o.z = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(acc[2]) * gexp);
    o.w = 1.0f;
    const vf4 ov = {o.x, o.y, o.z, o.w};
    // a streaming store: the radiance is not read again by the chain (1.4 % of the kernel against a cached store)
    __builtin_nontemporal_store(ov, (__attribute__((address_space(1))) vf4*)fd.out + p);
}

// Shades one pixel whose inputs are loaded: its list entries [off, end) come from the LDS stage (when `staged`,
// entry e at e - start) or from global memory.
__device__ __forceinline__ void shade_pixel(const FrameDev& fd, const PixelIn& px, uint2 nb, float4 ar, float metal,
                                            bool has_list, uint32_t off, uint32_t end, bool staged, uint32_t start,
                                            const StagedLight* s_light, uint64_t p)
{
    float acc[3];
    const ShadeSurface s = shade_surface(fd, px, nb, ar, metal, acc);
    if (has_list) {
        if (staged) shade_lights<true>(fd, px, s, off, end, start, s_light, acc);
        else shade_lights<false>(fd, px, s, off, end, start, s_light, acc);
    }
    shade_finish(fd, s, acc, p);
}

// ---- material resolve (shade.comp:89-96) -------------------------------------------------------------------------

typedef float vf2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) vf2 gvf2;
typedef const __attribute__((address_space(1))) uint16_t gu16;

// one RGBA8 UNORM texel (a 4-byte load) as floats c / 255
__device__ __forceinline__ float4 texel_rgba8(const vren_texture& t, uint32_t i, uint32_t j)
{
    const uint32_t v = *((guint*)(const uint32_t*)t.texels + ((uint64_t)j * t.width + i));
    constexpr float k = 1.0f / 255.0f;
    return make_float4((float)(v & 255u) * k, (float)((v >> 8) & 255u) * k, (float)((v >> 16) & 255u) * k,
                       (float)(v >> 24) * k);
}

// texture(textures[ti], (s, t)) at the base level: the texture's filter, address modes and texel selection of
// include/vren_hip/sampler.h; zeros for an index past the table or an empty texture
__device__ __forceinline__ float4 sample_texture(const FrameDev& fd, uint32_t ti, float s, float t)
{
    if (ti >= fd.n_textures) return make_float4(0.f, 0.f, 0.f, 0.f);
    const vren_texture tx = fd.textures[ti];
    if (!tx.texels || tx.width == 0 || tx.height == 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t i0, i1, j0, j1;
    float a, b;
    vren_texel_axis(s, tx.width, tx.filter, tx.address_u, &i0, &i1, &a);
    vren_texel_axis(t, tx.height, tx.filter, tx.address_v, &j0, &j1, &b);
    const float4 c00 = texel_rgba8(tx, i0, j0);
    if (tx.filter != VREN_FILTER_LINEAR) return c00;
    const float4 c10 = texel_rgba8(tx, i1, j0), c01 = texel_rgba8(tx, i0, j1), c11 = texel_rgba8(tx, i1, j1);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return make_float4(c00.x * w00 + c10.x * w10 + c01.x * w01 + c11.x * w11,
                       c00.y * w00 + c10.y * w10 + c01.y * w01 + c11.y * w11,
                       c00.z * w00 + c10.z * w10 + c01.z * w01 + c11.z * w11,
                       c00.w * w00 + c10.w * w10 + c01.w * w01 + c11.w * w11);
}

// albedo (rgb) + roughness (w) and metallic of a pixel from its texcoord and material index; a material index past
// the table reads a zero material
__device__ __forceinline__ void resolve_material(const FrameDev& fd, vf2 tc, uint32_t mi, float4& ar, float& metal)
{
    ar = make_float4(0.f, 0.f, 0.f, 0.f);
    metal = 0.0f;
    if (mi >= fd.n_materials) return;
    const vren_material m = fd.materials[mi];
    const float4 base = sample_texture(fd, m.base_color_texture_idx, tc.x, tc.y);
    const float4 mr = m.metallic_roughness_texture_idx == m.base_color_texture_idx
                          ? base
                          : sample_texture(fd, m.metallic_roughness_texture_idx, tc.x, tc.y);
    ar = make_float4(base.x * m.base_color_factor[0], base.y * m.base_color_factor[1],
                     base.z * m.base_color_factor[2], mr.y * m.roughness_factor);
    metal = mr.z * m.metallic_factor;
}

// [begin, end) of the index-buffer entries of a tile's light lists: the tile's clusters are the consecutive indices
// [c0, c0 + u) (clusters kernel) and their lists are consecutive in the index buffer, so the range runs from the
// first cluster's offset to the end of the last list (both clamped to the capacities like the per-pixel lists).
__global__ __launch_bounds__(256) void tile_range_kernel(const FrameDev* __restrict__ frames,
                                                         const uint32_t* __restrict__ tile_frame, uint32_t total_tiles,
                                                         const uint2* __restrict__ tile_cl, uint2* __restrict__ tile_range)
{
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= total_tiles) return;
    const FrameDev& fd = frames[frame_of_tile(tile_frame, g)];
    const uint2 c = tile_cl[g];
    uint32_t a = 0, b = 0;
    if (g - fd.tile_base < fd.Sx * fd.Sy && fd.L > 0 && c.y > 0 && c.x < fd.key_cap) { // not a padding tile
        const uint32_t c1 = min(c.x + c.y, fd.key_cap) - 1u;
        a = min(fd.offsets[c.x], fd.index_cap);
        const uint64_t e = (uint64_t)fd.offsets[c1] + fd.counts[c1];
        b = e > fd.index_cap ? fd.index_cap : (uint32_t)e;
    }
    tile_range[g] = make_uint2(a, b);
}

// A 256-thread workgroup shades PPT 32x8 strips of a 32x32 tile (PPT pixels per thread, their loads issued
// together), flattened over every frame of the table. A wave covers two whole 32-pixel rows of the tile, so each
// 4-byte G-buffer plane is read in whole 128-byte lines: with 16-pixel-wide pieces, the other half of every line
// belonged to the neighbouring workgroup, which runs on another XCD (another L2) and fetched the line again. Every pixel of a tile reads its light list from one
// contiguous run of the index buffer, which the workgroup stages in LDS as light records (position with 1 / radius,
// colour with radius) before shading each pixel from LDS. With tile_range (the chain, whose references are the
// clusters kernel's), the run is known from one load, so the staged indices are requested beside the pixel data and
// the light records beside the pixels' list bounds: two dependent rounds of loads. Without it (references from the
// caller), the workgroup first reduces the run from its pixels' lists.
// MAT: the table may hold frames on the material path (texcoords + material index + material and texture tables)
template <int PPT, bool MAT>
// at most 64 VGPRs: eight waves per SIMD (the split light loops would otherwise take 66 and seven)
__global__ __launch_bounds__(256, 8) void shade_kernel(const FrameDev* __restrict__ frames,
                                                    const uint32_t* __restrict__ tile_frame,
                                                    const uint2* __restrict__ tile_range, uint32_t total_tiles)
{
    constexpr uint32_t BPT = 4 / PPT; // workgroups per tile
    static_assert(kFwdStage == 256, "one staged entry per thread");
    __shared__ StagedLight s_light[kFwdStage];
    __shared__ uint32_t s_min[4], s_max[4];
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / VREN_WAVE;
    // the BPT workgroups of a tile run on one XCD (workgroup b runs on XCD b mod 8) and consecutive tiles on
    // consecutive XCDs, so that a tile's list range, indices and light records are fetched into one L2; the grid is
    // rounded up to whole rounds of 8 tiles
    const uint32_t k = blockIdx.x / 8;
    const uint32_t g = (k / BPT) * 8 + blockIdx.x % 8, sub = k % BPT;
    if (g >= total_tiles) return;
    const uint32_t f = frame_of_tile(tile_frame, g);
    const FrameDev& fd = frames[f];
    const uint32_t t = g - fd.tile_base;
    if (t >= fd.Sx * fd.Sy) return; // padding after the frame's tiles (the whole workgroup)
    const uint32_t ti = t % fd.Sx, tj = t / fd.Sx;
    uint32_t start = 0, nstage = 0;
    uint32_t li = 0xFFFFFFFFu;
    if (tile_range) {
        const uint2 rg = tile_range[g];
        start = rg.x;
        nstage = rg.y > rg.x ? rg.y - rg.x : 0u;
        if (nstage <= kFwdStage && tid < nstage) li = *((guint*)fd.indices + start + tid);
    }
    PixelIn px[PPT];
    uint2 nb[PPT];
    float4 ar[PPT];
    float metal[PPT];
    uint32_t cl_raw[PPT];
    uint64_t p[PPT];
    bool in[PPT];
    vf2 tc[PPT];
    uint32_t mi[PPT];
    const bool mat = MAT && fd.texcoords != nullptr;
    // every per-pixel load is issued before any of them is used
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        const uint32_t strip = sub * PPT + q;
        const uint32_t gx = ti * 32 + (tid & 31), gy = tj * 32 + strip * 8 + (tid >> 5);
        in[q] = gx < fd.W && gy < fd.H;
        p[q] = in[q] ? (uint64_t)gy * fd.W + gx : 0;
        px[q].shaded = false;
        nb[q] = make_uint2(0u, 0u);
        ar[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        metal[q] = 0.0f;
        cl_raw[q] = 0xFFFFFFFFu;
        tc[q] = vf2{0.0f, 0.0f};
        mi[q] = 0xFFFFFFFFu;
        if (in[q]) {
            const float depth = *((gfloat*)fd.depth + p[q]);
            const vu2 nv = *((gvu2*)fd.normals + p[q]);
            nb[q] = make_uint2(nv.x, nv.y);
            if (mat) {
                tc[q] = *((gvf2*)fd.texcoords + p[q]);
                mi[q] = *((gu16*)fd.mat_idx + p[q]);
            } else {
                const vf4 av = *((gvf4*)fd.albedo_rough + p[q]);
                ar[q] = make_float4(av.x, av.y, av.z, av.w);
                metal[q] = *((gfloat*)fd.metallic + p[q]);
            }
            cl_raw[q] = *((guint*)fd.refs + p[q]);
            px[q] = reconstruct(fd, gx, gy, depth);
        }
    }
    if (mat) {
#pragma unroll
        for (int q = 0; q < PPT; q++)
            if (px[q].shaded) resolve_material(fd, tc[q], mi[q], ar[q], metal[q]);
    }
    // each pixel's light list
    bool has_list[PPT];
    uint32_t my_off[PPT], my_end[PPT];
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        has_list[q] = px[q].shaded && fd.L > 0 && cl_raw[q] < fd.key_cap;
        my_off[q] = 0;
        my_end[q] = 0;
        if (has_list[q]) {
            const uint32_t o = *((guint*)fd.offsets + cl_raw[q]);
            my_off[q] = min(o, fd.index_cap);
            const uint64_t e = (uint64_t)o + *((guint*)fd.counts + cl_raw[q]);
            my_end[q] = e > fd.index_cap ? fd.index_cap : (uint32_t)e;
        }
    }
    if (!tile_range) {
        // the workgroup's range of list entries, reduced from its pixels' lists
        uint32_t rmin = 0xFFFFFFFFu, rmax = 0u;
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            if (has_list[q]) {
                rmin = min(rmin, my_off[q]);
                rmax = max(rmax, my_end[q]);
            }
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            rmin = min(rmin, shfl_xor(rmin, m));
            rmax = max(rmax, shfl_xor(rmax, m));
        }
        if (lane == 0) {
            s_min[wave] = rmin;
            s_max[wave] = rmax;
        }
        __syncthreads();
        const uint32_t a = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
        const uint32_t b = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
        start = a == 0xFFFFFFFFu ? 0u : a;
        nstage = (a != 0xFFFFFFFFu && b > a) ? b - a : 0u;
        if (nstage <= kFwdStage && tid < nstage) li = *((guint*)fd.indices + start + tid);
    }
    if (nstage <= kFwdStage) {
        if (tid < nstage) s_light[tid] = stage_light(fd, li);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PPT; q++)
            if (px[q].shaded)
                shade_pixel(fd, px[q], nb[q], ar[q], metal[q], has_list[q], my_off[q], my_end[q], true, start, s_light,
                            p[q]);
        return;
    }
    // a range longer than the stage (a few tiles of frames with tens of thousands of lights): the range in windows
    // of kFwdStage entries, each staged in turn, every pixel shading the part of its list inside the window (a loop
    // uniform over the workgroup, so the barriers are reached by every thread)
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        float acc[3] = {0.0f, 0.0f, 0.0f};
        ShadeSurface sf;
        if (px[q].shaded) sf = shade_surface(fd, px[q], nb[q], ar[q], metal[q], acc);
        for (uint32_t w0 = start; w0 - start < nstage; w0 += kFwdStage) {
            const uint32_t n = min(kFwdStage, nstage - (w0 - start));
            __syncthreads(); // the previous window's readers are done
            if (tid < n) s_light[tid] = stage_light(fd, *((guint*)fd.indices + w0 + tid));
            __syncthreads();
            if (px[q].shaded && has_list[q]) {
                const uint32_t a = max(my_off[q], w0), b = min(my_end[q], w0 + n);
                if (a < b) shade_lights<true>(fd, px[q], sf, a, b, w0, s_light, acc);
            }
        }
        if (px[q].shaded) shade_finish(fd, sf, acc, p[q]);
    }
}


// ---- adjoint ---------------------------------------------------------------------------------------------------

struct LightTerm {
    float L[3];       // unit direction light -> surface
    float dist, r, x, w, att;
    float inv_r, inv_dist;
    bool live;
};

// forward quantities of one (pixel, light) pair shared by the two passes of the adjoint; quotients through the
// hardware reciprocal (1 ulp), |d| in double from one Newton step on the float square root
__device__ __forceinline__ LightTerm light_term(const double P[3], const float4& lp, const float4& lc, uint32_t atten)
{
    LightTerm t;
    t.live = false;
    const double d[3] = {P[0] - (double)lp.x, P[1] - (double)lp.y, P[2] - (double)lp.z};
    const double dist2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    t.r = lc.w;
    if (!(dist2 > 0.0) || !(t.r > 0.0f) || !(dist2 < (double)t.r * (double)t.r)) return t;
    const float s0 = __builtin_amdgcn_sqrtf((float)dist2);
    // |d| = s0 + (|d|^2 - s0^2) / (2 s0): the correction is ~1e-7 s0, so a float reciprocal leaves ~1e-14
    const double dist = (double)s0 + (dist2 - (double)s0 * (double)s0) * (double)(0.5f * frcp(s0));
    t.inv_r = frcp(t.r);
    const float omx = (float)((double)t.r - dist) * t.inv_r;
    t.x = 1.0f - omx;
    if (atten == VREN_ATTEN_SMOOTH) {
        t.w = omx * (1.0f + t.x) * (1.0f + t.x * t.x);
        t.att = t.w * t.w;
    } else {
        t.w = 0.0f;
        t.att = omx * (1.0f + t.x + t.x * t.x) * frcp(1.0f + t.x * __builtin_amdgcn_sqrtf(t.x));
    }
    t.dist = (float)dist;
    t.inv_dist = frcp(t.dist);
    t.L[0] = (float)d[0] * t.inv_dist;
    t.L[1] = (float)d[1] * t.inv_dist;
    t.L[2] = (float)d[2] * t.inv_dist;
    t.live = t.att > 0.0f;
    return t;
}

// Adjoint of one pair's contribution (pbr.glsl:42-93 with the smooth window). gT = dLoss/d(contribution rgb).
// Accumulates d/d(light position), d/d(light rgb), d/d(radius) into gl[7] and d/d(albedo, roughness) into gp[4].
// Every quotient goes through one of four hardware reciprocals (1 / inner, 1 / (n.l (1 - k) + k), 1 / D and the
// per-pixel 1 / (n.v (1 - k) + k)) or the pair's 1 / r and 1 / |d|.
__device__ __forceinline__ void light_adjoint(const Surface& s, const LightTerm& t, const float4& lc, const float gT[3],
                                              float i_gv, float gl[7], float gp[4])
{
    const float mL[3] = {-t.L[0], -t.L[1], -t.L[2]};
    const float NdotL = dot3(s.N, mL);
    if (!(NdotL > 0.0f)) return;
    const float h[3] = {s.V[0] + mL[0], s.V[1] + mL[1], s.V[2] + mL[2]};
    const float hh = dot3(h, h);
    if (!(hh > 0.0f)) return;
    const float ih = __builtin_amdgcn_rsqf(hh);
    const float H[3] = {h[0] * ih, h[1] * ih, h[2] * ih};
    const float Nh = dot3(s.N, h);
    float NdotH = 0.0f, inner = 1.0f;
    if (Nh > 0.0f) {
        const float c[3] = {s.N[1] * h[2] - s.N[2] * h[1], s.N[2] * h[0] - s.N[0] * h[2], s.N[0] * h[1] - s.N[1] * h[0]};
        const float inv_hh = ih * ih;
        NdotH = Nh * ih;
        inner = s.nn + dot3(c, c) * inv_hh + s.a2 * (Nh * Nh) * inv_hh;
    }
    const float i_inner = frcp(inner);
    const float NDF = s.a2 * kInvPi * i_inner * i_inner;
    const float k = s.rough;
    const float gl_den = NdotL * (1.0f - k) + k;
    const float i_gl = frcp(gl_den);
    const float glv = NdotL * i_gl;
    const float G = s.ggxV * glv;
    const float HdotVraw = dot3(h, s.V) * ih;
    const float HdotV = fmaxf(HdotVraw, 0.0f);
    const float om = 1.0f - HdotV;
    const float om2 = om * om;
    const float f5 = om2 * om2 * om;
    const float D = 4.0f * s.NdotV * NdotL + 0.0001f;
    const float i_D = frcp(D);
    const float spec = NDF * G * i_D;
    const float km = 1.0f - s.metal;
    float g_att = 0.0f, g_NdotL = 0.0f, g_spec = 0.0f, g_f5 = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float F = s.F0[c] + (1.0f - s.F0[c]) * f5;
        const float kD = (1.0f - F) * km;
        const float B = kD * s.albedo[c] * kInvPi + F * spec;
        const float lcc = c == 0 ? lc.x : (c == 1 ? lc.y : lc.z);
        gl[3 + c] += gT[c] * B * t.att * NdotL;
        const float gB = gT[c] * lcc * t.att * NdotL;
        g_att += gT[c] * B * lcc * NdotL;
        g_NdotL += gT[c] * B * lcc * t.att;
        g_spec += gB * F;
        const float gF = gB * (spec - s.albedo[c] * km * kInvPi);
        gp[c] += gB * kD * kInvPi + gF * (1.0f - f5) * s.metal;
        g_f5 += gF * (1.0f - s.F0[c]);
    }
    const float g_HdotV = HdotVraw > 0.0f ? g_f5 * (-5.0f * om2 * om2) : 0.0f;
    const float g_NDF = g_spec * G * i_D;
    const float g_G = g_spec * NDF * i_D;
    g_NdotL += -g_spec * spec * i_D * 4.0f * s.NdotV;
    const float g_gl = g_G * s.ggxV;
    const float g_ggxV = g_G * glv;
    g_NdotL += g_gl * k * (i_gl * i_gl);
    float g_k = g_gl * (-NdotL * (1.0f - NdotL) * (i_gl * i_gl)) +
                g_ggxV * (-s.NdotV * (1.0f - s.NdotV) * (i_gv * i_gv));
    const float g_inner = g_NDF * (-2.0f * NDF * i_inner);
    const float g_a2 = g_NDF * kInvPi * i_inner * i_inner + g_inner * NdotH * NdotH;
    const float g_NdotH = Nh > 0.0f ? g_inner * 2.0f * NdotH * (s.a2 - 1.0f) : 0.0f;
    gp[3] += g_a2 * 2.0f * s.rough + g_k;
    // geometry: h = V - L, NdotL = -N.L, NdotH = N.h/|h|, HdotV = V.h/|h|
    float g_L[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float g_h = (g_NdotH * (s.N[i] - NdotH * H[i]) + g_HdotV * (s.V[i] - HdotV * H[i])) * ih;
        g_L[i] = -(g_h + g_NdotL * s.N[i]);
    }
    // att(x), x = |d| / r (smooth window: att = w^2, w = 1 - x^4)
    const float datt_dx = -8.0f * t.x * t.x * t.x * t.w;
    const float g_dist = g_att * datt_dx * t.inv_r;
    const float gLL = dot3(g_L, t.L);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float g_d = (g_L[i] - gLL * t.L[i]) * t.inv_dist + g_dist * t.L[i];
        gl[i] -= g_d; // d = p - light
    }
    gl[6] += g_att * datt_dx * (-t.x * t.inv_r);
}

// Sums each of the 7 values over the wave with 10 shuffles instead of 42: the first three butterfly stages (lane
// bits 5, 4, 3) each keep half of the remaining components, so lane l ends with component (l >> 3) & 7 summed
// over its 8-lane group, and three more stages sum the group. Every lane of the wave must be active.
__device__ __forceinline__ float wave_sum_components(const float g[7])
{
    const uint32_t lane = lane_id();
    const float v[8] = {g[0], g[1], g[2], g[3], g[4], g[5], g[6], 0.0f};
    const bool h32 = (lane & 32u) != 0, h16 = (lane & 16u) != 0, h8 = (lane & 8u) != 0;
    float w4[4], w2[2];
#pragma unroll
    for (int i = 0; i < 4; i++) w4[i] = (h32 ? v[i + 4] : v[i]) + shfl_xor(h32 ? v[i] : v[i + 4], 32);
#pragma unroll
    for (int i = 0; i < 2; i++) w2[i] = (h16 ? w4[i + 2] : w4[i]) + shfl_xor(h16 ? w4[i] : w4[i + 2], 16);
    float w = (h8 ? w2[1] : w2[0]) + shfl_xor(h8 ? w2[0] : w2[1], 8);
    w += shfl_xor(w, 4);
    w += shfl_xor(w, 2);
    w += shfl_xor(w, 1);
    return w;
}

__global__ __launch_bounds__(256) void shade_backward_kernel(const FrameDev* __restrict__ frames, uint32_t F,
                                                             uint32_t total_tiles, const float* grad_out,
                                                             float* g_pos, float* g_light, float* g_ar)
{
    __shared__ float4 s_lpos[kStage];
    __shared__ float4 s_lcol[kStage];
    __shared__ float s_g[kStage * 7];
    __shared__ uint32_t s_li[kStage];
    __shared__ uint32_t s_min[4], s_max[4];
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / VREN_WAVE;
    // a tile's four strips on one XCD, consecutive tiles on consecutive XCDs (as the forward shade), the grid rounded
    // up to whole rounds of 8 tiles
    const uint32_t k = blockIdx.x / 8;
    const uint32_t g = (k / 4) * 8 + blockIdx.x % 8, sub = k % 4;
    if (g >= total_tiles) return;
    const uint32_t f = find_frame_wave(frames, F, g, [](const FrameDev& d) { return d.tile_base; });
    const FrameDev& fd = frames[f];
    const uint32_t t = g - fd.tile_base;
    if (t >= fd.Sx * fd.Sy) return; // padding after the frame's tiles (the whole workgroup)
    const uint32_t ti = t % fd.Sx, tj = t / fd.Sx;
    const uint32_t gx = ti * 32 + (tid & 31), gy = tj * 32 + sub * 8 + (tid >> 5); // 32x8 strip `sub`
    const bool in = gx < fd.W && gy < fd.H;
    const uint64_t p = in ? (uint64_t)gy * fd.W + gx : 0;
    PixelIn px;
    px.shaded = false;
    uint2 nb = make_uint2(0u, 0u);
    float4 ar = make_float4(0.f, 0.f, 0.f, 0.f);
    float metal = 0.0f;
    float4 go = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 gar = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t cl_raw = 0xFFFFFFFFu;
    // every per-pixel load is issued before any of them is used: the reference and the pixel's running material
    // gradient along with the G-buffer, not after the reconstruction and the light loops
    if (in) {
        // global-address-space loads, as in the forward kernel
        const float depth = *((gfloat*)fd.depth + p);
        const vu2 nv = *((gvu2*)fd.normals + p);
        nb = make_uint2(nv.x, nv.y);
        const vf4 av = *((gvf4*)fd.albedo_rough + p);
        ar = make_float4(av.x, av.y, av.z, av.w);
        metal = *((gfloat*)fd.metallic + p);
        const vf4 gv = *((gvf4*)grad_out + p);
        go = make_float4(gv.x, gv.y, gv.z, gv.w);
        cl_raw = *((guint*)fd.refs + p);
        const vf4 gav = *((gvf4*)g_ar + p);
        gar = make_float4(gav.x, gav.y, gav.z, gav.w);
        px = reconstruct(fd, gx, gy, depth);
    }
    const bool has_list = px.shaded && fd.L > 0 && cl_raw < fd.key_cap;
    // the pixel's list [off, end) (clamped to the index capacity), then the workgroup's range of list entries
    // reduced from its pixels' lists, so that every list lies inside the staged range
    uint32_t off = 0, end = 0;
    if (has_list) {
        const uint32_t o = *((guint*)fd.offsets + cl_raw);
        off = min(o, fd.index_cap);
        const uint64_t e = (uint64_t)o + *((guint*)fd.counts + cl_raw);
        end = e > fd.index_cap ? fd.index_cap : (uint32_t)e;
    }
    uint32_t rmin = has_list ? off : 0xFFFFFFFFu, rmax = has_list ? end : 0u;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        rmin = min(rmin, shfl_xor(rmin, m));
        rmax = max(rmax, shfl_xor(rmax, m));
    }
    if (lane == 0) {
        s_min[wave] = rmin;
        s_max[wave] = rmax;
    }
    __syncthreads();
    const uint32_t ra = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
    const uint32_t rb = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
    const uint32_t start = ra == 0xFFFFFFFFu ? 0u : ra;
    const uint32_t nstage = (ra != 0xFFFFFFFFu && rb > ra) ? rb - ra : 0u;
    const bool staged = nstage <= kStage;
    // light records and list entries through global-address-space loads (flat loads would also wait on the LDS
    // counter)
    const auto lpos = [&](uint32_t li) {
        const vf4 v = *((gvf4*)fd.light_pos + li);
        return make_float4(v.x, v.y, v.z, v.w);
    };
    const auto lcol = [&](uint32_t li) {
        const vf4 v = *((gvf4*)fd.lights + li);
        return make_float4(v.x, v.y, v.z, v.w);
    };
    if (staged) {
        for (uint32_t i = tid; i < nstage; i += 256) {
            const uint32_t li = *((guint*)fd.indices + start + i);
            s_li[i] = li;
            const bool ok = li < fd.L;
            s_lpos[i] = ok ? lpos(li) : make_float4(0.f, 0.f, 0.f, 0.f);
            s_lcol[i] = ok ? lcol(li) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        for (uint32_t i = tid; i < nstage * 7; i += 256) s_g[i] = 0.0f;
    }
    __syncthreads();
    Surface s;
    uint32_t len = 0;
    float i_gv = 0.0f;
    float gT[3] = {0.0f, 0.0f, 0.0f};
    float gp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (px.shaded) {
        {
            const float V[3] = {(float)(fd.eye[0] - px.P[0]), (float)(fd.eye[1] - px.P[1]), (float)(fd.eye[2] - px.P[2])};
            const float il = __builtin_amdgcn_rsqf(dot3(V, V)); // |V| is scene-sized, never subnormal
            for (int i = 0; i < 3; i++) s.V[i] = V[i] * il;
        }
        s.N[0] = normal_h2f(nb.x & 0xFFFFu);
        s.N[1] = normal_h2f(nb.x >> 16);
        s.N[2] = normal_h2f(nb.y & 0xFFFFu);
        s.nn = (float)(1.0 - ((double)s.N[0] * s.N[0] + (double)s.N[1] * s.N[1] + (double)s.N[2] * s.N[2]));
        s.albedo[0] = ar.x;
        s.albedo[1] = ar.y;
        s.albedo[2] = ar.z;
        s.rough = ar.w;
        s.metal = metal;
        s.a2 = s.rough * s.rough;
        for (int c = 0; c < 3; c++) s.F0[c] = 0.04f * (1.0f - s.metal) + s.albedo[c] * s.metal;
        s.NdotV = fmaxf(dot3(s.N, s.V), 0.0f);
        i_gv = frcp(s.NdotV * (1.0f - s.rough) + s.rough);
        s.ggxV = s.NdotV * i_gv;
        // pass 1: the pixel's colour (the gamma derivative needs it)
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t e = off; e < end; e++) {
            uint32_t li;
            float4 lp, lc;
            if (staged) {
                lp = s_lpos[e - start];
                lc = s_lcol[e - start];
            } else {
                li = *((guint*)fd.indices + e);
                if (li >= fd.L) continue;
                lp = lpos(li);
                lc = lcol(li);
            }
            const LightTerm lt = light_term(px.P, lp, lc, fd.attenuation);
            if (!lt.live) continue;
            const float rad[3] = {lc.x * lt.att, lc.y * lt.att, lc.z * lt.att};
            apply_light(s, lt.L, rad, acc);
        }
        const float gexp = 1.0f / 2.2f;
        const float goc[3] = {go.x, go.y, go.z};
        for (int c = 0; c < 3; c++) {
            // col^(1/2.2 - 1) as 2^(log2(col) (1/2.2 - 1)) on the transcendental unit (~1e-6 relative)
            const float col = 0.03f * s.albedo[c] + acc[c];
            gT[c] = goc[c] * gexp * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(col) * (gexp - 1.0f));
            gp[c] += 0.03f * gT[c];
        }
        len = end > off ? end - off : 0u;
    }
    // pass 2: adjoint of every contribution, in a loop that is uniform over the wave (iteration k handles entry
    // off + k of every lane whose list is that long), so that when every such lane is at the same entry -- the
    // lanes of one cluster -- the wave sums its 7 light-gradient terms and one lane adds them to the staged slot
    uint32_t kmax = len;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) kmax = max(kmax, shfl_xor(kmax, m));
    kmax = __builtin_amdgcn_readfirstlane(kmax);
    for (uint32_t k = 0; k < kmax; k++) {
        const bool act = k < len;
        const uint32_t e = off + k;
        uint32_t li = 0xFFFFFFFFu;
        float gl[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        bool live = false;
        if (act) {
            float4 lp, lc;
            bool ok = true;
            if (staged) {
                lp = s_lpos[e - start];
                lc = s_lcol[e - start];
            } else {
                li = *((guint*)fd.indices + e);
                ok = li < fd.L;
                if (ok) {
                    lp = lpos(li);
                    lc = lcol(li);
                }
            }
            if (ok) {
                const LightTerm lt = light_term(px.P, lp, lc, fd.attenuation);
                if (lt.live) {
                    light_adjoint(s, lt, lc, gT, i_gv, gl, gp);
                    live = true;
                }
            }
        }
        const uint64_t lm = ballot(live);
        if (lm == 0) continue;
        if (staged) {
            const uint32_t first = (uint32_t)__builtin_ctzll(lm);
            const uint32_t ef = __builtin_amdgcn_readlane(e, first);
            if (ballot(live && e != ef) == 0) {
                const float sum = wave_sum_components(gl);
                const uint32_t c = (lane >> 3) & 7u;
                if ((lane & 7u) == 0 && c < 7 && sum != 0.0f) atomicAdd(&s_g[(ef - start) * 7 + c], sum);
            } else if (live) {
                float* slot = s_g + (e - start) * 7;
#pragma unroll
                for (int i = 0; i < 7; i++) atomicAdd(&slot[i], gl[i]);
            }
        } else if (live) {
#pragma unroll
            for (int i = 0; i < 3; i++) atomicAdd(&g_pos[4 * li + i], gl[i]);
#pragma unroll
            for (int i = 0; i < 4; i++) atomicAdd(&g_light[4 * li + i], gl[3 + i]);
        }
    }
    if (px.shaded) {
        // only this thread touches pixel p's material gradient, so the value loaded with the pixel is current
        gar.x += gp[0];
        gar.y += gp[1];
        gar.z += gp[2];
        gar.w += gp[3];
        reinterpret_cast<float4*>(g_ar)[p] = gar;
    }
    __syncthreads();
    if (staged) {
        for (uint32_t i = tid; i < nstage; i += 256) {
            const uint32_t li = s_li[i];
            if (li >= fd.L) continue;
            const float* slot = s_g + i * 7;
            for (int c = 0; c < 3; c++)
                if (slot[c] != 0.0f) atomicAdd(&g_pos[4 * li + c], slot[c]);
            for (int c = 0; c < 4; c++)
                if (slot[3 + c] != 0.0f) atomicAdd(&g_light[4 * li + c], slot[3 + c]);
        }
    }
}

} // namespace

int launch_shade(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w)
{
    if (w.tile_range) {
        tile_range_kernel<<<(w.total_tiles + 255) / 256, 256, 0, s>>>(d_frames, F > 1 ? w.tile_frame : nullptr,
                                                                    w.total_tiles, w.tile_cl, w.tile_range);
        VREN_LAUNCH_CHECK(ctx);
    }
    const uint32_t* tf = F > 1 ? w.tile_frame : nullptr;
    const uint32_t rounds = (w.total_tiles + 7) / 8; // rounds of 8 tiles, one per XCD
    if (w.shade_ppt == 2) {
        const uint32_t grid = rounds * 8 * 2;
        if (w.any_materials) shade_kernel<2, true><<<grid, 256, 0, s>>>(d_frames, tf, w.tile_range, w.total_tiles);
        else shade_kernel<2, false><<<grid, 256, 0, s>>>(d_frames, tf, w.tile_range, w.total_tiles);
    } else {
        const uint32_t grid = rounds * 8 * (4 / kShadePPT);
        if (w.any_materials)
            shade_kernel<kShadePPT, true><<<grid, 256, 0, s>>>(d_frames, tf, w.tile_range, w.total_tiles);
        else shade_kernel<kShadePPT, false><<<grid, 256, 0, s>>>(d_frames, tf, w.tile_range, w.total_tiles);
    }
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    return VREN_OK;
}

extern "C" VREN_EXPORT int vren_shade(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                      const vren_camera* camera, const vren_cluster_cfg* cfg, const float* depth,
                                      const uint16_t* normals, const float* albedo_roughness, const float* metallic,
                                      const uint32_t* cluster_refs, uint32_t key_capacity,
                                      const uint32_t* light_counts, const uint32_t* light_offsets,
                                      const uint32_t* light_indices, uint32_t index_capacity,
                                      const float* light_positions, const float* point_lights, uint32_t light_count,
                                      const float* directional_lights, uint32_t directional_light_count,
                                      float* output)
{
    if (!ctx) return VREN_EINVAL;
    int st = vren_validate_camera(ctx, camera);
    if (!st) st = vren_validate_cfg(ctx, cfg);
    if (st) return st;
    if (width == 0 || height == 0 || width > 8192 || height > 8192)
        return vren_fail(ctx, VREN_EINVAL, "shade: screen must be in [1, 8192]^2");
    if (!depth || !normals || !albedo_roughness || !metallic || !cluster_refs || !output)
        return vren_fail(ctx, VREN_EINVAL, "shade: null G-buffer, reference or output buffer");
    if (light_count > 0 && (!light_counts || !light_offsets || !light_indices || !light_positions || !point_lights))
        return vren_fail(ctx, VREN_EINVAL, "shade: null light buffer");
    if (directional_light_count > 0 && !directional_lights)
        return vren_fail(ctx, VREN_EINVAL, "shade: null directional light buffer");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "shade: context has no device");
    hipStream_t s = (hipStream_t)stream;
    FrameDev f;
    memset(&f, 0, sizeof(f));
    vren_frame_setup(&f, width, height, camera, cfg);
    f.depth = depth;
    f.normals = normals;
    f.albedo_rough = albedo_roughness;
    f.metallic = metallic;
    f.refs = const_cast<uint32_t*>(cluster_refs);
    f.key_cap = key_capacity;
    f.counts = const_cast<uint32_t*>(light_counts);
    f.offsets = const_cast<uint32_t*>(light_offsets);
    f.indices = const_cast<uint32_t*>(light_indices);
    f.index_cap = index_capacity;
    f.light_pos = light_positions;
    f.lights = point_lights;
    f.L = light_count;
    f.dir_lights = directional_lights;
    f.n_dir = directional_light_count;
    f.out = output;
    FrameDev* slot = vren_next_frame_slot(ctx);
    st = vren_upload_frames(ctx, s, slot, &f, 1);
    ChainWork w;
    memset(&w, 0, sizeof(w));
    w.total_tiles = f.Sx * f.Sy;
    if (!st) st = launch_shade(ctx, s, slot, 1, w);
    return st;
}

extern "C" VREN_EXPORT int vren_shade_materials(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                                const vren_camera* camera, const vren_cluster_cfg* cfg,
                                                const float* depth, const uint16_t* normals,
                                                const vren_material_set* materials, const uint32_t* cluster_refs,
                                                uint32_t key_capacity, const uint32_t* light_counts,
                                                const uint32_t* light_offsets, const uint32_t* light_indices,
                                                uint32_t index_capacity, const float* light_positions,
                                                const float* point_lights, uint32_t light_count,
                                                const float* directional_lights, uint32_t directional_light_count,
                                                float* output)
{
    if (!ctx) return VREN_EINVAL;
    int st = vren_validate_camera(ctx, camera);
    if (!st) st = vren_validate_cfg(ctx, cfg);
    if (st) return st;
    if (width == 0 || height == 0 || width > 8192 || height > 8192)
        return vren_fail(ctx, VREN_EINVAL, "shade_materials: screen must be in [1, 8192]^2");
    if (!depth || !normals || !materials || !cluster_refs || !output)
        return vren_fail(ctx, VREN_EINVAL, "shade_materials: null G-buffer, material set, reference or output");
    st = vren_validate_material_set(ctx, materials, "shade_materials");
    if (st) return st;
    if (light_count > 0 && (!light_counts || !light_offsets || !light_indices || !light_positions || !point_lights))
        return vren_fail(ctx, VREN_EINVAL, "shade_materials: null light buffer");
    if (directional_light_count > 0 && !directional_lights)
        return vren_fail(ctx, VREN_EINVAL, "shade_materials: null directional light buffer");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "shade_materials: context has no device");
    hipStream_t s = (hipStream_t)stream;
    FrameDev f;
    memset(&f, 0, sizeof(f));
    vren_frame_setup(&f, width, height, camera, cfg);
    f.depth = depth;
    f.normals = normals;
    vren_frame_set_materials(&f, materials);
    f.refs = const_cast<uint32_t*>(cluster_refs);
    f.key_cap = key_capacity;
    f.counts = const_cast<uint32_t*>(light_counts);
    f.offsets = const_cast<uint32_t*>(light_offsets);
    f.indices = const_cast<uint32_t*>(light_indices);
    f.index_cap = index_capacity;
    f.light_pos = light_positions;
    f.lights = point_lights;
    f.L = light_count;
    f.dir_lights = directional_lights;
    f.n_dir = directional_light_count;
    f.out = output;
    FrameDev* slot = vren_next_frame_slot(ctx);
    st = vren_upload_frames(ctx, s, slot, &f, 1);
    ChainWork w;
    memset(&w, 0, sizeof(w));
    w.total_tiles = f.Sx * f.Sy;
    w.any_materials = 1;
    if (!st) st = launch_shade(ctx, s, slot, 1, w);
    return st;
}

extern "C" VREN_EXPORT int vren_shade_backward(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                               const vren_camera* camera, const vren_cluster_cfg* cfg,
                                               const float* depth, const uint16_t* normals,
                                               const float* albedo_roughness, const float* metallic,
                                               const uint32_t* cluster_refs, uint32_t key_capacity,
                                               const uint32_t* light_counts, const uint32_t* light_offsets,
                                               const uint32_t* light_indices, uint32_t index_capacity,
                                               const float* light_positions, const float* point_lights,
                                               uint32_t light_count, const float* grad_output,
                                               float* grad_light_positions, float* grad_point_lights,
                                               float* grad_albedo_roughness)
{
    if (!ctx) return VREN_EINVAL;
    int st = vren_validate_camera(ctx, camera);
    if (!st) st = vren_validate_cfg(ctx, cfg);
    if (st) return st;
    if (width == 0 || height == 0 || width > 8192 || height > 8192)
        return vren_fail(ctx, VREN_EINVAL, "shade_backward: screen must be in [1, 8192]^2");
    if (!depth || !normals || !albedo_roughness || !metallic || !cluster_refs || !grad_output || !grad_albedo_roughness)
        return vren_fail(ctx, VREN_EINVAL, "shade_backward: null G-buffer, reference or gradient buffer");
    if (light_count > 0 && (!light_counts || !light_offsets || !light_indices || !light_positions || !point_lights ||
                            !grad_light_positions || !grad_point_lights))
        return vren_fail(ctx, VREN_EINVAL, "shade_backward: null light buffer");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "shade_backward: context has no device");
    hipStream_t s = (hipStream_t)stream;
    FrameDev f;
    memset(&f, 0, sizeof(f));
    vren_frame_setup(&f, width, height, camera, cfg);
    f.depth = depth;
    f.normals = normals;
    f.albedo_rough = albedo_roughness;
    f.metallic = metallic;
    f.refs = const_cast<uint32_t*>(cluster_refs);
    f.key_cap = key_capacity;
    f.counts = const_cast<uint32_t*>(light_counts);
    f.offsets = const_cast<uint32_t*>(light_offsets);
    f.indices = const_cast<uint32_t*>(light_indices);
    f.index_cap = index_capacity;
    f.light_pos = light_positions;
    f.lights = point_lights;
    f.L = light_count;
    FrameDev* slot = vren_next_frame_slot(ctx);
    st = vren_upload_frames(ctx, s, slot, &f, 1);
    if (st) return st;
    const uint32_t tiles = f.Sx * f.Sy;
    shade_backward_kernel<<<(tiles + 7) / 8 * 32, 256, 0, s>>>(slot, 1, tiles, grad_output, grad_light_positions,
                                                               grad_point_lights, grad_albedo_roughness);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}
