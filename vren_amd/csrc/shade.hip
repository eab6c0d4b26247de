// shade.hip -- shade (clustered_shading.cpp:709-811, shade.comp:69-141, pbr.glsl:6-112) with resolved per-pixel
// materials, for gfx950.
//
// One thread per pixel, one 1024-thread workgroup per 32x32 tile, flattened over every frame of a batch. The
// reference evaluates everything in fp32, including inverse(P) and inverse(V) per invocation; that formula loses
// ~1e-4 relative accuracy wherever a light sits close to the surface or a low roughness makes the GGX
// denominator cancel. This kernel keeps the reference's light model (constants PI = 3.14, EPSILON = 0.001,
// a = roughness, k = roughness, +0.0001 in the specular denominator, ambient 0.03 * albedo, gamma 1/2.2) and
// reorganises the arithmetic so the fp32 result stays within ~1e-6 of an FP64 evaluation:
//   - the pixel's world position is reconstructed in double (per pixel), and d = p - light, |d| and the
//     attenuation are evaluated in double per light (the attenuation cancels near the light's radius);
//   - the GGX denominator n.h^2 (a^2 - 1) + 1 is evaluated as (1 - |n|^2) + |n x h|^2 / |h|^2 + a^2 (n.h)^2 / |h|^2
//     with the unnormalised half vector h = v - l, a sum of non-negative terms (no cancellation near n.h = 1);
//   - contributions that are exactly zero in the reference formula (attenuation 0 or n.l <= 0) are skipped.
// The background test (view depth >= far - 0.001) uses the double reconstruction and is exactly the oracle's.
#include "chain.hpp"

namespace {

constexpr float kPi = 3.14f;

struct PixelIn {
    double P[3];
    bool shaded;
};

__device__ __forceinline__ PixelIn reconstruct(const FrameDev& fd, uint32_t gx, uint32_t gy, float depth)
{
#pragma clang fp contract(off)
    PixelIn o;
    double v[4] = {(double)gx / (double)fd.W, (double)gy / (double)fd.H, (double)depth, 1.0};
    v[1] = 1.0 - v[1];
    v[0] = v[0] * 2.0 - 1.0;
    v[1] = v[1] * 2.0 - 1.0;
    double r[4];
    for (int i = 0; i < 4; i++)
        r[i] = ((fd.inv_proj_d[0 * 4 + i] * v[0] + fd.inv_proj_d[1 * 4 + i] * v[1]) + fd.inv_proj_d[2 * 4 + i] * v[2]) +
               fd.inv_proj_d[3 * 4 + i] * v[3];
    const double w = r[3];
    for (int i = 0; i < 4; i++) r[i] = r[i] / w;
    o.shaded = r[2] < (double)fd.bg_threshold;
    for (int i = 0; i < 3; i++)
        o.P[i] = ((fd.inv_view_d[0 * 4 + i] * r[0] + fd.inv_view_d[1 * 4 + i] * r[1]) + fd.inv_view_d[2 * 4 + i] * r[2]) +
                 fd.inv_view_d[3 * 4 + i] * r[3];
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
    float inner = 1.0f;
    if (Nh > 0.0f && hh > 0.0f) {
        const float c[3] = {s.N[1] * h[2] - s.N[2] * h[1], s.N[2] * h[0] - s.N[0] * h[2], s.N[0] * h[1] - s.N[1] * h[0]};
        const float inv_hh = 1.0f / hh;
        inner = s.nn + dot3(c, c) * inv_hh + s.a2 * (Nh * Nh) * inv_hh;
    }
    const float NDF = s.a2 / (kPi * inner * inner);
    const float k = s.rough;
    const float G = s.ggxV * (NdotL / (NdotL * (1.0f - k) + k));
    const float HdotV = hh > 0.0f ? fmaxf(dot3(h, s.V) * rsqrtf(hh), 0.0f) : 0.0f;
    float om = 1.0f - HdotV;
    float om2 = om * om;
    const float f5 = om2 * om2 * om;
    const float spec = (NDF * G) / (4.0f * s.NdotV * NdotL + 0.0001f);
    const float km = 1.0f - s.metal;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float F = s.F0[c] + (1.0f - s.F0[c]) * f5;
        const float kD = (1.0f - F) * km;
        acc[c] += fmaxf((kD * s.albedo[c] / kPi + F * spec) * rad[c] * NdotL, 0.0f);
    }
}

__global__ __launch_bounds__(1024) void shade_kernel(const FrameDev* frames, uint32_t F)
{
    const uint32_t g = blockIdx.x;
    const uint32_t f = find_frame(frames, F, g, [](const FrameDev& d) { return d.tile_base; });
    const FrameDev& fd = frames[f];
    const uint32_t t = g - fd.tile_base;
    const uint32_t ti = t % fd.Sx, tj = t / fd.Sx;
    const uint32_t gx = ti * 32 + (threadIdx.x & 31), gy = tj * 32 + (threadIdx.x >> 5);
    if (gx >= fd.W || gy >= fd.H) return;
    const uint64_t p = (uint64_t)gy * fd.W + gx;
    const PixelIn px = reconstruct(fd, gx, gy, fd.depth[p]);
    if (!px.shaded) return;
    Surface s;
    double V[3], vl = 0.0;
    for (int i = 0; i < 3; i++) {
        V[i] = fd.eye[i] - px.P[i];
        vl += V[i] * V[i];
    }
    vl = sqrt(vl);
    for (int i = 0; i < 3; i++) s.V[i] = (float)(V[i] / vl);
    const uint2 nb = reinterpret_cast<const uint2*>(fd.normals)[p];
    s.N[0] = vren_half_to_float((uint16_t)(nb.x & 0xFFFFu));
    s.N[1] = vren_half_to_float((uint16_t)(nb.x >> 16));
    s.N[2] = vren_half_to_float((uint16_t)(nb.y & 0xFFFFu));
    s.nn = (float)(1.0 - ((double)s.N[0] * s.N[0] + (double)s.N[1] * s.N[1] + (double)s.N[2] * s.N[2]));
    const float4 ar = reinterpret_cast<const float4*>(fd.albedo_rough)[p];
    s.albedo[0] = ar.x;
    s.albedo[1] = ar.y;
    s.albedo[2] = ar.z;
    s.rough = ar.w;
    s.metal = fd.metallic[p];
    s.a2 = s.rough * s.rough;
    for (int c = 0; c < 3; c++) s.F0[c] = 0.04f * (1.0f - s.metal) + s.albedo[c] * s.metal;
    s.NdotV = fmaxf(dot3(s.N, s.V), 0.0f);
    s.ggxV = s.NdotV / (s.NdotV * (1.0f - s.rough) + s.rough);

    float acc[3] = {0.0f, 0.0f, 0.0f};
    const uint32_t cl = fd.refs[p];
    if (fd.L > 0 && cl < fd.key_cap) {
        const uint32_t cnt = fd.counts[cl];
        const uint32_t off = fd.offsets[cl];
        const uint32_t end = (uint64_t)off + cnt > fd.index_cap ? fd.index_cap : off + cnt;
        const float4* lpos = reinterpret_cast<const float4*>(fd.light_pos);
        const float4* lcol = reinterpret_cast<const float4*>(fd.lights);
        for (uint32_t e = off; e < end; e++) {
            const uint32_t li = fd.indices[e];
            if (li >= fd.L) continue;
            const float4 lp = lpos[li];
            const float4 lc = lcol[li];
            // d = p - light (pbr.glsl:87), its length and the attenuation in double: near the edge of the light's
            // radius 1 - (d / r)^1.5 (or 1 - (d / r)^4) cancels, and a float distance would leave ~1e-7 / (1 - d / r)
            // relative error on a contribution that can dominate the pixel
            const double d[3] = {px.P[0] - (double)lp.x, px.P[1] - (double)lp.y, px.P[2] - (double)lp.z};
            const double dist2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            if (!(dist2 > 0.0)) continue;
            const double dist = sqrt(dist2);
            const double q = dist / (double)lc.w;
            double attd;
            if (fd.attenuation == VREN_ATTEN_SMOOTH) {
                const double q2 = q * q;
                const double w = 1.0 - q2 * q2;
                attd = w > 0.0 ? w * w : 0.0;
            } else {
                const double x = q < 1.0 ? q : 1.0;
                attd = 1.0 - x * sqrt(x);
            }
            if (!(attd > 0.0)) continue;
            const double inv = 1.0 / dist;
            const float L[3] = {(float)(d[0] * inv), (float)(d[1] * inv), (float)(d[2] * inv)};
            const float att = (float)attd;
            const float rad[3] = {lc.x * att, lc.y * att, lc.z * att};
            apply_light(s, L, rad, acc);
        }
    }
    for (uint32_t i = 0; i < fd.n_dir; i++) {
        const float4* D = reinterpret_cast<const float4*>(fd.dir_lights) + 2 * i;
        const float4 dir = D[0], col = D[1];
        const float L[3] = {dir.x, dir.y, dir.z};
        const float rad[3] = {col.x, col.y, col.z};
        apply_light(s, L, rad, acc);
    }
    const float gexp = 1.0f / 2.2f;
    float4 o;
    o.x = powf(0.03f * s.albedo[0] + acc[0], gexp);
    o.y = powf(0.03f * s.albedo[1] + acc[1], gexp);
    o.z = powf(0.03f * s.albedo[2] + acc[2], gexp);
    o.w = 1.0f;
    reinterpret_cast<float4*>(fd.out)[p] = o;
}

} // namespace

int launch_shade(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w)
{
    shade_kernel<<<w.total_tiles, 1024, 0, s>>>(d_frames, F);
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
