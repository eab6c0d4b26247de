// synth.hip -- device generators of the synthetic inputs defined in include/vren_hip/synth.h (lights, procedural
// box scene depth/normals, per-pixel materials). Generating on the device keeps multi-frame 4K batches cheap to
// set up; the CPU oracle evaluates the same definitions and the tests compare the bytes.
#include <cmath>
#include <cstring>

#include "../../include/vren_hip/synth.h"
#include "chain.hpp"

namespace {

struct PixelParams {
    double inv_view[16];
    double tan_half, aspect, near_plane, far_plane, proj_c, proj_d;
    uint32_t W, H, n_boxes;
};

__global__ void synth_lights_kernel(uint64_t initstate, uint32_t L, float4* pos, float4* lights)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    float p[4], l[4];
    vren_synth_light(initstate, i, p, l);
    pos[i] = make_float4(p[0], p[1], p[2], p[3]);
    lights[i] = make_float4(l[0], l[1], l[2], l[3]);
}

__global__ void synth_boxes_kernel(uint64_t initstate, PixelParams pp, double* boxes)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= pp.n_boxes) return;
    vren_synth_box(initstate, b, pp.inv_view, pp.tan_half, pp.aspect, boxes + 3 * b, boxes + 3 * (pp.n_boxes + b));
}

__global__ __launch_bounds__(256) void synth_pixels_kernel(PixelParams pp, const double* boxes, float* depth,
                                                           uint16_t* normals)
{
    __shared__ double s_c[3 * VREN_SYNTH_MAX_BOXES];
    __shared__ double s_h[3 * VREN_SYNTH_MAX_BOXES];
    for (uint32_t i = threadIdx.x; i < 3 * pp.n_boxes; i += 256) {
        s_c[i] = boxes[i];
        s_h[i] = boxes[3 * pp.n_boxes + i];
    }
    __syncthreads();
    const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (uint64_t)pp.W * pp.H) return;
    const uint32_t x = (uint32_t)(p % pp.W), y = (uint32_t)(p / pp.W);
    uint16_t n[4];
    float d;
    vren_synth_pixel(x, y, pp.W, pp.H, pp.inv_view, pp.tan_half, pp.aspect, pp.near_plane, pp.far_plane, pp.proj_c,
                     pp.proj_d, s_c, s_h, pp.n_boxes, &d, n);
    depth[p] = d;
    reinterpret_cast<uint2*>(normals)[p] = make_uint2((uint32_t)n[0] | ((uint32_t)n[1] << 16), (uint32_t)n[2] | ((uint32_t)n[3] << 16));
}

__global__ void synth_materials_kernel(uint64_t initstate, uint64_t n, float4* ar, float* metal)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float a[4], m;
    vren_synth_material(initstate, p, a, &m);
    ar[p] = make_float4(a[0], a[1], a[2], a[3]);
    metal[p] = m;
}

// PCG32 draws [chunk * 64, chunk * 64 + 64) of stream (initstate, initseq): one jump-ahead per 64 draws
__global__ void synth_keys_kernel(uint64_t initstate, uint64_t initseq, uint64_t n, uint32_t* out)
{
    const uint64_t chunk = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t start = chunk * 64;
    if (start >= n) return;
    vren_pcg32 r = vren_pcg32_at(initstate, initseq, start);
    const uint64_t end = start + 64 < n ? start + 64 : n;
    for (uint64_t i = start; i < end; i++) out[i] = vren_pcg32_next(&r);
}

} // namespace

extern "C" {

VREN_EXPORT int vren_synth_keys(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint64_t initseq,
                                uint64_t n, uint32_t* out)
{
    if (!ctx) return VREN_EINVAL;
    if (n == 0) return VREN_OK;
    if (!out) return vren_fail(ctx, VREN_EINVAL, "synth_keys: null buffer");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "synth_keys: context has no device");
    const uint64_t chunks = (n + 63) / 64;
    synth_keys_kernel<<<(uint32_t)((chunks + 255) / 256), 256, 0, (hipStream_t)stream>>>(initstate, initseq, n, out);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

VREN_EXPORT int vren_synth_lights(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint32_t light_count,
                                  float* light_positions, float* point_lights)
{
    if (!ctx) return VREN_EINVAL;
    if (light_count == 0) return VREN_OK;
    if (!light_positions || !point_lights) return vren_fail(ctx, VREN_EINVAL, "synth_lights: null buffer");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "synth_lights: context has no device");
    synth_lights_kernel<<<(light_count + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        initstate, light_count, reinterpret_cast<float4*>(light_positions), reinterpret_cast<float4*>(point_lights));
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

VREN_EXPORT int vren_synth_frame(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint32_t width,
                                 uint32_t height, const vren_camera* camera, uint32_t n_boxes, float* depth,
                                 uint16_t* normals)
{
    if (!ctx) return VREN_EINVAL;
    int st = vren_validate_camera(ctx, camera);
    if (st) return st;
    if (width == 0 || height == 0 || !depth || !normals || n_boxes > VREN_SYNTH_MAX_BOXES)
        return vren_fail(ctx, VREN_EINVAL, "synth_frame: empty frame, null buffer or more than %u boxes",
                         VREN_SYNTH_MAX_BOXES);
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "synth_frame: context has no device");
    hipStream_t s = (hipStream_t)stream;
    vren_cluster_cfg cfg;
    vren_cluster_cfg_default(&cfg);
    FrameDev f;
    memset(&f, 0, sizeof(f));
    vren_frame_setup(&f, width, height, camera, &cfg);
    float view[16], proj[16];
    vren_camera_matrices(camera, view, proj);
    PixelParams pp;
    memcpy(pp.inv_view, f.inv_view_d, sizeof(pp.inv_view));
    pp.tan_half = (double)(float)std::tan((double)(camera->fov_y / 2.0f));
    pp.aspect = (double)camera->aspect_ratio;
    pp.near_plane = (double)camera->near_plane;
    pp.far_plane = (double)camera->far_plane;
    pp.proj_c = (double)proj[10];
    pp.proj_d = (double)proj[14];
    pp.W = width;
    pp.H = height;
    pp.n_boxes = n_boxes;
    st = vren_ensure_scratch(ctx, sizeof(double) * 6 * (n_boxes + 1), s);
    if (st) return st;
    double* boxes = static_cast<double*>(ctx->scratch);
    if (n_boxes) {
        synth_boxes_kernel<<<1, 256, 0, s>>>(initstate, pp, boxes);
        VREN_LAUNCH_CHECK(ctx);
    }
    const uint64_t px = (uint64_t)width * height;
    synth_pixels_kernel<<<(uint32_t)((px + 255) / 256), 256, 0, s>>>(pp, boxes, depth, normals);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

VREN_EXPORT int vren_synth_materials(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint64_t n_pixels,
                                     float* albedo_roughness, float* metallic)
{
    if (!ctx) return VREN_EINVAL;
    if (n_pixels == 0) return VREN_OK;
    if (!albedo_roughness || !metallic) return vren_fail(ctx, VREN_EINVAL, "synth_materials: null buffer");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "synth_materials: context has no device");
    synth_materials_kernel<<<(uint32_t)((n_pixels + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        initstate, n_pixels, reinterpret_cast<float4*>(albedo_roughness), metallic);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

} // extern "C"
