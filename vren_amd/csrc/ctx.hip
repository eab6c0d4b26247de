// ctx.hip -- context lifetime, status reporting and the host-only helpers of the C ABI.
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "clustered.hpp"
#include "common.hpp"

int vren_fail(vren_ctx* ctx, int status, const char* fmt, ...)
{
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->last_error, sizeof(ctx->last_error), fmt, ap);
        va_end(ap);
    }
    return status;
}

int vren_hip_fail(vren_ctx* ctx, hipError_t err, const char* what)
{
    return vren_fail(ctx, err == hipErrorOutOfMemory ? VREN_ENOMEM : VREN_EHIP, "%s: %s", what,
                     hipGetErrorString(err));
}

// The context's scratch serves one stream at a time (a context is used from one stream, like the reference's
// single command buffer per recording thread). Growing it synchronises the device and frees the old buffer, which
// must not happen while `s` is being captured: the capture would break, and a graph captured earlier would keep a
// pointer into freed memory. A capture therefore needs the scratch already grown by an earlier call of the same
// size outside the capture.
int vren_ensure_scratch(vren_ctx* ctx, size_t bytes, hipStream_t s)
{
    if (bytes <= ctx->scratch_bytes) return VREN_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    VREN_HIP(ctx, hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return vren_fail(ctx, VREN_EINVAL,
                         "context scratch must grow to %zu bytes during a stream capture; make the same call once "
                         "outside the capture first",
                         bytes);
    if (ctx->scratch) {
        // a grown scratch replaces the old one only after the device is done with it
        VREN_HIP(ctx, hipDeviceSynchronize());
        VREN_HIP(ctx, hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t sz = bytes < (1u << 20) ? (1u << 20) : bytes;
    VREN_HIP(ctx, hipMalloc(&ctx->scratch, sz));
    ctx->scratch_bytes = sz;
    return VREN_OK;
}

extern "C" {

VREN_EXPORT int vren_create(int hip_device, vren_ctx** out)
{
    if (!out) return VREN_EINVAL;
    *out = nullptr;
    if (hip_device == -1) {
        // host-only context: argument validation and the host helpers work, every device call returns EINVAL
        vren_ctx* ctx = new vren_ctx();
        ctx->device = -1;
        *out = ctx;
        return VREN_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || hip_device < 0 || hip_device >= count) return VREN_EINVAL;
    vren_ctx* ctx = new vren_ctx();
    ctx->device = hip_device;
    hipError_t e = hipSetDevice(hip_device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, hip_device);
    if (e == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    if (e == hipSuccess) e = hipMalloc(&ctx->d_flags, 64);
    if (e == hipSuccess) e = hipMemset(ctx->d_flags, 0, 64);
    if (e == hipSuccess) e = hipMalloc(&ctx->frame_slots, sizeof(FrameDev) * VREN_FRAME_SLOTS);
    if (e == hipSuccess && vren_noalloc_scan_setup(ctx) != VREN_OK) e = hipErrorInvalidSymbol;
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        vren_destroy(ctx);
        return e == hipErrorOutOfMemory ? VREN_ENOMEM : VREN_EHIP;
    }
    *out = ctx;
    return VREN_OK;
}

VREN_EXPORT void vren_destroy(vren_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->device < 0) {
        delete ctx;
        return;
    }
    vren_teardown(hipSetDevice(ctx->device));
    vren_teardown(hipDeviceSynchronize());
    if (ctx->d_flags) vren_teardown(hipFree(ctx->d_flags));
    if (ctx->scratch) vren_teardown(hipFree(ctx->scratch));
    if (ctx->frame_slots) vren_teardown(hipFree(ctx->frame_slots));
    delete ctx;
}

VREN_EXPORT const char* vren_status_string(int status)
{
    switch (status) {
    case VREN_OK: return "ok";
    case VREN_EINVAL: return "invalid argument";
    case VREN_ECAPACITY: return "output capacity exceeded";
    case VREN_EHIP: return "HIP runtime error";
    case VREN_ENOMEM: return "out of device memory";
    default: return "unknown status";
    }
}

VREN_EXPORT const char* vren_last_error(vren_ctx* ctx) { return ctx ? ctx->last_error : ""; }

VREN_EXPORT int vren_sync_status(vren_ctx* ctx, vren_stream_t stream)
{
    if (!ctx) return VREN_EINVAL;
    if (ctx->device < 0) return VREN_OK;
    hipStream_t s = (hipStream_t)stream;
    uint32_t flags = 0;
    VREN_HIP(ctx, hipMemcpyAsync(&flags, ctx->d_flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    VREN_HIP(ctx, hipStreamSynchronize(s));
    if (flags != 0) {
        VREN_HIP(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(uint32_t), s));
        VREN_HIP(ctx, hipStreamSynchronize(s));
        return vren_fail(ctx, VREN_ECAPACITY, "device capacity flags 0x%x (%s%s)", flags,
                         (flags & VREN_FLAG_KEY_CAPACITY) ? "unique cluster keys " : "",
                         (flags & VREN_FLAG_INDEX_CAPACITY) ? "assigned light indices" : "");
    }
    return VREN_OK;
}

VREN_EXPORT void vren_cluster_cfg_default(vren_cluster_cfg* cfg)
{
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->key_mode = VREN_KEYS_REFERENCE;
    cfg->depth_slices = 32;
    cfg->cluster_aabb_mode = VREN_CLUSTER_AABB_REFERENCE;
    cfg->morton_bits = 5;
    cfg->morton_aabb_mode = VREN_MORTON_LIGHTS_AABB;
    cfg->attenuation = VREN_ATTEN_REFERENCE;
}

// build_bvh.cpp:101-136, in integer arithmetic (the reference's double log/ceil gives the same values for
// every leaf count up to 32^6)
VREN_EXPORT uint32_t vren_calc_bvh_padded_leaf_count(uint32_t leaf_count)
{
    if (leaf_count <= 1) return 32u;
    uint64_t p = 32;
    while (p < leaf_count) p *= 32;
    return (uint32_t)p;
}

VREN_EXPORT uint32_t vren_calc_bvh_buffer_length(uint32_t leaf_count)
{
    uint32_t p = vren_calc_bvh_padded_leaf_count(leaf_count);
    uint32_t len = 0;
    while (p != 0) {
        len += p;
        p >>= 5;
    }
    return len;
}

VREN_EXPORT size_t vren_calc_bvh_buffer_size(uint32_t leaf_count)
{
    return (size_t)vren_calc_bvh_buffer_length(leaf_count) * sizeof(vren_bvh_node);
}

VREN_EXPORT uint32_t vren_calc_bvh_root_index(uint32_t leaf_count)
{
    return vren_calc_bvh_buffer_length(leaf_count) - 1;
}

VREN_EXPORT uint32_t vren_calc_bvh_level_count(uint32_t leaf_count)
{
    return ilog2_u64(vren_calc_bvh_padded_leaf_count(leaf_count)) / 5;
}

// camera.cpp:9-50. The rotation part is composed in double and rounded once; for yaw = pitch = 0 the view
// matrix is the exact translation and the projection is exactly the float expression of camera.cpp:40-50.
VREN_EXPORT void vren_camera_matrices(const vren_camera* cam, float view[16], float proj[16])
{
    double cy = std::cos((double)cam->yaw), sy = std::sin((double)cam->yaw);
    double cp = std::cos(-(double)cam->pitch), sp = std::sin(-(double)cam->pitch);
    // orientation = Ry(yaw) * Rx(-pitch); view = transpose(orientation) * translate(-position)
    double o[9];
    // Ry columns: (cy,0,-sy) (0,1,0) (sy,0,cy); Rx columns: (1,0,0) (0,cp,sp) (0,-sp,cp)
    double ry[9] = {cy, 0, -sy, 0, 1, 0, sy, 0, cy};
    double rx[9] = {1, 0, 0, 0, cp, sp, 0, -sp, cp};
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += ry[k * 3 + r] * rx[c * 3 + k];
            o[c * 3 + r] = s;
        }
    double p[3] = {-(double)cam->position[0], -(double)cam->position[1], -(double)cam->position[2]};
    for (int i = 0; i < 16; i++) view[i] = 0.0f;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) view[c * 4 + r] = (float)o[r * 3 + c];
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += o[r * 3 + k] * p[k];
        view[12 + r] = (float)s;
    }
    view[15] = 1.0f;
    float tan_half = (float)std::tan((double)(cam->fov_y / 2.0f));
    for (int i = 0; i < 16; i++) proj[i] = 0.0f;
    proj[0] = 1.0f / (tan_half * cam->aspect_ratio);
    proj[5] = 1.0f / tan_half;
    proj[10] = cam->far_plane / (cam->far_plane - cam->near_plane);
    proj[11] = 1.0f;
    proj[14] = -(cam->far_plane * cam->near_plane) / (cam->far_plane - cam->near_plane);
}

} // extern "C"
