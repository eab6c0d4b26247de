// frame.hip -- host-side setup of FrameDev records and their upload to device memory.
#include <cmath>
#include <cstring>

#include "chain.hpp"

namespace {

void mat4_inverse_d(const double* m, double* inv)
{
    double a[16];
    a[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    a[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    a[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    a[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    a[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    a[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    a[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    a[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    a[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    a[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    a[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    a[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    a[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    a[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    a[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    a[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * a[0] + m[1] * a[4] + m[2] * a[8] + m[3] * a[12];
    for (int i = 0; i < 16; i++) inv[i] = a[i] / det;
}

__global__ void upload_frame_kernel(FrameDev* dst, FrameDev src) { *dst = src; }

} // namespace

void vren_frame_setup(FrameDev* f, uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg)
{
    f->W = W;
    f->H = H;
    f->Sx = (W + 31) / 32;
    f->Sy = (H + 31) / 32;
    float proj[16];
    vren_camera_matrices(cam, f->view, proj);
    double proj_d[16], view_d[16];
    for (int i = 0; i < 16; i++) {
        proj_d[i] = proj[i];
        view_d[i] = f->view[i];
    }
    mat4_inverse_d(proj_d, f->inv_proj_d);
    mat4_inverse_d(view_d, f->inv_view_d);
    for (int i = 0; i < 16; i++) f->inv_proj[i] = (float)f->inv_proj_d[i];
    f->proj_zw_xy_zero = (f->inv_proj[0 * 4 + 2] == 0.0f && f->inv_proj[1 * 4 + 2] == 0.0f &&
                          f->inv_proj[0 * 4 + 3] == 0.0f && f->inv_proj[1 * 4 + 3] == 0.0f) ? 1u : 0u;
    for (int i = 0; i < 3; i++) f->eye[i] = (double)cam->position[i];
    f->inv_W = 1.0 / (double)W;
    f->inv_H = 1.0 / (double)H;
    {
        // r_x = ip[0] (2 gx / W - 1) + ip[12], r_y = ip[5] (1 - 2 gy / H) + ip[13] (column-major ip)
        const double* ip = f->inv_proj_d;
        f->proj_xy_sparse = (ip[4] == 0.0 && ip[8] == 0.0 && ip[1] == 0.0 && ip[9] == 0.0) ? 1u : 0u;
        f->rec_xy[0] = 2.0 * ip[0] * f->inv_W;
        f->rec_xy[1] = ip[12] - ip[0];
        f->rec_xy[2] = -2.0 * ip[5] * f->inv_H;
        f->rec_xy[3] = ip[5] + ip[13];
    }
    f->near_plane = cam->near_plane;
    f->far_plane = cam->far_plane;
    float tan_half = (float)std::tan((double)(cam->fov_y / 2.0f));
    f->slice_a = (2.0f * tan_half) / (float)f->Sy + 1.0f;
    f->log_a = vren_logf(1.0f + (2.0f * tan_half) / (float)f->Sy);
    f->log_fn = vren_logf(cam->far_plane / cam->near_plane);
    f->bg_threshold = cam->far_plane - 0.001f;
    f->key_mode = cfg->key_mode;
    f->depth_slices = cfg->depth_slices;
    f->aabb_mode = cfg->cluster_aabb_mode;
    f->morton_bits = cfg->morton_bits;
    f->morton_aabb_mode = cfg->morton_aabb_mode;
    f->attenuation = cfg->attenuation;
    for (int i = 0; i < 3; i++) {
        f->morton_min[i] = cfg->morton_aabb_min[i];
        f->morton_max[i] = cfg->morton_aabb_max[i];
    }
}

int vren_validate_cfg(vren_ctx* ctx, const vren_cluster_cfg* cfg)
{
    if (!cfg) return vren_fail(ctx, VREN_EINVAL, "null cluster cfg");
    if (cfg->key_mode > VREN_KEYS_SLICED) return vren_fail(ctx, VREN_EINVAL, "bad key_mode %u", cfg->key_mode);
    if (cfg->key_mode == VREN_KEYS_SLICED && (cfg->depth_slices == 0 || cfg->depth_slices > 256))
        return vren_fail(ctx, VREN_EINVAL, "depth_slices must be in [1, 256]");
    if (cfg->cluster_aabb_mode > VREN_CLUSTER_AABB_TIGHT) return vren_fail(ctx, VREN_EINVAL, "bad cluster_aabb_mode");
    if (cfg->morton_bits < 1 || cfg->morton_bits > 10) return vren_fail(ctx, VREN_EINVAL, "morton_bits must be in [1, 10]");
    if (cfg->morton_aabb_mode > VREN_MORTON_GIVEN_AABB) return vren_fail(ctx, VREN_EINVAL, "bad morton_aabb_mode");
    if (cfg->attenuation > VREN_ATTEN_SMOOTH) return vren_fail(ctx, VREN_EINVAL, "bad attenuation");
    return VREN_OK;
}

int vren_validate_camera(vren_ctx* ctx, const vren_camera* cam)
{
    if (!cam) return vren_fail(ctx, VREN_EINVAL, "null camera");
    if (!(cam->near_plane > 0.0f) || !(cam->far_plane > cam->near_plane) || !(cam->fov_y > 0.0f) ||
        !(cam->fov_y < 3.14159f) || !(cam->aspect_ratio > 0.0f))
        return vren_fail(ctx, VREN_EINVAL, "camera needs 0 < near < far, 0 < fov_y < pi and aspect > 0");
    return VREN_OK;
}

int vren_validate_material_set(vren_ctx* ctx, const vren_material_set* m, const char* what)
{
    if (!m->texcoords || !m->material_indices)
        return vren_fail(ctx, VREN_EINVAL, "%s: the material path needs texcoords and material indices", what);
    if ((m->material_count > 0 && !m->materials) || (m->texture_count > 0 && !m->textures))
        return vren_fail(ctx, VREN_EINVAL, "%s: null material or texture table with a non-zero count", what);
    if ((reinterpret_cast<uintptr_t>(m->texcoords) & 7) || (reinterpret_cast<uintptr_t>(m->material_indices) & 1) ||
        (reinterpret_cast<uintptr_t>(m->materials) & 15) || (reinterpret_cast<uintptr_t>(m->textures) & 15))
        return vren_fail(ctx, VREN_EINVAL, "%s: texcoords must be 8-byte, material indices 2-byte, material and "
                                           "texture tables 16-byte aligned", what);
    return VREN_OK;
}

void vren_frame_set_materials(FrameDev* f, const vren_material_set* m)
{
    f->texcoords = m->texcoords;
    f->mat_idx = m->material_indices;
    f->materials = m->materials;
    f->n_materials = m->materials ? m->material_count : 0;
    f->textures = m->textures;
    f->n_textures = m->textures ? m->texture_count : 0;
}

int vren_upload_frames(vren_ctx* ctx, hipStream_t s, FrameDev* dst, const FrameDev* src, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) {
        upload_frame_kernel<<<1, 1, 0, s>>>(dst + i, src[i]);
        VREN_LAUNCH_CHECK(ctx);
    }
    return VREN_OK;
}

FrameDev* vren_next_frame_slot(vren_ctx* ctx)
{
    FrameDev* slot = ctx->frame_slots + ctx->frame_slot_next;
    ctx->frame_slot_next = (ctx->frame_slot_next + 1) % VREN_FRAME_SLOTS;
    return slot;
}
