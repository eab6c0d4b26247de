// clustered.hpp -- per-frame device descriptor of the clustered-shading chain.
//
// Every clustered-shading kernel reads its frame's sizes, camera constants and buffer pointers from a table of
// FrameDev records in device memory and takes the number of frames in the table, so a single frame (the
// reference's per-frame call) and a batch of frames run the same code. Constants that the reference recomputes
// per invocation in GLSL (inverse(P), tan(fov/2), the depth-slice base) are computed once on the host.
#pragma once

#include "common.hpp"

// light lists of at most this many entries are kept by the counting traversal and copied by the writing pass; a
// longer list is traversed again by the writing pass. 192 entries keep every list of a 65 536-light 4K frame (the
// longest lists of the heterogeneous 64-frame batch hold 147 lights; 0.27 % of its clusters hold more than 64, and
// traversing those again cost 0.2 ms per batch) while the counting kernel's per-half LDS lists still let 8
// workgroups share a CU
#define VREN_LIST_SLOT 192u

struct FrameDev {
    // sizes
    uint32_t W, H, Sx, Sy;
    uint32_t tile_base;     // first flattened 32x32 tile of this frame in the table
    uint32_t leaf_base;     // first flattened leaf slot (256-aligned) of this frame's BVH leaves
    uint32_t live_leaf_base; // the same over the leaf groups that hold lights only (a batch's chain launch)
    uint32_t offs_base;     // first flattened tile of this frame's light-offset scan
    // camera constants
    float view[16];         // camera::get_view()
    float inv_proj[16];     // inverse(camera::get_projection()), rounded from double
    uint32_t proj_zw_xy_zero; // inv_proj's z and w rows have zero x and y coefficients (every perspective camera)
    double inv_proj_d[16];  // shade reconstruction (double)
    double inv_view_d[16];
    double eye[3];
    double inv_W, inv_H;    // 1 / W, 1 / H (shade: pixel x/y only feed the zero coefficients of the z/w rows)
    // inv_proj's x row depends on NDC x alone and its y row on NDC y alone (every perspective camera): shade then
    // takes r_x = rec_xy[0] gx + rec_xy[1] and r_y = rec_xy[2] gy + rec_xy[3], the NDC transform folded in
    uint32_t proj_xy_sparse;
    double rec_xy[4];
    float near_plane, far_plane;
    float slice_a;          // 2 tan(fov/2) / Sy + 1
    float log_a;            // vren_logf(slice_a)
    float log_fn;           // vren_logf(far / near)
    float bg_threshold;     // far - 0.001
    // configuration
    uint32_t key_mode, depth_slices, aabb_mode, morton_bits, morton_aabb_mode, attenuation;
    float morton_min[3], morton_max[3];
    // lights
    const float* light_pos;   // float4[L] world
    const float* lights;      // {rgb, radius}[L]
    uint32_t L;
    uint32_t n_dir;
    const float* dir_lights;  // {dir, pad, rgb, pad}[n_dir]
    float* view_pos;          // float4[L]
    uint32_t* codes;          // uint2[L] (code, light index), sorted by code
    uint32_t* codes_tmp;      // uint2[L] scratch for the sort
    vren_bvh_node* bvh;
    uint32_t bvh_padded, bvh_root, bvh_levels, bvh_length;
    float* leaf_pos;          // optional float4[L]: view position of sorted leaf i (= view_pos[codes[i].y])
    // G-buffer
    const float* depth;
    const uint16_t* normals;
    const float* albedo_rough;
    const float* metallic;
    // material path (shade.comp:89-96): texcoords != nullptr selects it over albedo_rough / metallic
    const float* texcoords;          // float2[W*H]
    const uint16_t* mat_idx;         // u16[W*H]
    const vren_material* materials;
    const vren_texture* textures;
    uint32_t n_materials, n_textures;
    // clusters
    uint32_t* keys;
    uint32_t* cluster_count;  // device u32
    uint32_t* refs;
    uint32_t key_cap;
    uint32_t index_cap;
    // light lists
    uint32_t* counts;
    uint32_t* offsets;
    uint32_t* indices;
    uint32_t* assigned_total; // device u32
    uint32_t* list_slots;     // optional: VREN_LIST_SLOT entries per cluster for clusters below slot_cap
    uint32_t slot_cap;
    uint64_t* offset_status;  // look-back statuses of the offset scan (ceil(key_cap / 4096) words)
    // output
    float* out;
    // error flags
    uint32_t* flags;
};

// host: fill the camera/config part of a FrameDev (sizes, matrices, derived constants)
void vren_frame_setup(FrameDev* f, uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg);
