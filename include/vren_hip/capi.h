/*
 * vren_hip/capi.h -- the C ABI of the MI355X (gfx950) clustered-shading path.
 *
 * Every entry point below replaces one operator of the reference (loryruta/vren). The reference records
 * Vulkan dispatches into a caller-supplied VkCommandBuffer and returns void; here each call enqueues HIP
 * kernels on the caller's stream and returns a status code. Conventions:
 *   - all buffers are device pointers owned by the caller (the reference's vk_utils::buffer + byte offset);
 *   - calls are asynchronous on `stream` (a hipStream_t; NULL = the legacy default stream);
 *   - return 0 (VREN_OK) or a negative vren_status; arguments are validated on the host before any HIP call,
 *     so invalid lengths / capacities are reported without touching the device;
 *   - a vren_ctx is not thread-safe (one per host thread), mirroring single-threaded command recording, and is
 *     used from one stream at a time: the primitives that take no caller scratch share the context's scratch
 *     buffer. Growing it synchronises the device, so a call that would grow it during a stream capture fails
 *     with VREN_EINVAL (make the same call once outside the capture first);
 *   - capacity overflows that only the device can see (unique clusters, assigned lights) never write out of
 *     bounds: they raise a device-side flag that vren_sync_status() reports as VREN_ECAPACITY
 *     (the reference writes past its buffers silently, config.hpp:23-24).
 */
#ifndef VREN_HIP_CAPI_H_
#define VREN_HIP_CAPI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VREN_EXPORT __attribute__((visibility("default")))

typedef void* vren_stream_t; /* hipStream_t */

typedef enum vren_status {
    VREN_OK = 0,
    VREN_EINVAL = -1,    /* precondition violated (reference: assert / std::invalid_argument) */
    VREN_ECAPACITY = -2, /* an output exceeded its capacity (reference: silent out-of-bounds write) */
    VREN_EHIP = -3,      /* a HIP runtime call failed (reference: VREN_CHECK -> std::runtime_error) */
    VREN_ENOMEM = -4,    /* device allocation failed */
} vren_status;

enum { VREN_REDUCE_ADD = 0, VREN_REDUCE_MIN = 1, VREN_REDUCE_MAX = 2 };

/* bvh_node (build_bvh.hpp:8-18): 32 bytes */
#define VREN_BVH_LEAF_NODE 0xFFFFFFFFu
#define VREN_BVH_INVALID_NODE 0xFFFFFFFEu
typedef struct vren_bvh_node {
    float min[3];
    uint32_t next; /* first child index, VREN_BVH_LEAF_NODE or VREN_BVH_INVALID_NODE */
    float max[3];
    uint32_t pad;
} vren_bvh_node;

/* camera (camera.hpp fields used by camera.cpp:9-50) */
typedef struct vren_camera {
    float position[3];
    float fov_y;
    float aspect_ratio;
    float near_plane;
    float far_plane;
    float yaw;
    float pitch;
} vren_camera;

/* cluster-key encoders */
enum {
    VREN_KEYS_REFERENCE = 0, /* 8b tile x | 8b tile y | 10b log slice | 6b cube-face normal code (find_unique_clusters.comp:72-76) */
    VREN_KEYS_SLICED = 1,    /* 8b tile x | 8b tile y | 8b log slice over [near, far] in `depth_slices` | 3b face (0..5, 7 = null) */
};
/* cluster AABB used by light assignment */
enum {
    VREN_CLUSTER_AABB_REFERENCE = 0, /* two corners as written in clustered_shading.glsl:193-231 (y range inverted) */
    VREN_CLUSTER_AABB_TIGHT = 1,     /* min/max over the 8 corners of the frustum slice */
};
/* Morton code domain */
enum {
    VREN_MORTON_LIGHTS_AABB = 0, /* AABB of the view-space light positions (clustered_shading.cpp:137-224) */
    VREN_MORTON_GIVEN_AABB = 1,  /* caller-given view-space AABB (e.g. the scene bounds) */
};
/* point-light attenuation in shade */
enum {
    VREN_ATTEN_REFERENCE = 0, /* 1 - min(d / r, 1)^1.5 (pbr.glsl:89) */
    VREN_ATTEN_SMOOTH = 1,    /* (1 - (d / r)^4)_+^2, differentiable everywhere */
};

typedef struct vren_cluster_cfg {
    uint32_t key_mode;          /* VREN_KEYS_* */
    uint32_t depth_slices;      /* VREN_KEYS_SLICED only (1..256) */
    uint32_t cluster_aabb_mode; /* VREN_CLUSTER_AABB_* */
    uint32_t morton_bits;       /* bits per axis: 5 (reference, 15-bit codes) .. 10 (30-bit codes) */
    uint32_t morton_aabb_mode;  /* VREN_MORTON_* */
    float morton_aabb_min[3];   /* view space, VREN_MORTON_GIVEN_AABB only */
    float morton_aabb_max[3];
    uint32_t attenuation;       /* VREN_ATTEN_* */
} vren_cluster_cfg;

typedef struct vren_ctx vren_ctx;

/* hip_device = -1 creates a host-only context: argument validation and the host helpers work and every call
 * that would touch a device returns VREN_EINVAL after validating its arguments (for tests without a GPU) */
VREN_EXPORT int vren_create(int hip_device, vren_ctx** out);
VREN_EXPORT void vren_destroy(vren_ctx* ctx);
VREN_EXPORT const char* vren_status_string(int status);
/* last host-side error message of this context (argument validation / HIP error text) */
VREN_EXPORT const char* vren_last_error(vren_ctx* ctx);
/* synchronise `stream` and report (then clear) device-side capacity flags raised by calls on this ctx */
VREN_EXPORT int vren_sync_status(vren_ctx* ctx, vren_stream_t stream);
/* the reference's default configuration (reference keys / AABB / 15-bit Morton / reference attenuation) */
VREN_EXPORT void vren_cluster_cfg_default(vren_cluster_cfg* cfg);

/* ---------------------------------------------------------------------------------------------------------
 * Parallel primitives (vren/vren/primitives)
 * ------------------------------------------------------------------------------------------------------- */

/* reduce<uint, op>::operator() (reduce.hpp:15-49, reduce.cpp:32-114): writes the in-place Blelloch up-sweep
 * tree of each of `blocks` segments. Segment b reads in[b * n ...], padded with the op identity to
 * p = next_pow2(n), and writes out_tree[b * p ... b * p + p); the segment total is at out_tree[b * p + p - 1].
 * in == out_tree (in-place overload, reduce.cpp:116-128) is allowed when blocks == 1 or n is a power of 2. */
VREN_EXPORT int vren_reduce_u32(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in, uint32_t n,
                                uint32_t* out_tree, int op, uint32_t blocks);
/* reduce<vec4, op>: same over float4 elements (16 bytes); identities 0, +1e35, -1e35 (VRen.cmake:70-75) */
VREN_EXPORT int vren_reduce_vec4(vren_ctx* ctx, vren_stream_t stream, const float* in, uint32_t n,
                                 float* out_tree, int op, uint32_t blocks);
/* scalar reduction of n uint32 (no tree): *out = op over in[0..n) (mod 2^32 for add) */
VREN_EXPORT int vren_reduce_total_u32(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in, uint64_t n,
                                      uint32_t* out, int op);

/* blelloch_scan::operator() (blelloch_scan.hpp:9-49, blelloch_scan.cpp:141-166): exclusive prefix sum
 * mod 2^32 of each of `blocks` contiguous segments of n elements. in == out (in place) is allowed.
 * Any n >= 1 (the reference requires a power of 2). */
VREN_EXPORT int vren_exclusive_scan_u32(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in, uint32_t* out,
                                        uint64_t n, uint32_t blocks);
/* blelloch_scan::downsweep (blelloch_scan.hpp:31-39, blelloch_scan.cpp:57-139, blelloch_scan_downsweep.comp):
 * in place on `blocks` segments of `length` (a power of 2) elements. Every element i becomes
 * P + sum over the set bits l of i of d[((i >> (l+1)) << (l+1)) + 2^l - 1] (the reference's level-by-level swaps
 * and adds, unrolled), where P = 0 if clear_last or length < 1024 (the reference zero-pads shorter segments in its
 * 1024-wide LDS pass), else the segment's last element. On an up-sweep tree (vren_reduce_u32 add) with clear_last
 * this is the exclusive prefix sum of the leaves. */
VREN_EXPORT int vren_blelloch_downsweep_u32(vren_ctx* ctx, vren_stream_t stream, uint32_t* data, uint32_t length,
                                            uint32_t blocks, int clear_last);
/* No counterpart in the reference: vren_exclusive_scan_u32 for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid hipMalloc). The same kernels and results, with
 * the look-back workspace in a fixed device array (one copy per GPU, 512 KiB + 256 B, resolved by vren_create), so
 * the call makes no hipMalloc, hipFree or other allocation. At most 2^16 tiles of 16384 keys over all segments:
 * ceil(n / 16384) * blocks <= 65536 (n * blocks <= 2^30), else VREN_EINVAL. Calls that overlap on one GPU (other
 * streams or contexts) share the array and race; calls in sequence on one stream do not. */
VREN_EXPORT int vren_exclusive_scan_u32_noalloc(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in,
                                                uint32_t* out, uint64_t n, uint32_t blocks);

/* radix_sort (radix_sort.hpp:9-53, radix_sort.cpp:149-338): stable LSD sort of uint32 keys (and, if vals is
 * not NULL, uint32 values carried with them) over key bits [begin_bit, end_bit). The result is left in
 * keys/vals, as the reference leaves it in `buffer`. scratch: vren_radix_sort_scratch_bytes(n, vals != NULL). */
VREN_EXPORT size_t vren_radix_sort_scratch_bytes(uint64_t n, int with_values);
VREN_EXPORT int vren_radix_sort(vren_ctx* ctx, vren_stream_t stream, uint32_t* keys, uint32_t* vals, uint64_t n,
                                uint32_t begin_bit, uint32_t end_bit, void* scratch, size_t scratch_bytes);

/* Radix-sort autotuner. No reference counterpart (the reference's digit width is fixed at 4 bits,
 * radix_sort.hpp:15-16). Configurations are (digit bits, keys per thread, waves per workgroup, algorithm: 0 onesweep
 * with a per-digit decoupled look-back, 1 reduce-then-scan: tile counts, one scan, scatter); every configuration
 * gives the same (stable) output. vren_autotune_radix_sort times them all on pseudo-random keys of n elements and
 * remembers the fastest for that size class in the context. Out-pointers may be NULL. */
VREN_EXPORT int vren_radix_sort_config_count(void);
VREN_EXPORT int vren_radix_sort_config_info(int config, uint32_t* radix_bits, uint32_t* keys_per_thread,
                                            uint32_t* waves_per_block, uint32_t* algorithm);
VREN_EXPORT int vren_radix_sort_force_config(vren_ctx* ctx, int config); /* -1 = automatic */
VREN_EXPORT int vren_autotune_radix_sort(vren_ctx* ctx, vren_stream_t stream, uint64_t n, int with_values,
                                         uint32_t key_bits, int* best_config, float* best_ms);
VREN_EXPORT int vren_autotune_save(vren_ctx* ctx, const char* path);
VREN_EXPORT int vren_autotune_load(vren_ctx* ctx, const char* path);

/* bucket_sort::operator() (bucket_sort.hpp:9-45, bucket_sort.cpp:72-161): counting sort of (key, value)
 * uint32 pairs into n_buckets (a power of 2, <= 65536) buckets by (key >> key_shift) & (n_buckets - 1).
 * Stable (the reference's atomic scatter is not). bucket_offsets[n_buckets] receives the exclusive bucket
 * starts (the reference appends them to its output buffer, bucket_sort.cpp:67-70). Pair buffers are 8-byte
 * aligned and distinct (VREN_EINVAL otherwise). */
VREN_EXPORT int vren_bucket_sort(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in_pairs, uint32_t n,
                                 uint32_t* out_pairs, uint32_t* bucket_offsets, uint32_t n_buckets,
                                 uint32_t key_shift);

/* vren_demo::point_light_bouncer::bounce (point_light_bouncer.cpp:17-57) and bounce_point_lights.comp:33-73: moves
 * n point lights in place by speed * dt inside [aabb_min, aabb_max] (each position first clamped into it),
 * reflecting the direction components of the walls reached, at most 32 segments. positions and directions are
 * vec4[n] (16-byte aligned); w is written as 1 and 0. The per-frame light update ahead of the light BVH build. */
VREN_EXPORT int vren_bounce_point_lights(vren_ctx* ctx, vren_stream_t stream, float* positions, float* directions,
                                         uint32_t n, const float* aabb_min, const float* aabb_max, float speed,
                                         float dt);

/* calc_bvh_* (build_bvh.hpp:52-56, build_bvh.cpp:101-136), host only */
VREN_EXPORT uint32_t vren_calc_bvh_padded_leaf_count(uint32_t leaf_count);
VREN_EXPORT uint32_t vren_calc_bvh_buffer_length(uint32_t leaf_count);
VREN_EXPORT size_t vren_calc_bvh_buffer_size(uint32_t leaf_count);
VREN_EXPORT uint32_t vren_calc_bvh_root_index(uint32_t leaf_count);
VREN_EXPORT uint32_t vren_calc_bvh_level_count(uint32_t leaf_count);

/* build_bvh::operator() (build_bvh.cpp:38-99): builds the parent levels of a 32-ary BVH whose
 * padded_leaf_count (a power of 32, >= 32) leaves are already in nodes[0 .. padded_leaf_count). A parent's
 * AABB is the min/max over its valid children, next = its first child, or VREN_BVH_INVALID_NODE with the
 * empty AABB (min = +inf, max = -inf) when no child is valid. */
VREN_EXPORT int vren_build_bvh(vren_ctx* ctx, vren_stream_t stream, vren_bvh_node* nodes,
                               uint32_t padded_leaf_count);

/* Karras binary radix-tree BVH (LBVH; no reference counterpart: a binary BVH of the lights, beside the reference's
 * 32-ary tree). Leaves are the n >= 1 sorted (Morton code, light index) pairs light_codes[n] (uint2; the 64-bit keys
 * code << 32 | index must be strictly increasing, as vren_construct_point_light_bvh emits them); leaf i's AABB is
 * view_positions[light index] -+ radius (point_lights[.].w), as in the 32-ary leaves.
 * nodes[2n - 1]: internal nodes 0 .. n-2 (root 0; for n == 1 the root is the only leaf), then leaf i at n - 1 + i.
 * An internal node holds its children's node indices; a leaf has left = VREN_BVH_LEAF_NODE and right = its light
 * index. Every internal AABB is the exact min/max of its children (refitted bottom-up). parents[2n - 1] receives
 * each node's parent (0xFFFFFFFF for the root). Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees,
 * and k-d Trees", HPG 2012. */
typedef struct vren_lbvh_node {
    float min[3];
    uint32_t left;
    float max[3];
    uint32_t right;
} vren_lbvh_node;
VREN_EXPORT int vren_build_lbvh(vren_ctx* ctx, vren_stream_t stream, const uint32_t* light_codes, uint32_t n,
                                const float* view_positions, const float* point_lights, vren_lbvh_node* nodes,
                                uint32_t* parents);

/* ---------------------------------------------------------------------------------------------------------
 * Clustered shading (vren/vren/pipeline/clustered_shading.{hpp,cpp})
 * ------------------------------------------------------------------------------------------------------- */

/* camera::get_view / get_projection (camera.cpp:32-50), column-major float[16] */
VREN_EXPORT void vren_camera_matrices(const vren_camera* camera, float view[16], float proj[16]);

/* construct_point_light_bvh::operator() (clustered_shading.cpp:60-346):
 *   view_positions[L] (float4) = view * (p, 1);  light_codes[L] (uint2) = (Morton code, light index) stably
 *   sorted by code;  bvh[vren_calc_bvh_buffer_length(L)] = 32-ary BVH over the sorted lights (leaf AABB =
 *   view position -+ radius; padding leaves invalid). point_lights = {float color[3]; float intensity}[L]
 *   where intensity is the influence radius (init_light_array_bvh.comp:52-53). */
VREN_EXPORT int vren_construct_point_light_bvh(vren_ctx* ctx, vren_stream_t stream, const float* light_positions,
                                               const float* point_lights, uint32_t light_count,
                                               const vren_camera* camera, const vren_cluster_cfg* cfg,
                                               float* view_positions, uint32_t* light_codes, vren_bvh_node* bvh);

/* find_unique_cluster_list::operator() (clustered_shading.cpp:363-448): per-pixel cluster keys from the D32
 * depth and RGBA16F normals, unique keys per 32x32 tile in ascending order, tiles in row-major order.
 * cluster_keys[key_capacity] receives the unique keys, *cluster_count (device) their number, and
 * cluster_refs[W*H] each pixel's index into cluster_keys. */
VREN_EXPORT int vren_find_unique_cluster_list(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                              const vren_camera* camera, const vren_cluster_cfg* cfg,
                                              const float* depth, const uint16_t* normals, uint32_t* cluster_keys,
                                              uint32_t key_capacity, uint32_t* cluster_count,
                                              uint32_t* cluster_refs);

/* assign_lights::operator() (clustered_shading.cpp:473-690): per unique cluster, traverse the light BVH and
 * write the indices of the lights whose sphere meets the cluster AABB, each list in ascending light index.
 * light_counts / light_offsets [key_capacity], light_indices[index_capacity], *assigned_total (device). */
VREN_EXPORT int vren_assign_lights(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                   const vren_camera* camera, const vren_cluster_cfg* cfg,
                                   const uint32_t* cluster_keys, const uint32_t* cluster_count, uint32_t key_capacity,
                                   const vren_bvh_node* bvh, uint32_t light_count, const uint32_t* light_codes,
                                   const float* view_positions, uint32_t* light_counts, uint32_t* light_offsets,
                                   uint32_t* light_indices, uint32_t index_capacity, uint32_t* assigned_total);

/* shade::operator() (clustered_shading.cpp:709-811, shade.comp:69-141) with resolved per-pixel materials:
 * albedo_roughness[W*H] = (r, g, b, roughness), metallic[W*H]. Writes RGBA32F output for every pixel whose
 * view depth is below far - 0.001 (background pixels are left untouched, as in the reference). */
VREN_EXPORT int vren_shade(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                           const vren_camera* camera, const vren_cluster_cfg* cfg, const float* depth,
                           const uint16_t* normals, const float* albedo_roughness, const float* metallic,
                           const uint32_t* cluster_refs, uint32_t key_capacity, const uint32_t* light_counts,
                           const uint32_t* light_offsets, const uint32_t* light_indices, uint32_t index_capacity,
                           const float* light_positions, const float* point_lights, uint32_t light_count,
                           const float* directional_lights, uint32_t directional_light_count, float* output);

/* ---- material resolve (shade.comp:89-96) ---------------------------------------------------------------------
 * The reference's G-buffer carries a texcoord (RG32F) and a material index (R16_UINT) per pixel (gbuffer.cpp:14-27);
 * shade looks the material up in the material buffer (material.hpp:19-58) and samples its base-colour and
 * metallic-roughness textures from the bindless texture array (texture_manager.hpp:32):
 *   albedo = base_colour_texture(uv).rgb * base_color_factor.rgb
 *   metallic = metallic_roughness_texture(uv).b * metallic_factor
 *   roughness = metallic_roughness_texture(uv).g * roughness_factor
 * Textures are RGBA8 UNORM with one mip level (tinygltf_parser.cpp:157); texel selection and filtering are those
 * of include/vren_hip/sampler.h. A material index at or past material_count, or a texture index at or past
 * texture_count, reads zeros (robust buffer access), as does a texture with a zero size or no texels. */
enum {
    VREN_FILTER_NEAREST = 0,
    VREN_FILTER_LINEAR = 1,
};
enum {
    VREN_ADDRESS_REPEAT = 0,
    VREN_ADDRESS_MIRRORED_REPEAT = 1,
    VREN_ADDRESS_CLAMP_TO_EDGE = 2,
};

typedef struct vren_material {   /* vren::material / GLSL Material (common.glsl:61-68): 32 bytes */
    uint32_t base_color_texture_idx;
    uint32_t metallic_roughness_texture_idx;
    float metallic_factor;
    float roughness_factor;
    float base_color_factor[4];
} vren_material;

typedef struct vren_texture {    /* one vk_utils::texture and its sampler: 32 bytes */
    const uint8_t* texels;       /* device, 4-byte aligned: height rows of width RGBA8 texels */
    uint32_t width, height;
    uint32_t filter;             /* VREN_FILTER_* (the magnification filter: compute shaders sample at lambda = 0) */
    uint32_t address_u, address_v; /* VREN_ADDRESS_* */
    uint32_t pad;
} vren_texture;

typedef struct vren_material_set {
    const float* texcoords;           /* f32x2[W*H] (RG32F G-buffer texcoords) */
    const uint16_t* material_indices; /* u16[W*H] (R16_UINT G-buffer material indices) */
    const vren_material* materials;   /* device, material_count records (vren::material_buffer::m_buffer) */
    uint32_t material_count;
    uint32_t texture_count;
    const vren_texture* textures;     /* device, texture_count descriptors (texture_manager::m_textures) */
} vren_material_set;

/* shade (clustered_shading.hpp:114-140) with the reference's material path: as vren_shade, with the per-pixel
 * albedo, roughness and metallic resolved from `materials` (shade.comp:89-96) instead of read from buffers. */
VREN_EXPORT int vren_shade_materials(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                     const vren_camera* camera, const vren_cluster_cfg* cfg, const float* depth,
                                     const uint16_t* normals, const vren_material_set* materials,
                                     const uint32_t* cluster_refs, uint32_t key_capacity,
                                     const uint32_t* light_counts, const uint32_t* light_offsets,
                                     const uint32_t* light_indices, uint32_t index_capacity,
                                     const float* light_positions, const float* point_lights, uint32_t light_count,
                                     const float* directional_lights, uint32_t directional_light_count,
                                     float* output);

/* shade_backward: adjoint of vren_shade with the light lists held fixed (no reference counterpart; the reference
 * renders forward only). Given dL/d(output) (RGBA32F, alpha ignored) it ADDS into
 *   grad_light_positions[L] (float4: d/dx, d/dy, d/dz, 0), grad_point_lights[L] (float4: d/d colour*intensity rgb,
 *   d/d radius) and grad_albedo_roughness[W*H] (float4: d/d albedo rgb, d/d roughness).
 * The gradient buffers must be zeroed by the caller before the first accumulation. Use VREN_ATTEN_SMOOTH: the
 * reference attenuation has a kink at d = r. */
VREN_EXPORT int vren_shade_backward(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                    const vren_camera* camera, const vren_cluster_cfg* cfg, const float* depth,
                                    const uint16_t* normals, const float* albedo_roughness, const float* metallic,
                                    const uint32_t* cluster_refs, uint32_t key_capacity, const uint32_t* light_counts,
                                    const uint32_t* light_offsets, const uint32_t* light_indices,
                                    uint32_t index_capacity, const float* light_positions, const float* point_lights,
                                    uint32_t light_count, const float* grad_output, float* grad_light_positions,
                                    float* grad_point_lights, float* grad_albedo_roughness);

/* ---------------------------------------------------------------------------------------------------------
 * cluster_and_shade (clustered_shading.hpp:147-182, clustered_shading.cpp:817-1172) over a batch of frames.
 * A batch owns every intermediate buffer of its frames (sized per frame, not at the reference's fixed maxima)
 * and runs the whole chain -- light BVH, unique clusters, light assignment, shade -- either frame by frame
 * (each frame's launches followed by a stream synchronisation, the reference's submit + fence wait) or for all
 * frames at once (one launch per stage), optionally replayed from a captured HIP graph. All modes produce the
 * same bytes per frame.
 * ------------------------------------------------------------------------------------------------------- */
typedef struct vren_frame_desc {
    uint32_t width, height;
    vren_camera camera;
    const float* depth;              /* f32[W*H] NDC depth (D32) */
    const uint16_t* normals;         /* f16x4[W*H] (RGBA16F) */
    const float* albedo_roughness;   /* f32x4[W*H] */
    const float* metallic;           /* f32[W*H] */
    const float* light_positions;    /* f32x4[L] world space */
    const float* point_lights;       /* {f32 colour[3]; f32 radius}[L] */
    uint32_t light_count;
    const float* directional_lights; /* {f32 dir[3]; pad; f32 colour[3]; pad}[D] */
    uint32_t directional_light_count;
    float* output;                   /* f32x4[W*H] RGBA32F */
    uint32_t key_capacity;           /* 0: min(tiles * 1024, 2^20) */
    uint32_t index_capacity;         /* 0: min(key_capacity * L, 2^24) */
    const vren_material_set* materials; /* NULL: albedo_roughness + metallic; else the material path (then those
                                           two may be NULL). Read when the batch is created. */
} vren_frame_desc;

typedef struct vren_frame_outputs {
    float* view_positions;
    uint32_t* light_codes;
    vren_bvh_node* bvh;
    uint32_t bvh_length;
    uint32_t* cluster_keys;
    uint32_t* cluster_count;
    uint32_t* cluster_refs;
    uint32_t key_capacity;
    uint32_t* light_counts;
    uint32_t* light_offsets;
    uint32_t* light_indices;
    uint32_t index_capacity;
    uint32_t* assigned_total;
} vren_frame_outputs;

typedef struct vren_batch vren_batch;
enum {
    VREN_RUN_PER_FRAME = 0, /* per frame: launch its chain, hipStreamSynchronize */
    VREN_RUN_BATCHED = 1,   /* every frame in one launch per stage, asynchronous */
    VREN_RUN_GRAPH = 2,     /* VREN_RUN_BATCHED captured once into a hipGraph, then replayed */
};
VREN_EXPORT int vren_batch_create(vren_ctx* ctx, const vren_frame_desc* frames, uint32_t n_frames,
                                  const vren_cluster_cfg* cfg, vren_batch** out);
VREN_EXPORT int vren_batch_run(vren_batch* batch, vren_stream_t stream, int mode);
/* VREN_RUN_BATCHED with a HIP event after every kernel; kernel_ms[VREN_CHAIN_KERNEL_COUNT] receives the
 * durations of chain init, lights, BVH leaves, BVH upper levels, clusters, light counts, light offsets,
 * light lists and shade (the reference's GPU timestamp profiler, profiler.cpp:32-104, has no slot wired on
 * this path). Synchronises the stream. */
#define VREN_CHAIN_KERNEL_COUNT 9
VREN_EXPORT int vren_batch_run_profiled(vren_batch* batch, vren_stream_t stream, float* kernel_ms);
VREN_EXPORT int vren_batch_frame_outputs(vren_batch* batch, uint32_t frame, vren_frame_outputs* out);
VREN_EXPORT void vren_batch_destroy(vren_batch* batch);

/* The chain's launch shapes, picked by the chain autotuner. No reference counterpart: the reference fixes its
 * workgroup sizes in the shaders (32x32 tiles, 1024-thread sorts, one wave32 per cluster); here a batch runs its
 * chain with one shape, chosen per size class. Every shape gives the same bytes. Zero fields mean the defaults. */
typedef struct vren_chain_shape {
    uint32_t shade_pixels_per_thread; /* 1 or 2: 32x8 strips per 256-thread shade workgroup (default 1) */
    uint32_t assign_waves_per_cu;     /* 8, 16, 24 or 32: persistent light-assignment waves per CU (default 32) */
    uint32_t light_sort_bits;         /* 4 or 8: digit width of each frame's Morton-code sort (default 8) */
    uint32_t light_stage_concurrent;  /* 0 or 1: batched runs build the light BVH on a second stream beside the
                                         cluster stage (a fork and join in the captured graph; default 0) */
} vren_chain_shape;
/* the batch's current shape (the context's tuned shape for the batch's size class when one exists, else the
 * defaults) and a forced one (invalid fields: VREN_EINVAL); a new shape drops the captured graph */
VREN_EXPORT int vren_batch_get_shape(vren_batch* batch, vren_chain_shape* shape);
VREN_EXPORT int vren_batch_set_shape(vren_batch* batch, const vren_chain_shape* shape);
/* times the batched chain (VREN_RUN_BATCHED, `reps` timed runs after one warm-up) with every shape of the grid
 * {1, 2} x {8, 16, 24, 32} x {4, 8} x {0, 1}, keeps the fastest for the batch and caches it in the context under the batch's
 * size class (log2 of its pixels, log2 of its largest light count, log2 of its frame count), where batches created
 * later in that class find it; vren_autotune_save / _load keep the chain shapes beside the radix configurations.
 * Synchronises the stream; the outputs hold the last run's results. */
VREN_EXPORT int vren_batch_autotune(vren_batch* batch, vren_stream_t stream, uint32_t reps, vren_chain_shape* best,
                                    float* best_ms);

/* ---------------------------------------------------------------------------------------------------------
 * Depth-buffer pyramid (depth_buffer_pyramid.hpp:19-80, depth_buffer_pyramid.cpp:14-23, 177-299)
 * The Hi-Z max pyramid of a D32 depth buffer: levels 0 .. n-1, n = floor(log2(max(W, H))) + 1, level l of
 * max(W >> l, 1) x max(H >> l, 1) texels, packed one after another (row-major) per image. Level 0 is the depth
 * buffer (depth_buffer_copy.comp); level l + 1 texel (x, y) is max(0, the level-l texels (2x + i, 2y + j) inside
 * level l) with GLSL max (depth_buffer_reduce.comp: NaN and negative depths never win; an odd last row or column
 * of a level is not read by the next). `images` depth buffers of W * H floats each, stacked, give `images`
 * pyramids of vren_depth_pyramid_length(W, H) floats each.
 * ------------------------------------------------------------------------------------------------------- */
VREN_EXPORT uint32_t vren_depth_pyramid_level_count(uint32_t width, uint32_t height);
VREN_EXPORT uint64_t vren_depth_pyramid_level_offset(uint32_t width, uint32_t height, uint32_t level);
VREN_EXPORT uint64_t vren_depth_pyramid_length(uint32_t width, uint32_t height);
VREN_EXPORT int vren_depth_pyramid(vren_ctx* ctx, vren_stream_t stream, const float* depth, uint32_t width,
                                   uint32_t height, uint32_t images, float* pyramid);

/* ---------------------------------------------------------------------------------------------------------
 * Debug consumers of the demo (vren_demo): image and line dumps of the clustered-shading outputs
 * ------------------------------------------------------------------------------------------------------- */
/* show_clusters (clustered_shading_debug.cpp:52-100, show_clusters.comp:27-126): RGBA32F image of W*H pixels,
 * pixel p shows the cluster cluster_refs[p] (reference key layout): mode 0 hash(key), 1 (i / tiles_x, j / tiles_y,
 * k / 1024), 2 (face / 6, u / 3, v / 3) of the normal code, 3 min(count / 64, 1), 4 hash(offset), 5 hash(xor of
 * the cluster's light indices), any other mode (1, 0, 0); alpha 1. hash(x) = float(triple32(x)) / float(2^32 - 1).
 * A reference at or past key_capacity shows (1, 0, 0). */
VREN_EXPORT int vren_show_clusters(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height, int32_t mode,
                                   const uint32_t* cluster_refs, const uint32_t* cluster_keys, uint32_t key_capacity,
                                   const uint32_t* light_counts, const uint32_t* light_offsets,
                                   const uint32_t* light_indices, uint32_t index_capacity, float* output);
/* visualize_bvh (visualize_bvh.cpp:21-90, show_bvh.comp): for every node of a 32-ary BVH of level_count + 1
 * levels (leaves first, as build_bvh lays them out), 12 lines = 24 vertices {f32 x, y, z; u32 colour} (16 bytes,
 * debug_renderer.hpp:25-30) at vertices[24 * node]: the node's AABB edges in show_bvh.comp's order, coloured by
 * level (leaves 0x00ff00, then 0x0000ff, 0x00ffff, 0xff00ff, 0xffffff); invalid nodes give zero vertices of
 * colour 0. level_count = vren_calc_bvh_level_count(lights) <= 5. */
VREN_EXPORT int vren_visualize_bvh(vren_ctx* ctx, vren_stream_t stream, const vren_bvh_node* bvh, uint32_t level_count,
                                   float* vertices);

/* ---------------------------------------------------------------------------------------------------------
 * Synthetic inputs on the device (include/vren_hip/synth.h; no reference counterpart: the reference draws its
 * inputs from a rasterised glTF scene). initstate = 42 + 1000 * config (+ frame).
 * ------------------------------------------------------------------------------------------------------- */
/* n uint32 PCG32 draws of stream (initstate, initseq) (the sort workload's keys: initstate 42, initseq 54) */
VREN_EXPORT int vren_synth_keys(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint64_t initseq, uint64_t n,
                                uint32_t* out);
VREN_EXPORT int vren_synth_lights(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint32_t light_count,
                                  float* light_positions, float* point_lights);
VREN_EXPORT int vren_synth_frame(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint32_t width,
                                 uint32_t height, const vren_camera* camera, uint32_t n_boxes, float* depth,
                                 uint16_t* normals);
VREN_EXPORT int vren_synth_materials(vren_ctx* ctx, vren_stream_t stream, uint64_t initstate, uint64_t n_pixels,
                                     float* albedo_roughness, float* metallic);

#ifdef __cplusplus
}
#endif

#endif /* VREN_HIP_CAPI_H_ */
