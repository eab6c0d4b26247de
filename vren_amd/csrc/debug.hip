// debug.hip -- the demo's debug consumers of the clustered-shading outputs, as device kernels whose results can be
// dumped and compared: show_clusters (clustered_shading_debug.cpp:52-100, show_clusters.comp:27-126), a per-pixel
// view of the cluster keys and light lists, and visualize_bvh (visualize_bvh.cpp:21-90, show_bvh.comp), the light
// BVH as a line list of node boxes.
#include "common.hpp"

namespace {

// triple32 integer hash (show_clusters.comp:48-58) and hash(x) = float(triple32(x)) / float(0xffffffff)
__host__ __device__ inline uint32_t triple32(uint32_t x)
{
    x ^= x >> 17;
    x *= 0xed5ad4bbu;
    x ^= x >> 11;
    x *= 0xac4c1b51u;
    x ^= x >> 15;
    x *= 0x31848babu;
    x ^= x >> 14;
    return x;
}

__device__ __forceinline__ float hash01(uint32_t x) { return (float)triple32(x) / (float)0xffffffffu; }

__global__ __launch_bounds__(256) void show_clusters_kernel(uint32_t W, uint32_t H, int32_t mode, uint32_t tiles_x,
                                                            uint32_t tiles_y, const uint32_t* __restrict__ refs,
                                                            const uint32_t* __restrict__ keys, uint32_t key_cap,
                                                            const uint32_t* __restrict__ counts,
                                                            const uint32_t* __restrict__ offsets,
                                                            const uint32_t* __restrict__ indices, uint32_t index_cap,
                                                            float4* __restrict__ out)
{
    const uint32_t x = blockIdx.x * 32 + threadIdx.x % 32, y = blockIdx.y * 8 + threadIdx.x / 32;
    if (x >= W || y >= H) return;
    const uint64_t p = (uint64_t)y * W + x;
    const uint32_t ci = refs[p];
    float3 c = make_float3(1.0f, 0.0f, 0.0f); // an unknown mode (and a reference past the key table) shows red
    if (ci < key_cap) {
        const uint32_t key = keys[ci];
        const uint32_t i = key & 0xFFu, j = (key >> 8) & 0xFFu, k = (key >> 16) & 0x3FFu, n = key >> 26;
        if (mode == 0) {
            const float h = hash01(key);
            c = make_float3(h, h, h);
        } else if (mode == 1) {
            c = make_float3((float)i / (float)tiles_x, (float)j / (float)tiles_y, (float)k / 1024.0f);
        } else if (mode == 2) {
            c = make_float3((float)(n / 9) / 6.0f, (float)((n / 3) % 3) / 3.0f, (float)(n % 3) / 3.0f);
        } else if (mode == 3) {
            const float v = fminf((float)counts[ci] / 64.0f, 1.0f);
            c = make_float3(v, v, v);
        } else if (mode == 4) {
            const float h = hash01(offsets[ci]);
            c = make_float3(h, h, h);
        } else if (mode == 5) {
            uint32_t hx = 0;
            const uint32_t o = offsets[ci], e = counts[ci];
            for (uint32_t q = 0; q < e && (uint64_t)o + q < index_cap; q++) hx ^= indices[o + q];
            const float h = hash01(hx);
            c = make_float3(h, h, h);
        }
    }
    out[p] = make_float4(c.x, c.y, c.z, 1.0f);
}

struct DebugVertex {
    float x, y, z;
    uint32_t color;
};

// the 12 edges of an AABB in the order of show_bvh.comp's VREN_WRITE_DEBUG_DRAW_BUFFER_AABB: corner bits (x, y, z)
// of each edge's two ends, 1 = max
__constant__ uint8_t kEdges[12][2] = {{0b000, 0b100}, {0b100, 0b101}, {0b101, 0b001}, {0b001, 0b000},
                                      {0b010, 0b110}, {0b110, 0b111}, {0b111, 0b011}, {0b011, 0b010},
                                      {0b000, 0b010}, {0b100, 0b110}, {0b101, 0b111}, {0b001, 0b011}};

// one thread per (node, edge): writes the edge's two vertices (32 contiguous bytes)
__global__ __launch_bounds__(256) void visualize_bvh_kernel(const vren_bvh_node* __restrict__ bvh, uint32_t level_count,
                                                            uint64_t n_nodes, DebugVertex* __restrict__ out)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_nodes * 12) return;
    const uint64_t node = g / 12;
    const uint32_t edge = (uint32_t)(g - node * 12);
    // levels are written from the leaves (level_count) up to the root (0); colours by level as visualize_bvh.cpp:69-74
    const uint32_t colors[7] = {0xff0000u, 0xffff00u, 0x00ff00u, 0x0000ffu, 0x00ffffu, 0xff00ffu, 0xffffffu};
    uint64_t start = 0;
    uint32_t level = level_count;
    for (;; level--) {
        const uint64_t cnt = 1ull << (5 * level);
        if (node < start + cnt || level == 0) break;
        start += cnt;
    }
    const uint32_t ci = level_count - level + 2;
    const vren_bvh_node nd = bvh[node];
    DebugVertex a, b;
    if (nd.next != VREN_BVH_INVALID_NODE) {
        const uint32_t color = ci < 7 ? colors[ci] : 0xffffffu; // past the reference's table: white
        const uint8_t e0 = kEdges[edge][0], e1 = kEdges[edge][1];
        a = DebugVertex{(e0 & 4) ? nd.max[0] : nd.min[0], (e0 & 2) ? nd.max[1] : nd.min[1],
                        (e0 & 1) ? nd.max[2] : nd.min[2], color};
        b = DebugVertex{(e1 & 4) ? nd.max[0] : nd.min[0], (e1 & 2) ? nd.max[1] : nd.min[1],
                        (e1 & 1) ? nd.max[2] : nd.min[2], color};
    } else {
        a = DebugVertex{0.0f, 0.0f, 0.0f, 0u};
        b = a;
    }
    out[2 * g] = a;
    out[2 * g + 1] = b;
}

} // namespace

extern "C" {

VREN_EXPORT int vren_show_clusters(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height, int32_t mode,
                                   const uint32_t* cluster_refs, const uint32_t* cluster_keys, uint32_t key_capacity,
                                   const uint32_t* light_counts, const uint32_t* light_offsets,
                                   const uint32_t* light_indices, uint32_t index_capacity, float* output)
{
    if (!ctx) return VREN_EINVAL;
    if (width == 0 || height == 0 || width > 8192 || height > 8192)
        return vren_fail(ctx, VREN_EINVAL, "show_clusters: screen must be in [1, 8192]^2");
    if (!cluster_refs || !cluster_keys || !output || ((mode == 3 || mode == 4 || mode == 5) &&
                                                       (!light_counts || !light_offsets || (mode == 5 && !light_indices))))
        return vren_fail(ctx, VREN_EINVAL, "show_clusters: null buffer for mode %d", mode);
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "show_clusters: context has no device");
    const uint32_t tx = (width + 31) / 32, ty = (height + 31) / 32; // clustered_shading_debug.cpp:68-73
    dim3 grid(tx, (height + 7) / 8);
    show_clusters_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(width, height, mode, tx, ty, cluster_refs,
                                                                cluster_keys, key_capacity, light_counts,
                                                                light_offsets, light_indices, index_capacity,
                                                                reinterpret_cast<float4*>(output));
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

VREN_EXPORT int vren_visualize_bvh(vren_ctx* ctx, vren_stream_t stream, const vren_bvh_node* bvh, uint32_t level_count,
                                   float* vertices)
{
    if (!ctx) return VREN_EINVAL;
    if (!bvh || !vertices || level_count > 5)
        return vren_fail(ctx, VREN_EINVAL, "visualize_bvh: null buffer or level_count > 5");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "visualize_bvh: context has no device");
    uint64_t n = 0;
    for (uint32_t l = 0; l <= level_count; l++) n += 1ull << (5 * l);
    const uint64_t threads = n * 12;
    visualize_bvh_kernel<<<(uint32_t)((threads + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        bvh, level_count, n, reinterpret_cast<DebugVertex*>(vertices));
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

} // extern "C"
