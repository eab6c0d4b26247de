// lbvh.hip -- Karras binary radix-tree BVH over sorted light codes, with a float32 AABB refit (no reference
// counterpart: a binary BVH of the 4096-light workload, beside the reference's 32-ary one). Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees", HPG 2012: every internal node finds its key range and split independently (figure 4), so the
// topology is built with one thread per internal node; the AABBs are refitted bottom-up, each leaf climbing
// until it is the first of its parent's two children to arrive (an agent-scope acquire-release counter per
// internal node), the second arrival computing the parent's box from both children. min/max are exact and
// order-independent, so the result does not depend on the arrival order.
#include "common.hpp"

namespace {

__device__ __forceinline__ uint64_t lbvh_key(const uint2* codes, uint32_t i)
{
    const uint2 c = codes[i];
    return ((uint64_t)c.x << 32) | c.y;
}

__device__ __forceinline__ int lbvh_delta(const uint2* codes, int64_t n, uint64_t ki, int64_t j)
{
    if (j < 0 || j >= n) return -1;
    const uint64_t x = ki ^ lbvh_key(codes, (uint32_t)j);
    return x ? __builtin_clzll(x) : 64;
}

__global__ __launch_bounds__(256) void lbvh_internal_kernel(const uint2* codes, uint32_t n, vren_lbvh_node* nodes,
                                                            uint32_t* parents, uint32_t* counters)
{
    const int64_t N = n;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) parents[0] = 0xFFFFFFFFu; // the root (an internal node, or the only leaf when n == 1)
    if (i + 1 >= N) return;
    counters[i] = 0u;
    const uint64_t ki = lbvh_key(codes, (uint32_t)i);
    const int d = lbvh_delta(codes, N, ki, i + 1) - lbvh_delta(codes, N, ki, i - 1) > 0 ? 1 : -1;
    const int dmin = lbvh_delta(codes, N, ki, i - d);
    int64_t lmax = 2;
    while (lbvh_delta(codes, N, ki, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (lbvh_delta(codes, N, ki, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = lbvh_delta(codes, N, ki, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (lbvh_delta(codes, N, ki, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const uint32_t left = (uint32_t)(min(i, j) == gamma ? N - 1 + gamma : gamma);
    const uint32_t right = (uint32_t)(max(i, j) == gamma + 1 ? N - 1 + gamma + 1 : gamma + 1);
    nodes[i].left = left;
    nodes[i].right = right;
    parents[left] = (uint32_t)i;
    parents[right] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void lbvh_refit_kernel(const uint2* codes, uint32_t n, const float4* view_pos,
                                                         const float4* lights, vren_lbvh_node* nodes,
                                                         const uint32_t* parents, uint32_t* counters)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t li = codes[i].y;
    const float4 vp = view_pos[li];
    const float r = lights[li].w;
    vren_lbvh_node& leaf = nodes[n - 1 + i];
    leaf.min[0] = vp.x - r;
    leaf.min[1] = vp.y - r;
    leaf.min[2] = vp.z - r;
    leaf.max[0] = vp.x + r;
    leaf.max[1] = vp.y + r;
    leaf.max[2] = vp.z + r;
    leaf.left = VREN_BVH_LEAF_NODE;
    leaf.right = li;
    uint32_t node = parents[n - 1 + i];
    // at most 64 levels (one key bit per level); the bound only guards against malformed input
    for (int level = 0; level < 80 && node != 0xFFFFFFFFu; level++) {
        // release this subtree's box, acquire the sibling's
        if (__hip_atomic_fetch_add(&counters[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0u) break;
        vren_lbvh_node& p = nodes[node];
        const vren_lbvh_node& a = nodes[p.left];
        const vren_lbvh_node& b = nodes[p.right];
        for (int ax = 0; ax < 3; ax++) {
            p.min[ax] = tmin(a.min[ax], b.min[ax]);
            p.max[ax] = tmax(a.max[ax], b.max[ax]);
        }
        node = parents[node];
    }
}

} // namespace

extern "C" VREN_EXPORT int vren_build_lbvh(vren_ctx* ctx, vren_stream_t stream, const uint32_t* light_codes,
                                           uint32_t n, const float* view_positions, const float* point_lights,
                                           vren_lbvh_node* nodes, uint32_t* parents)
{
    if (!ctx) return VREN_EINVAL;
    if (!light_codes || !view_positions || !point_lights || !nodes || !parents || n == 0)
        return vren_fail(ctx, VREN_EINVAL, "build_lbvh: null buffer or n == 0");
    if (n > (1u << 20)) return vren_fail(ctx, VREN_EINVAL, "build_lbvh: at most 2^20 leaves");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "build_lbvh: context has no device");
    hipStream_t s = (hipStream_t)stream;
    int st = vren_ensure_scratch(ctx, (size_t)n * 4, s);
    if (st) return st;
    uint32_t* counters = static_cast<uint32_t*>(ctx->scratch);
    const uint2* codes = reinterpret_cast<const uint2*>(light_codes);
    const uint32_t blocks = (n + 255) / 256;
    lbvh_internal_kernel<<<blocks, 256, 0, s>>>(codes, n, nodes, parents, counters);
    VREN_LAUNCH_CHECK(ctx);
    lbvh_refit_kernel<<<blocks, 256, 0, s>>>(codes, n, reinterpret_cast<const float4*>(view_positions),
                                             reinterpret_cast<const float4*>(point_lights), nodes, parents, counters);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}
