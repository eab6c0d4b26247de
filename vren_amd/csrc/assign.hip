// assign.hip -- assign_lights (clustered_shading.cpp:473-690, assign_lights.comp:77-241,
// clustered_shading.glsl:58-110).
//
// Reference: one wave32 per unique cluster (indirect dispatch), a DFS over the 32-ary BVH with a per-level
// ballot stack, run twice (COUNT then WRITE) around a fixed 2^17-entry Blelloch scan of the counts; lists come
// out in traversal order with reversed lanes inside each leaf block.
// Here: persistent kernels, one wave64 per cluster, the wave tests the 32 children of TWO pending BVH blocks
// at once (lanes 0-31 and 32-63) from a per-wave LDS stack, so no lane idles at the 32-wide fan-out; the
// number of clusters is read on the device (it never returns to the host). COUNT counts leaf hits; one
// workgroup per frame scans the counts into offsets; WRITE sets one bit per hit
// light in a per-wave LDS bitmask and emits the list in ascending light index (the canonical order) by a
// popcount prefix over the bitmask words.
// The cluster AABB and both tests are restated operation for operation (including the inverted y range of
// the two-corner AABB); padding leaves and invalid nodes (empty AABB) never match.
#include "chain.hpp"
#include "scan_device.hpp"

namespace {

constexpr uint32_t kStack = 384; // per-wave stack entries (levels are nondecreasing on the stack: <= 64 per level)
constexpr uint32_t kLevelShift = 28;
constexpr uint32_t kAddrMask = (1u << kLevelShift) - 1u;

struct Box {
    float mn[3], mx[3];
};

__device__ Box cluster_aabb(const FrameDev& fd, uint32_t key)
{
    uint32_t ci = key & 0xFFu, cj = (key >> 8) & 0xFFu, ck;
    float near_k, far_k;
    if (fd.key_mode == VREN_KEYS_SLICED) {
        ck = (key >> 16) & 0xFFu;
        float S = (float)fd.depth_slices;
        near_k = fd.near_plane * vren_expf((float)ck / S * fd.log_fn);
        far_k = fd.near_plane * vren_expf((float)(ck + 1) / S * fd.log_fn);
    } else {
        ck = (key >> 16) & 0x3FFu;
        near_k = fd.near_plane * vren_powi(fd.slice_a, ck);
        far_k = near_k * fd.slice_a;
    }
    const float fSx = (float)fd.Sx, fSy = (float)fd.Sy;
    const float x0 = (float)ci / fSx, y0 = (float)cj / fSy;
    const float x1 = (float)(ci + 1) / fSx, y1 = (float)(cj + 1) / fSy;
    Box b;
    if (fd.aabb_mode == VREN_CLUSTER_AABB_TIGHT) {
        for (int a = 0; a < 3; a++) {
            b.mn[a] = VREN_F32_INF;
            b.mx[a] = -VREN_F32_INF;
        }
        for (int corner = 0; corner < 4; corner++) {
            float cx = (corner & 1) ? x1 : x0, cy = (corner & 2) ? y1 : y0;
            float v[4] = {cx * 2.0f - 1.0f, (1.0f - cy) * 2.0f - 1.0f, 0.0f, 1.0f};
            float r[4];
            vren_mat4_mul_vec4(fd.inv_proj, v, r);
            float p[3] = {r[0] / r[3], r[1] / r[3], r[2] / r[3]};
            float d[3];
            vren_normalize3(p, d);
            for (int plane = 0; plane < 2; plane++) {
                float zp = plane ? far_k : near_k;
                float sc = zp / d[2];
                for (int a = 0; a < 3; a++) {
                    float q = sc * d[a];
                    b.mn[a] = tmin(b.mn[a], q);
                    b.mx[a] = tmax(b.mx[a], q);
                }
            }
        }
        return b;
    }
    float vmin[4] = {x0, 1.0f - y0, 0.0f, 1.0f};
    vmin[0] = vmin[0] * 2.0f - 1.0f;
    vmin[1] = vmin[1] * 2.0f - 1.0f;
    float vmax[4] = {x1, 1.0f - y1, 0.0f, 1.0f};
    vmax[0] = vmax[0] * 2.0f - 1.0f;
    vmax[1] = vmax[1] * 2.0f - 1.0f;
    float rmin[4], rmax[4];
    vren_mat4_mul_vec4(fd.inv_proj, vmin, rmin);
    vren_mat4_mul_vec4(fd.inv_proj, vmax, rmax);
    float pmin[3] = {rmin[0] / rmin[3], rmin[1] / rmin[3], rmin[2] / rmin[3]};
    float pmax[3] = {rmax[0] / rmax[3], rmax[1] / rmax[3], rmax[2] / rmax[3]};
    float d1[3], d2[3];
    vren_normalize3(pmin, d1);
    vren_normalize3(pmax, d2);
    const float s1 = near_k / d1[2], s2 = far_k / d2[2];
    for (int a = 0; a < 3; a++) {
        b.mn[a] = s1 * d1[a];
        b.mx[a] = s2 * d2[a];
    }
    return b;
}

__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }

// test_aabb_aabb (assign_lights.comp:84-94), cluster first
__device__ __forceinline__ bool test_aabb_aabb(const Box& c, const float4& lo, const float4& hi)
{
    return c.mx[0] >= lo.x && c.mn[0] <= hi.x && c.mx[1] >= lo.y && c.mn[1] <= hi.y && c.mx[2] >= lo.z &&
           c.mn[2] <= hi.z;
}

// assign_lights.comp:97-102 with r = (max.x - min.x) / 2 (assign_lights.comp:214)
__device__ __forceinline__ bool test_leaf(const Box& c, const float4& o, float leaf_min_x, float leaf_max_x)
{
    const float r = (leaf_max_x - leaf_min_x) / 2.0f;
    float d[3];
    const float oo[3] = {o.x, o.y, o.z};
    for (int a = 0; a < 3; a++) {
        float p = gmax(c.mn[a], gmin(oo[a], c.mx[a]));
        d[a] = p - oo[a];
    }
    return vren_sqrtf(vren_dot3(d, d)) < r;
}

// Traverse the BVH for one cluster with the whole wave. MODE 0 counts hits, MODE 1 sets bits in `bitmask`.
template <int MODE>
__device__ uint32_t traverse(const FrameDev& fd, const Box& box, uint32_t* stack, uint32_t* bitmask)
{
    const uint32_t lane = lane_id();
    const uint32_t sub = lane & 31u;
    const uint32_t leaf_level = fd.bvh_levels - 1;
    const vren_bvh_node* bvh = fd.bvh;
    const uint2* codes = reinterpret_cast<const uint2*>(fd.codes);
    const float4* vpos = reinterpret_cast<const float4*>(fd.view_pos);
    // level 0 = the root's children (assign_lights.comp:108-119): the first child of the root
    const uint32_t root_first = __float_as_uint(reinterpret_cast<const float4*>(&bvh[fd.bvh_root])[0].w);
    uint32_t sp = 0;
    if (root_first == VREN_BVH_INVALID_NODE) return 0u;
    if (lane == 0) stack[0] = root_first; // level 0
    sp = 1;
    uint32_t hits = 0;
    __builtin_amdgcn_wave_barrier();
    while (sp > 0) {
        uint32_t e0, e1 = 0xFFFFFFFFu;
        if (sp >= 2) {
            e0 = stack[sp - 2];
            e1 = stack[sp - 1];
            sp -= 2;
        } else {
            e0 = stack[sp - 1];
            sp -= 1;
        }
        const uint32_t e = lane < 32 ? e0 : e1;
        // entries always address a whole 32-node block inside the BVH; anything else is ignored
        const bool active = e != 0xFFFFFFFFu && (e >> kLevelShift) <= leaf_level &&
                            (e & kAddrMask) + 32u <= fd.bvh_length;
        bool leaf_hit = false, push = false;
        uint32_t child = 0, level = 0;
        if (active) {
            level = e >> kLevelShift;
            const uint32_t addr = (e & kAddrMask) + sub;
            const bool at_leaf = level == leaf_level && addr < fd.L;
            // a leaf's light index and view position are addressed by the leaf slot, so they load together with
            // the node (leaf_pos holds the positions in leaf order; without it the position needs the index)
            uint32_t li = 0xFFFFFFFFu;
            float4 lp = make_float4(0.f, 0.f, 0.f, 0.f);
            if (at_leaf) {
                li = codes[addr].y;
                if (fd.leaf_pos) lp = reinterpret_cast<const float4*>(fd.leaf_pos)[addr];
            }
            const float4* q = reinterpret_cast<const float4*>(&bvh[addr]);
            const float4 lo = q[0], hi = q[1];
            const uint32_t next = __float_as_uint(lo.w);
            if (level == leaf_level) {
                if (next == VREN_BVH_LEAF_NODE && at_leaf && li < fd.L) {
                    if (!fd.leaf_pos) lp = vpos[li];
                    leaf_hit = test_leaf(box, lp, lo.x, hi.x);
                    if (MODE == 1 && leaf_hit) atomicOr(&bitmask[li >> 5], 1u << (li & 31u));
                }
            } else if (next != VREN_BVH_INVALID_NODE && next != VREN_BVH_LEAF_NODE) {
                push = test_aabb_aabb(box, lo, hi);
                child = next | ((level + 1) << kLevelShift);
            }
        }
        hits += (uint32_t)__popcll(ballot(leaf_hit));
        const uint64_t pm = ballot(push);
        if (pm) {
            const uint32_t r = mbcnt(pm);
            if (push && sp + r < kStack) stack[sp + r] = child;
            sp += (uint32_t)__popcll(pm);
            if (sp > kStack) sp = kStack; // unreachable by the level bound; never write past the stack
        }
        __builtin_amdgcn_wave_barrier();
    }
    return hits;
}

// Clusters of every frame of the table are processed as one flattened range: each wave loads the device-side
// cluster counts of 64 frames at once (one per lane) and scans them, instead of walking the frames one dependent
// load after another. Calls fn(fd, cluster) for the wave's share (clusters gw, gw + stride, ...).
template <typename Fn>
__device__ __forceinline__ void for_each_cluster(const FrameDev* frames, uint32_t F, uint32_t gw, uint32_t stride, Fn fn)
{
    const uint32_t lane = lane_id();
    for (uint32_t f0 = 0; f0 < F; f0 += VREN_WAVE) {
        const uint32_t fi = f0 + lane;
        uint32_t nf = 0;
        if (fi < F) {
            const FrameDev& d = frames[fi];
            if (d.L != 0 && d.cluster_count != nullptr) nf = min(*d.cluster_count, d.key_cap);
        }
        const uint32_t incl = wave_inclusive_sum(nf);
        const uint32_t total = shfl(incl, VREN_WAVE - 1);
        for (uint32_t c = gw; c < total; c += stride) {
            const uint32_t fl = (uint32_t)__builtin_ctzll(ballot(incl > c));
            const uint32_t base = shfl(incl - nf, (int)fl);
            fn(frames[f0 + fl], c - base);
        }
    }
}

// Emits the lights set in a wave's bitmask in ascending index to dst[0 ..) (nothing when dst is null), clearing
// the mask. Returns the number emitted.
__device__ __forceinline__ uint32_t emit_bitmask(uint32_t* bitmask, uint32_t words, uint32_t* dst, uint32_t cap)
{
    const uint32_t lane = lane_id();
    uint32_t out = 0;
    for (uint32_t w0 = 0; w0 < words; w0 += VREN_WAVE) {
        const uint32_t wi = w0 + lane;
        uint32_t bits = wi < words ? bitmask[wi] : 0u;
        if (wi < words && bits) bitmask[wi] = 0u;
        const uint32_t cnt = (uint32_t)__popc(bits);
        const uint32_t incl = wave_inclusive_sum(cnt);
        uint32_t pos = out + incl - cnt;
        if (dst) {
            while (bits) {
                const uint32_t b = (uint32_t)__builtin_ctz(bits);
                bits &= bits - 1u;
                if (pos < cap) dst[pos] = wi * 32u + b;
                pos++;
            }
        }
        out += shfl(incl, VREN_WAVE - 1);
    }
    return out;
}

// COUNT (assign_lights.comp, count pass): the traversal marks the hit lights in the wave's LDS bitmask; the count
// is stored, and a list of at most VREN_LIST_SLOT lights is kept in the cluster's slot (ascending), so the writing
// pass copies it instead of traversing again.
__global__ void assign_count_kernel(const FrameDev* __restrict__ frames, uint32_t F, uint32_t mask_words)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t wpb = blockDim.x / VREN_WAVE;
    const uint32_t wave = wave_id();
    uint32_t* stack = s_mem + wave * kStack;
    uint32_t* bitmask = s_mem + wpb * kStack + wave * mask_words;
    for (uint32_t i = lane_id(); i < mask_words; i += VREN_WAVE) bitmask[i] = 0u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t gw = blockIdx.x * wpb + wave;
    for_each_cluster(frames, F, gw, gridDim.x * wpb, [&](const FrameDev& fd, uint32_t c) {
        const Box box = cluster_aabb(fd, fd.keys[c]);
        const uint32_t h = traverse<1>(fd, box, stack, bitmask);
        __builtin_amdgcn_wave_barrier();
        const bool keep = fd.list_slots && c < fd.slot_cap && h <= VREN_LIST_SLOT;
        emit_bitmask(bitmask, (fd.L + 31) / 32, keep ? fd.list_slots + (size_t)c * VREN_LIST_SLOT : nullptr,
                     VREN_LIST_SLOT);
        if (lane_id() == 0) fd.counts[c] = h;
        __builtin_amdgcn_wave_barrier();
    });
}

// exclusive scan of counts[0 .. min(cluster_count, key_cap)) -> offsets and assigned_total: one workgroup per
// frame walks its counts 4096 at a time with a running carry (no cross-workgroup chaining, no tickets; a frame
// holds a few thousand clusters, at most key_cap)
__global__ __launch_bounds__(1024) void assign_offsets_kernel(const FrameDev* __restrict__ frames)
{
    __shared__ uint32_t s_wave[16];
    const FrameDev& fd = frames[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const uint32_t n = (fd.L == 0 || fd.cluster_count == nullptr) ? 0u : min(*fd.cluster_count, fd.key_cap);
    uint32_t carry = 0;
    for (uint32_t start = 0; start < n; start += 4096) {
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = start + tid * 4 + k;
            v[k] = i < n ? fd.counts[i] : 0u;
            sum += v[k];
        }
        uint32_t agg;
        uint32_t run = carry + block_exclusive_sum<1024>(sum, s_wave, &agg);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = start + tid * 4 + k;
            if (i < n) fd.offsets[i] = run;
            run += v[k];
        }
        carry += agg;
    }
    if (tid == 0) {
        *fd.assigned_total = carry;
        if (carry > fd.index_cap) atomic_or_agent(fd.flags, VREN_FLAG_INDEX_CAPACITY);
    }
}

__global__ void assign_write_kernel(const FrameDev* __restrict__ frames, uint32_t F, uint32_t mask_words)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t wpb = blockDim.x / VREN_WAVE;
    const uint32_t wave = wave_id();
    const uint32_t lane = lane_id();
    uint32_t* stack = s_mem + wave * kStack;
    uint32_t* bitmask = s_mem + wpb * kStack + wave * mask_words;
    for (uint32_t i = lane; i < mask_words; i += VREN_WAVE) bitmask[i] = 0u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t gw = blockIdx.x * wpb + wave;
    for_each_cluster(frames, F, gw, gridDim.x * wpb, [&](const FrameDev& fd, uint32_t c) {
        const uint32_t cnt = fd.counts[c];
        const uint32_t out = fd.offsets[c];
        if (fd.list_slots && c < fd.slot_cap && cnt <= VREN_LIST_SLOT) {
            // the counting pass kept this list
            if (lane < cnt && (uint64_t)out + lane < fd.index_cap)
                fd.indices[out + lane] = fd.list_slots[(size_t)c * VREN_LIST_SLOT + lane];
            return;
        }
        const Box box = cluster_aabb(fd, fd.keys[c]);
        traverse<1>(fd, box, stack, bitmask);
        __builtin_amdgcn_wave_barrier();
        // emit set bits in ascending light index, clearing the mask as we go
        const uint32_t room = out < fd.index_cap ? fd.index_cap - out : 0u;
        emit_bitmask(bitmask, (fd.L + 31) / 32, fd.indices + out, room);
        __builtin_amdgcn_wave_barrier();
    });
}

// enough resident waves per CU (up to 32) to hide the dependent node loads of the traversals
uint32_t persistent_blocks(vren_ctx* ctx, uint32_t wpb) { return (uint32_t)ctx->num_cus * (32u / wpb); }

} // namespace

int launch_assign(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w)
{
    if (w.max_L == 0) return VREN_OK;
    const uint32_t mask_words = (w.max_L + 31) / 32;
    const size_t per_wave = (size_t)(kStack + mask_words) * 4;
    uint32_t wpb = 4;
    while (wpb > 1 && per_wave * wpb > 150 * 1024) wpb--;
    if (per_wave > 150 * 1024) return vren_fail(ctx, VREN_EINVAL, "assign_lights: too many lights for the LDS bitmask");
    assign_count_kernel<<<persistent_blocks(ctx, wpb), wpb * VREN_WAVE, per_wave * wpb, s>>>(d_frames, F, mask_words);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    assign_offsets_kernel<<<F, 1024, 0, s>>>(d_frames);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    assign_write_kernel<<<persistent_blocks(ctx, wpb), wpb * VREN_WAVE, per_wave * wpb, s>>>(d_frames, F, mask_words);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    return VREN_OK;
}

extern "C" VREN_EXPORT int vren_assign_lights(vren_ctx* ctx, vren_stream_t stream, uint32_t width, uint32_t height,
                                              const vren_camera* camera, const vren_cluster_cfg* cfg,
                                              const uint32_t* cluster_keys, const uint32_t* cluster_count,
                                              uint32_t key_capacity, const vren_bvh_node* bvh, uint32_t light_count,
                                              const uint32_t* light_codes, const float* view_positions,
                                              uint32_t* light_counts, uint32_t* light_offsets, uint32_t* light_indices,
                                              uint32_t index_capacity, uint32_t* assigned_total)
{
    if (!ctx) return VREN_EINVAL;
    int st = vren_validate_camera(ctx, camera);
    if (!st) st = vren_validate_cfg(ctx, cfg);
    if (st) return st;
    if (width == 0 || height == 0 || width > 8192 || height > 8192)
        return vren_fail(ctx, VREN_EINVAL, "assign_lights: screen must be in [1, 8192]^2");
    if (!cluster_keys || !cluster_count || !light_counts || !light_offsets || !assigned_total || key_capacity == 0)
        return vren_fail(ctx, VREN_EINVAL, "assign_lights: null buffer or zero key capacity");
    if (light_count > 0 && (!bvh || !light_codes || !view_positions || !light_indices || index_capacity == 0))
        return vren_fail(ctx, VREN_EINVAL, "assign_lights: null light buffer or zero index capacity");
    if (light_count > (1u << 20)) return vren_fail(ctx, VREN_EINVAL, "assign_lights: at most 2^20 lights");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "assign_lights: context has no device");
    hipStream_t s = (hipStream_t)stream;
    FrameDev f;
    memset(&f, 0, sizeof(f));
    vren_frame_setup(&f, width, height, camera, cfg);
    f.keys = const_cast<uint32_t*>(cluster_keys);
    f.cluster_count = const_cast<uint32_t*>(cluster_count);
    f.key_cap = key_capacity;
    f.bvh = const_cast<vren_bvh_node*>(bvh);
    f.L = light_count;
    f.codes = const_cast<uint32_t*>(light_codes);
    f.view_pos = const_cast<float*>(view_positions);
    f.bvh_padded = vren_calc_bvh_padded_leaf_count(light_count);
    f.bvh_root = vren_calc_bvh_root_index(light_count);
    f.bvh_levels = vren_calc_bvh_level_count(light_count);
    f.bvh_length = vren_calc_bvh_buffer_length(light_count);
    f.counts = light_counts;
    f.offsets = light_offsets;
    f.indices = light_indices;
    f.index_cap = index_capacity;
    f.assigned_total = assigned_total;
    f.flags = ctx->d_flags;
    f.offs_base = 0;
    ChainWork w;
    memset(&w, 0, sizeof(w));
    const uint32_t offs_tiles = (key_capacity + VREN_SCAN_TILE - 1) / VREN_SCAN_TILE;
    st = vren_ensure_scratch(ctx, chain_work_bytes(0, offs_tiles, 0));
    if (st) return st;
    chain_work_bind(&w, ctx->scratch, 0, offs_tiles, 0);
    w.max_L = light_count;
    FrameDev* slot = vren_next_frame_slot(ctx);
    st = vren_upload_frames(ctx, s, slot, &f, 1);
    if (!st) st = launch_chain_init(ctx, s, w);
    if (st) return st;
    if (light_count == 0) {
        // no lights: every list is empty
        VREN_HIP(ctx, hipMemsetAsync(light_counts, 0, (size_t)key_capacity * 4, s));
        VREN_HIP(ctx, hipMemsetAsync(light_offsets, 0, (size_t)key_capacity * 4, s));
        VREN_HIP(ctx, hipMemsetAsync(assigned_total, 0, 4, s));
        return VREN_OK;
    }
    return launch_assign(ctx, s, slot, 1, w);
}
