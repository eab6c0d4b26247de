// assign.hip -- assign_lights (clustered_shading.cpp:473-690, assign_lights.comp:77-241,
// clustered_shading.glsl:58-110).
//
// Reference: one wave32 per unique cluster (indirect dispatch), a DFS over the 32-ary BVH with a per-level
// ballot stack, run twice (COUNT then WRITE) around a fixed 2^17-entry Blelloch scan of the counts; lists come
// out in traversal order with reversed lanes inside each leaf block.
// Here: persistent kernels over the clusters of every frame (their number is read on the device; it never returns
// to the host), taken in chunks of up to 64 whose per-cluster work (the cluster AABB, the list bounds) is done one
// cluster per lane. COUNT runs two traversals per wave, one per 32-lane half (the BVH's fan-out), each from its own
// LDS stack, and keeps lists of at most VREN_LIST_SLOT (192) lights, sorted by rank, in per-cluster slots; one
// workgroup per frame scans the counts into offsets; WRITE copies the kept lists and traverses again only for longer
// ones, with the whole wave testing two BVH blocks per step, one bit per hit light in a per-wave LDS bitmask, emitted
// in ascending light index (the canonical order) by a popcount prefix over the bitmask words.
// The cluster AABB and both tests are restated operation for operation (including the inverted y range of
// the two-corner AABB); padding leaves and invalid nodes (empty AABB) never match.
#include "chain.hpp"
#include "scan_device.hpp"

namespace {

constexpr uint32_t kStack = 384; // per-wave stack entries (levels are nondecreasing on the stack: <= 64 per level)
// per-half stack entries of the counting pass: a half pops one block and pushes at most 32, so the stack holds at
// most 31 entries of each level above the deepest and 32 of the deepest, 31 * 3 + 32 = 125 for the four block
// levels of 2^20 lights
constexpr uint32_t kStackHalf = 192;
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

// Traverse the BVH for one cluster with the whole wave, setting one bit per hit light in the wave's LDS bitmask (and
// the word's bit in the summary, one bit per bitmask word), and return the number of hit lights.
__device__ uint32_t traverse(const FrameDev& fd, const Box& box, uint32_t* stack, uint32_t* bitmask,
                             uint32_t* summary)
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
                    if (leaf_hit) {
                        atomicOr(&bitmask[li >> 5], 1u << (li & 31u));
                        atomicOr(&summary[li >> 10], 1u << ((li >> 5) & 31u));
                    }
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

// The BVH of a frame as one half-wave's traversal needs it (per lane, so the two halves may serve two frames)
struct TravFrame {
    const vren_bvh_node* bvh;
    const uint2* codes;
    const float4* leaf_pos;
    const float4* vpos;
    uint32_t L, leaf_level, root, length;
};

__device__ __forceinline__ TravFrame trav_frame(const FrameDev& fd)
{
    return TravFrame{fd.bvh, reinterpret_cast<const uint2*>(fd.codes), reinterpret_cast<const float4*>(fd.leaf_pos),
                     reinterpret_cast<const float4*>(fd.view_pos), fd.L, fd.bvh_levels - 1, fd.bvh_root, fd.bvh_length};
}

// Two traversals at once, one per 32-lane half of the wave (lanes 0-31 and 32-63, each with its own frame, cluster
// box, stack and hit list; a half with live = false does nothing). Each step pops one 32-node block per half and
// tests its children with the half's 32 lanes, so the dependent node loads of two clusters overlap. Counts the hit
// lights of each half (returned to every lane of the half) and keeps the first VREN_LIST_SLOT hit light indices of
// each half in `list` (traversal order). Same node tests and visiting rules as traverse.
__device__ uint32_t traverse_halves(const TravFrame& tf, const Box& box, bool live, uint32_t* stack, uint32_t* list)
{
    const uint32_t lane = lane_id();
    const uint32_t sub = lane & 31u;
    const uint64_t half = lane < 32 ? 0x00000000FFFFFFFFull : 0xFFFFFFFF00000000ull;
    // level 0 = the root's children (assign_lights.comp:108-119): the first child of the root
    uint32_t sp = 0;
    if (live) {
        const uint32_t root_first = __float_as_uint(reinterpret_cast<const float4*>(&tf.bvh[tf.root])[0].w);
        if (root_first != VREN_BVH_INVALID_NODE) {
            if (sub == 0) stack[0] = root_first;
            sp = 1;
        }
    }
    uint32_t hits = 0;
    __builtin_amdgcn_wave_barrier();
    while (ballot(sp > 0)) {
        uint32_t e = 0xFFFFFFFFu;
        if (sp > 0) {
            e = stack[sp - 1];
            sp -= 1;
        }
        // entries always address a whole 32-node block inside the BVH; anything else is ignored
        const bool active = e != 0xFFFFFFFFu && (e >> kLevelShift) <= tf.leaf_level &&
                            (e & kAddrMask) + 32u <= tf.length;
        bool leaf_hit = false, push = false;
        uint32_t child = 0, li = 0xFFFFFFFFu;
        if (active) {
            const uint32_t level = e >> kLevelShift;
            const uint32_t addr = (e & kAddrMask) + sub;
            const bool at_leaf = level == tf.leaf_level && addr < tf.L;
            vf4 lpv = {0.f, 0.f, 0.f, 0.f};
            if (at_leaf) {
                li = *((guint*)tf.codes + 2 * (uint64_t)addr + 1);
                if (tf.leaf_pos) lpv = *((gvf4*)tf.leaf_pos + addr);
            }
            const vf4 lov = *((gvf4*)&tf.bvh[addr]), hiv = *((gvf4*)&tf.bvh[addr] + 1);
            const float4 lo = make_float4(lov.x, lov.y, lov.z, lov.w), hi = make_float4(hiv.x, hiv.y, hiv.z, hiv.w);
            float4 lp = make_float4(lpv.x, lpv.y, lpv.z, lpv.w);
            const uint32_t next = __float_as_uint(lo.w);
            if (level == tf.leaf_level) {
                if (next == VREN_BVH_LEAF_NODE && at_leaf && li < tf.L) {
                    if (!tf.leaf_pos) {
                        const vf4 v = *((gvf4*)tf.vpos + li);
                        lp = make_float4(v.x, v.y, v.z, v.w);
                    }
                    leaf_hit = test_leaf(box, lp, lo.x, hi.x);
                }
            } else if (next != VREN_BVH_INVALID_NODE && next != VREN_BVH_LEAF_NODE) {
                push = test_aabb_aabb(box, lo, hi);
                child = next | ((level + 1) << kLevelShift);
            }
        }
        const uint64_t hm = ballot(leaf_hit) & half;
        if (leaf_hit) {
            const uint32_t pos = hits + mbcnt(hm);
            if (pos < VREN_LIST_SLOT) list[pos] = li;
        }
        hits += (uint32_t)__popcll(hm);
        const uint64_t pm = ballot(push) & half;
        if (push) {
            const uint32_t r = mbcnt(pm);
            if (sp + r < kStackHalf) stack[sp + r] = child;
        }
        sp = min(sp + (uint32_t)__popcll(pm), kStackHalf); // the level bound keeps it below kStackHalf
        __builtin_amdgcn_wave_barrier();
    }
    return hits;
}

// Clusters per chunk: about four chunks per wave, strided over the whole range, so that every wave samples every
// part of the frames (the costly clusters, those covering many lights, sit together in the image), at most 64
// (one per lane).
__device__ __forceinline__ uint32_t chunk_size(uint32_t total, uint32_t waves)
{
    // 32-bit arithmetic: total <= 64 frames x 2^20 clusters and waves < 2^20, so nothing here wraps
    return max(1u, min(VREN_WAVE, (total + 4u * waves - 1u) / (4u * waves)));
}

// Clusters of every frame of the table are processed as one flattened range, in chunks of at most 64 consecutive
// clusters per wave: lane i takes cluster i of the chunk, finds its frame among 64 frames at a time (their
// device-side cluster counts, one per lane, prefix-summed) and computes that cluster's per-lane prologue `pre` (the
// cluster AABB, list bounds), so work that is uniform per cluster is done for a whole chunk at once. fn(fd, c, k, v)
// then runs for each cluster k of the chunk with the whole wave (v = lane k's prologue, read with readlane).
template <typename T, typename Pre, typename Fn>
__device__ __forceinline__ void for_each_cluster_chunk(const FrameDev* __restrict__ frames, uint32_t F, uint32_t gw,
                                                       uint32_t stride, Pre pre, Fn fn)
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
        const uint32_t excl = incl - nf;
        const uint32_t total = __builtin_amdgcn_readlane(incl, VREN_WAVE - 1);
                const uint32_t cs = chunk_size(total, stride);
        const uint32_t chunks = (total + cs - 1u) / cs;
        for (uint32_t chunk = gw; chunk < chunks; chunk += stride) {
            const uint32_t c0 = chunk * cs;
            const uint32_t c = c0 + lane;
            // frame of cluster c: the number of frames whose inclusive count is <= c (counts are non-decreasing)
            uint32_t pos = 0;
#pragma unroll
            for (uint32_t step = VREN_WAVE / 2; step >= 1; step >>= 1)
                if (shfl(incl, (int)(pos + step - 1)) <= c) pos += step;
            const uint32_t fl = min(pos, VREN_WAVE - 1);
            const uint32_t base = shfl(excl, (int)fl);
            T v{};
            if (lane < cs && c < total) v = pre(frames[f0 + fl], c - base);
            const uint32_t n = min(total - c0, cs);
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t fk = __builtin_amdgcn_readlane(fl, k);
                const uint32_t bk = __builtin_amdgcn_readlane(base, k);
                fn(frames[f0 + fk], c0 + k - bk, k, v);
            }
        }
    }
}

// for_each_cluster_chunk for the paired count pass: pre(fd, c, frame) runs per lane (lane i: cluster i of the
// chunk), then fn(k, n, cA, cB) for k = 0, 2, 4, ... with cA, cB the frame-local indices of clusters k and k + 1.
template <typename Pre, typename Fn>
__device__ __forceinline__ void for_each_cluster_chunk2(const FrameDev* __restrict__ frames, uint32_t F, uint32_t gw,
                                                        uint32_t stride, Pre pre, Fn fn)
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
        const uint32_t excl = incl - nf;
        const uint32_t total = __builtin_amdgcn_readlane(incl, VREN_WAVE - 1);
        const uint32_t cs = chunk_size(total, stride);
        const uint32_t chunks = (total + cs - 1u) / cs;
        // the chunks in eight contiguous blocks, one per XCD (workgroup b runs on XCD b mod 8), each block strided
        // over that XCD's waves: an L2 then holds the BVHs of an eighth of the frames (a grid that is not a
        // multiple of 8 workgroups keeps the plain stride)
        uint32_t first = gw, end = chunks, step = stride;
        if (gridDim.x % 8u == 0u) {
            const uint32_t xcd = blockIdx.x % 8u, wpb = blockDim.x / VREN_WAVE;
            first = (uint32_t)((uint64_t)chunks * xcd / 8u) + (blockIdx.x / 8u) * wpb + gw % wpb;
            end = (uint32_t)((uint64_t)chunks * (xcd + 1u) / 8u);
            step = stride / 8u;
        }
        for (uint32_t chunk = first; chunk < end; chunk += step) {
            const uint32_t c0 = chunk * cs;
            const uint32_t c = c0 + lane;
            uint32_t pos = 0;
#pragma unroll
            for (uint32_t step = VREN_WAVE / 2; step >= 1; step >>= 1)
                if (shfl(incl, (int)(pos + step - 1)) <= c) pos += step;
            const uint32_t fl = min(pos, VREN_WAVE - 1);
            const uint32_t base = shfl(excl, (int)fl);
            if (lane < cs && c < total) pre(frames[f0 + fl], c - base, f0 + fl);
            const uint32_t local = c - base;
            const uint32_t n = min(total - c0, cs);
            for (uint32_t k = 0; k < n; k += 2)
                fn(k, n, __builtin_amdgcn_readlane(local, k), __builtin_amdgcn_readlane(local, min(k + 1, n - 1)));
        }
    }
}



// Emits the lights set in a wave's bitmask in ascending index to dst[0 ..) (nothing when dst is null), clearing
// the bitmask and its summary, and returns their number. Only the words the summary marks are read: lane l of a
// step takes summary word s0 + l (bitmask words 32 (s0 + l) ..), counts its lights, and writes them after the
// lights of the lanes before it, so a cluster's list costs its hit words, not the L / 32 words of the mask.
__device__ __forceinline__ uint32_t emit_bitmask(uint32_t* bitmask, uint32_t* summary, uint32_t summary_words,
                                                 uint32_t* dst, uint32_t cap)
{
    const uint32_t lane = lane_id();
    uint32_t out = 0;
    for (uint32_t s0 = 0; s0 < summary_words; s0 += VREN_WAVE) {
        const uint32_t si = s0 + lane;
        const uint32_t sw = si < summary_words ? summary[si] : 0u;
        if (sw) summary[si] = 0u;
        uint32_t cnt = 0;
        for (uint32_t m = sw; m; m &= m - 1u) cnt += (uint32_t)__popc(bitmask[si * 32u + __builtin_ctz(m)]);
        const uint32_t incl = wave_inclusive_sum(cnt);
        uint32_t pos = out + incl - cnt;
        for (uint32_t m = sw; m; m &= m - 1u) {
            const uint32_t wi = si * 32u + (uint32_t)__builtin_ctz(m);
            uint32_t bits = bitmask[wi];
            bitmask[wi] = 0u;
            if (dst) {
                while (bits) {
                    const uint32_t b = (uint32_t)__builtin_ctz(bits);
                    bits &= bits - 1u;
                    if (pos < cap) dst[pos] = wi * 32u + b;
                    pos++;
                }
            }
        }
        out += __builtin_amdgcn_readlane(incl, VREN_WAVE - 1);
    }
    return out;
}

// COUNT (assign_lights.comp, count pass): two clusters per wave at a time (one per half, traverse_halves), each
// counting its hit lights and keeping the first VREN_LIST_SLOT of them in the half's LDS list; a list of at most
// VREN_LIST_SLOT lights is sorted (the rank of each index among the list) into the cluster's slot, so the writing
// pass copies it instead of traversing again.
__global__ __launch_bounds__(256, 8) void assign_count_kernel(const FrameDev* __restrict__ frames, uint32_t F,
                                                             uint32_t mask_words)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t wpb = blockDim.x / VREN_WAVE;
    const uint32_t wave = wave_id();
    const uint32_t lane = lane_id();
    const uint32_t hb = lane < 32 ? 0u : 1u; // half of the wave
    uint32_t* stack = s_mem + (wave * 2 + hb) * kStackHalf;
    uint32_t* list = s_mem + wpb * 2 * kStackHalf + (wave * 2 + hb) * VREN_LIST_SLOT;
    Box* boxes = reinterpret_cast<Box*>(s_mem + wpb * 2 * (kStackHalf + VREN_LIST_SLOT)) + wave * VREN_WAVE;
    uint32_t* cl_frame = s_mem + wpb * 2 * (kStackHalf + VREN_LIST_SLOT) + wpb * VREN_WAVE * 6 + wave * VREN_WAVE;
    const uint32_t gw = blockIdx.x * wpb + wave;
    // the chunk's cluster AABBs and frames wait in LDS (one per lane) while the wave traverses two at a time
    for_each_cluster_chunk2(
        frames, F, gw, gridDim.x * wpb,
        [&](const FrameDev& fdl, uint32_t c, uint32_t f) {
            boxes[lane] = cluster_aabb(fdl, fdl.keys[c]);
            cl_frame[lane] = f;
        },
        [&](uint32_t k, uint32_t n, uint32_t cA, uint32_t cB) {
            // half 0: cluster k of the chunk, half 1: cluster k + 1 (when the chunk has one)
            __builtin_amdgcn_wave_barrier();
            const uint32_t kk = k + hb;
            const bool live = kk < n;
            const uint32_t c = hb ? cB : cA;
            const FrameDev& fd = frames[cl_frame[live ? kk : k]];
            const Box box = boxes[live ? kk : k];
            const uint32_t h = traverse_halves(trav_frame(fd), box, live, stack, list);
            __builtin_amdgcn_wave_barrier();
            if (live && fd.list_slots && c < fd.slot_cap && h <= VREN_LIST_SLOT) {
                // ascending light index: each entry's rank among the (distinct) entries of the half's list, entries
                // sub and sub + 32 on lane sub of the half
                for (uint32_t q = lane & 31u; q < h; q += 32u) {
                    const uint32_t li = list[q];
                    uint32_t rank = 0;
                    for (uint32_t j = 0; j < h; j++) rank += list[j] < li ? 1u : 0u;
                    fd.list_slots[(size_t)c * VREN_LIST_SLOT + rank] = li;
                }
            }
            if (live && (lane & 31u) == 0) fd.counts[c] = h;
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

struct ListBounds {
    uint32_t count, offset;
};

// WRITE: each cluster's list at its offset: copied from the slot the counting pass kept, or traversed again
// (bitmask, emitted in ascending light index) when it was longer than a slot
__global__ __launch_bounds__(256, 8) void assign_write_kernel(const FrameDev* __restrict__ frames, uint32_t F, uint32_t mask_words)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t wpb = blockDim.x / VREN_WAVE;
    const uint32_t wave = wave_id();
    const uint32_t lane = lane_id();
    const uint32_t summary_words = (mask_words + 31) / 32;
    uint32_t* stack = s_mem + wave * kStack;
    uint32_t* bitmask = s_mem + wpb * kStack + wave * mask_words;
    uint32_t* summary = s_mem + wpb * (kStack + mask_words) + wave * summary_words;
    for (uint32_t i = lane; i < mask_words; i += VREN_WAVE) bitmask[i] = 0u;
    for (uint32_t i = lane; i < summary_words; i += VREN_WAVE) summary[i] = 0u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t gw = blockIdx.x * wpb + wave;
    for_each_cluster_chunk<ListBounds>(
        frames, F, gw, gridDim.x * wpb,
        [&](const FrameDev& fdl, uint32_t c) { return ListBounds{fdl.counts[c], fdl.offsets[c]}; },
        [&](const FrameDev& fd, uint32_t c, uint32_t k, const ListBounds& mine) {
            const uint32_t cnt = __builtin_amdgcn_readlane(mine.count, k);
            const uint32_t out = __builtin_amdgcn_readlane(mine.offset, k);
            if (fd.list_slots && c < fd.slot_cap && cnt <= VREN_LIST_SLOT) {
                // the counting pass kept this list
#pragma unroll
                for (uint32_t r = 0; r < VREN_LIST_SLOT / VREN_WAVE; r++) {
                    const uint32_t i = lane + r * VREN_WAVE;
                    if (i < cnt && (uint64_t)out + i < fd.index_cap)
                        fd.indices[out + i] = fd.list_slots[(size_t)c * VREN_LIST_SLOT + i];
                }
                return;
            }
            const Box box = cluster_aabb(fd, fd.keys[c]);
            traverse(fd, box, stack, bitmask, summary);
            __builtin_amdgcn_wave_barrier();
            // emit set bits in ascending light index, clearing the mask as we go
            const uint32_t room = out < fd.index_cap ? fd.index_cap - out : 0u;
            emit_bitmask(bitmask, summary, ((fd.L + 31) / 32 + 31) / 32, fd.indices + out, room);
            __builtin_amdgcn_wave_barrier();
        });
}

// enough resident waves per CU (32 by default, the chain shape's choice otherwise) to hide the dependent node loads
// of the traversals
uint32_t persistent_blocks(vren_ctx* ctx, uint32_t wpb, uint32_t waves_per_cu)
{
    const uint32_t wpc = waves_per_cu ? waves_per_cu : 32u;
    return (uint32_t)ctx->num_cus * (wpc / wpb > 0 ? wpc / wpb : 1u);
}

} // namespace

int launch_assign(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w)
{
    if (w.max_L == 0) return VREN_OK;
    const uint32_t mask_words = (w.max_L + 31) / 32;
    const size_t per_wave = (size_t)(kStack + mask_words + (mask_words + 31) / 32) * 4;
    uint32_t wpb = 4;
    while (wpb > 1 && per_wave * wpb > 150 * 1024) wpb--;
    if (per_wave > 150 * 1024) return vren_fail(ctx, VREN_EINVAL, "assign_lights: too many lights for the LDS bitmask");
    // two stacks and two lists per wave (one per half), the chunk's boxes and cluster frames
    constexpr size_t count_per_wave = (size_t)2 * (kStackHalf + VREN_LIST_SLOT) * 4 + VREN_WAVE * (sizeof(float) * 6 + 4);
    static_assert(count_per_wave * 4 <= 160 * 1024 / 8, "assign_count: 8 workgroups of 4 waves must fit a CU's LDS");
    assign_count_kernel<<<persistent_blocks(ctx, 4, w.assign_wpc), 4 * VREN_WAVE, count_per_wave * 4, s>>>(d_frames, F,
                                                                                                    mask_words);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    assign_offsets_kernel<<<F, 1024, 0, s>>>(d_frames);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    assign_write_kernel<<<persistent_blocks(ctx, wpb, w.assign_wpc), wpb * VREN_WAVE, per_wave * wpb, s>>>(d_frames, F,
                                                                                                   mask_words);
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
    st = vren_ensure_scratch(ctx, chain_work_bytes(0, offs_tiles, 0), s);
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
