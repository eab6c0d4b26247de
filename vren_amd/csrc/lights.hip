// lights.hip -- construct_point_light_bvh (clustered_shading.cpp:60-346) and build_bvh (build_bvh.cpp:38-99).
//
// Reference: 8 recorded steps and ~10 dispatches per frame (view transform, two vec4 reduce trees, a copy,
// Morton codes, a 65 536-bucket atomic bucket sort with a fixed 2^16 scan, leaf init, then one dispatch per BVH
// level). Here, per frame:
//   lights_kernel   one workgroup: view transform, AABB (order-independent min/max), Morton codes, and a stable
//                   LSD radix sort of (code, light index) in LDS (global ping-pong above 4096 lights);
//   bvh_leaf_kernel every leaf of every frame (flattened grid): leaf AABB = view position -+ radius, and the first
//                   parent level straight from registers (32-lane min/max over half-waves);
//   bvh_upper_kernel one workgroup per frame for the remaining (<= 32^k / 32) parent levels.
// A batch of frames runs the same three launches with one workgroup (or leaf range) per frame.
#include "chain.hpp"
#include "radix_device.hpp"
#include "scan_device.hpp"

namespace {

constexpr int kLT = 1024;           // threads of the per-frame light workgroup
constexpr int kLKPT = 4;            // keys per thread per sort chunk
constexpr int kLChunk = kLT * kLKPT;
constexpr uint32_t kLdsCap = 4096;  // lights sorted entirely in LDS

__device__ __forceinline__ uint32_t morton_interleave(uint32_t qx, uint32_t qy, uint32_t qz, uint32_t bits)
{
    uint32_t code = 0;
    for (uint32_t i = 0; i < bits; i++) {
        code |= ((qx >> i) & 1u) << (i * 3);
        code |= ((qy >> i) & 1u) << (i * 3 + 1);
        code |= ((qz >> i) & 1u) << (i * 3 + 2);
    }
    return code;
}

// one stable LSD pass of a BITS-bit digit (BITS <= 8) over a segment of L (code, index) pairs, src -> dst (generic
// pointers: LDS or global)
template <int BITS>
__device__ void block_sort_pass(const uint2* src, uint2* dst, uint32_t L, uint32_t shift, uint32_t width,
                                uint32_t* s_whist, uint32_t* s_base, uint32_t* s_cnt, uint32_t* s_scan)
{
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / VREN_WAVE;
    const uint32_t mask = (1u << width) - 1u;
    if (tid < 256) s_cnt[tid] = 0u;
    __syncthreads();
    for (uint32_t i = tid; i < L; i += kLT) atomicAdd(&s_cnt[(src[i].x >> shift) & mask], 1u);
    __syncthreads();
    uint32_t tot;
    uint32_t ex = block_exclusive_sum<kLT>(tid < 256 ? s_cnt[tid] : 0u, s_scan, &tot);
    if (tid < 256) s_base[tid] = ex;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < L; c0 += kLChunk) {
        const uint32_t cnt = min((uint32_t)kLChunk, L - c0);
        for (uint32_t i = tid; i < (uint32_t)(kLT / VREN_WAVE) * 256; i += kLT) s_whist[i] = 0u;
        __syncthreads();
        const uint32_t wbase = wave * kLKPT * VREN_WAVE;
        uint2 e[kLKPT];
        uint32_t digit[kLKPT], local[kLKPT];
#pragma unroll
        for (int j = 0; j < kLKPT; j++) {
            uint32_t pos = wbase + j * VREN_WAVE + lane;
            e[j] = pos < cnt ? src[c0 + pos] : make_uint2(0u, 0u);
            digit[j] = (e[j].x >> shift) & mask;
        }
        rank_wave_keys<BITS, kLKPT>(digit, wbase, cnt, &s_whist[wave * 256], local);
        __syncthreads();
        if (tid < 256) {
            uint32_t run = 0;
            for (int w = 0; w < kLT / VREN_WAVE; w++) {
                uint32_t c = s_whist[w * 256 + tid];
                s_whist[w * 256 + tid] = run;
                run += c;
            }
            s_cnt[tid] = run;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kLKPT; j++) {
            uint32_t pos = wbase + j * VREN_WAVE + lane;
            if (pos < cnt) {
                uint32_t d = digit[j];
                dst[s_base[d] + s_whist[wave * 256 + d] + local[j]] = e[j];
            }
        }
        __syncthreads();
        if (tid < 256) s_base[tid] += s_cnt[tid];
        __syncthreads();
    }
}

// BITS: digit width of the Morton-code sort (8 by default; 4 is a chain-shape choice)
template <int BITS>
__global__ __launch_bounds__(kLT) void lights_kernel(const FrameDev* __restrict__ frames)
{
    __shared__ uint2 s_a[kLdsCap];
    __shared__ uint2 s_b[kLdsCap];
    __shared__ uint32_t s_whist[(kLT / VREN_WAVE) * 256];
    __shared__ uint32_t s_base[256], s_cnt[256], s_scan[kLT / VREN_WAVE];
    __shared__ float s_red[6][kLT / VREN_WAVE];
    __shared__ float s_box[6];

    const FrameDev& fd = frames[blockIdx.x];
    const uint32_t L = fd.L;
    if (L == 0) return;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / VREN_WAVE;
    const float4* pos = reinterpret_cast<const float4*>(fd.light_pos);
    float4* vpos = reinterpret_cast<float4*>(fd.view_pos);

    // 1. view transform (point_light_position_to_view_space.comp:26-32) + AABB partials
    float mn[3] = {1e35f, 1e35f, 1e35f}, mx[3] = {-1e35f, -1e35f, -1e35f};
    for (uint32_t i = tid; i < L; i += kLT) {
        float4 p = pos[i];
        float v[4] = {p.x, p.y, p.z, 1.0f}, r[4];
        vren_mat4_mul_vec4(fd.view, v, r);
        vpos[i] = make_float4(r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = tmin(mn[a], r[a]);
            mx[a] = tmax(mx[a], r[a]);
        }
    }
    // 2. AABB (reduce_vec4 min/max of clustered_shading.cpp:137-191; min/max are exact in any order)
    if (fd.morton_aabb_mode == VREN_MORTON_GIVEN_AABB) {
        if (tid < 3) {
            s_box[tid] = fd.morton_min[tid];
            s_box[3 + tid] = fd.morton_max[tid];
        }
    } else {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                mn[a] = tmin(mn[a], shfl_xor(mn[a], m));
                mx[a] = tmax(mx[a], shfl_xor(mx[a], m));
            }
        }
        if (lane == 0) {
            for (int a = 0; a < 3; a++) {
                s_red[a][wave] = mn[a];
                s_red[3 + a][wave] = mx[a];
            }
        }
        __syncthreads();
        if (tid < 3) {
            float m0 = 1e35f, m1 = -1e35f;
            for (int w = 0; w < kLT / VREN_WAVE; w++) {
                m0 = tmin(m0, s_red[tid][w]);
                m1 = tmax(m1, s_red[3 + tid][w]);
            }
            s_box[tid] = m0;
            s_box[3 + tid] = m1;
        }
    }
    __syncthreads();

    // 3. Morton codes (discretize_point_light_positions.comp:26-43), paired with the light index
    const uint32_t bits = fd.morton_bits;
    const uint32_t total_bits = 3 * bits;
    const uint32_t passes = (total_bits + BITS - 1) / BITS;
    const float scale = (float)(1u << bits);
    const bool in_lds = L <= kLdsCap;
    uint2* A;
    uint2* B;
    if (in_lds) {
        A = s_a;
        B = s_b;
    } else {
        // end the even/odd ping-pong in fd.codes
        A = reinterpret_cast<uint2*>((passes & 1) ? fd.codes_tmp : fd.codes);
        B = reinterpret_cast<uint2*>((passes & 1) ? fd.codes : fd.codes_tmp);
    }
    for (uint32_t i = tid; i < L; i += kLT) {
        float4 p = pos[i];
        float v[4] = {p.x, p.y, p.z, 1.0f}, r[4];
        vren_mat4_mul_vec4(fd.view, v, r);
        uint32_t q[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float t = (r[a] - s_box[a]) / (s_box[3 + a] - s_box[a]) * scale;
            q[a] = vren_f2u_sat(vren_floorf(t));
        }
        A[i] = make_uint2(morton_interleave(q[0], q[1], q[2], bits), i);
    }
    __syncthreads();
    // 4. stable sort by code (ties keep light-index order)
    uint2* src = A;
    uint2* dst = B;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t shift = BITS * p;
        block_sort_pass<BITS>(src, dst, L, shift, min((uint32_t)BITS, total_bits - shift), s_whist, s_base, s_cnt,
                              s_scan);
        uint2* t = src;
        src = dst;
        dst = t;
    }
    if (in_lds) {
        uint2* out = reinterpret_cast<uint2*>(fd.codes);
        for (uint32_t i = tid; i < L; i += kLT) out[i] = src[i];
    }
}

__device__ __forceinline__ void half_wave_minmax(float (&mn)[3], float (&mx)[3])
{
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = tmin(mn[a], shfl_xor(mn[a], m));
            mx[a] = tmax(mx[a], shfl_xor(mx[a], m));
        }
    }
}

__device__ __forceinline__ void store_node(vren_bvh_node* n, const float (&mn)[3], const float (&mx)[3], uint32_t next)
{
    float4* q = reinterpret_cast<float4*>(n);
    q[0] = make_float4(mn[0], mn[1], mn[2], __uint_as_float(next));
    q[1] = make_float4(mx[0], mx[1], mx[2], 0.0f);
}

// leaves (init_light_array_bvh.comp:42-65) + the first parent level (build_bvh.comp:32-55). `span` selects the
// 32-leaf groups written: every group, only those holding a light (the groups past the last light are all INVALID
// leaves under an INVALID parent, which depend on the light count alone: a batch writes them once, with
// kLeavesPadding, when it is created), or only those padding groups.
__global__ __launch_bounds__(256) void bvh_leaf_kernel(const FrameDev* __restrict__ frames, uint32_t F, bool from_lights,
                                                       uint32_t span)
{
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    // the live span's grid covers each frame's light-holding groups only (numbered by live_leaf_base)
    const bool live_grid = span == kLeavesLive;
    const uint32_t f = live_grid ? find_frame(frames, F, blockIdx.x * 256, [](const FrameDev& d) { return d.live_leaf_base; })
                                 : find_frame(frames, F, blockIdx.x * 256, [](const FrameDev& d) { return d.leaf_base; });
    const FrameDev& fd = frames[f];
    const uint32_t i = g - (live_grid ? fd.live_leaf_base : fd.leaf_base);
    const uint32_t padded = fd.bvh_padded;
    if (i >= padded) return; // whole half-waves leave together (padded is a multiple of 32)
    const uint32_t live_end = (fd.L + 31u) & ~31u; // first 32-leaf group without a light
    if (span == kLeavesLive && i >= live_end) return;
    if (span == kLeavesPadding) {
        if (i < live_end) return;
        const float inf[3] = {VREN_F32_INF, VREN_F32_INF, VREN_F32_INF};
        const float ninf[3] = {-VREN_F32_INF, -VREN_F32_INF, -VREN_F32_INF};
        store_node(&fd.bvh[i], inf, ninf, VREN_BVH_INVALID_NODE);
        if ((lane_id() & 31) == 0) store_node(&fd.bvh[padded + i / 32], inf, ninf, VREN_BVH_INVALID_NODE);
        return;
    }
    float mn[3], mx[3];
    bool valid;
    if (from_lights) {
        valid = i < fd.L;
        if (valid) {
            uint32_t li = reinterpret_cast<const uint2*>(fd.codes)[i].y;
            float4 vp = reinterpret_cast<const float4*>(fd.view_pos)[li];
            float r = reinterpret_cast<const float4*>(fd.lights)[li].w;
            if (fd.leaf_pos) reinterpret_cast<float4*>(fd.leaf_pos)[i] = vp;
            mn[0] = vp.x - r;
            mn[1] = vp.y - r;
            mn[2] = vp.z - r;
            mx[0] = vp.x + r;
            mx[1] = vp.y + r;
            mx[2] = vp.z + r;
        } else {
            mn[0] = mn[1] = mn[2] = VREN_F32_INF;
            mx[0] = mx[1] = mx[2] = -VREN_F32_INF;
        }
        store_node(&fd.bvh[i], mn, mx, valid ? VREN_BVH_LEAF_NODE : VREN_BVH_INVALID_NODE);
    } else {
        const vren_bvh_node& n = fd.bvh[i];
        valid = n.next != VREN_BVH_INVALID_NODE;
        for (int a = 0; a < 3; a++) {
            mn[a] = valid ? n.min[a] : VREN_F32_INF;
            mx[a] = valid ? n.max[a] : -VREN_F32_INF;
        }
    }
    const uint64_t vm = ballot(valid);
    const uint32_t half = (lane_id() >> 5) << 5;
    const bool any = ((vm >> half) & 0xFFFFFFFFull) != 0ull;
    half_wave_minmax(mn, mx);
    if ((lane_id() & 31) == 0) {
        uint32_t gi = i / 32;
        store_node(&fd.bvh[padded + gi], mn, mx, any ? i : VREN_BVH_INVALID_NODE);
    }
}

// remaining parent levels, one workgroup per frame
__global__ __launch_bounds__(1024) void bvh_upper_kernel(const FrameDev* __restrict__ frames)
{
    const FrameDev& fd = frames[blockIdx.x];
    const uint32_t padded = fd.bvh_padded;
    uint32_t src = padded, size = padded / 32;
    const uint32_t tid = threadIdx.x;
    while (size > 1) {
        const uint32_t dst = src + size;
        // one half-wave per parent
        for (uint32_t base = 0; base < size; base += 1024) {
            uint32_t c = base + tid;
            bool in = c < size;
            float mn[3] = {VREN_F32_INF, VREN_F32_INF, VREN_F32_INF};
            float mx[3] = {-VREN_F32_INF, -VREN_F32_INF, -VREN_F32_INF};
            bool valid = false;
            if (in) {
                const float4* q = reinterpret_cast<const float4*>(&fd.bvh[src + c]);
                float4 lo = q[0], hi = q[1];
                valid = __float_as_uint(lo.w) != VREN_BVH_INVALID_NODE;
                if (valid) {
                    mn[0] = lo.x;
                    mn[1] = lo.y;
                    mn[2] = lo.z;
                    mx[0] = hi.x;
                    mx[1] = hi.y;
                    mx[2] = hi.z;
                }
            }
            const uint64_t vm = ballot(valid);
            const uint32_t half = (lane_id() >> 5) << 5;
            const bool any = ((vm >> half) & 0xFFFFFFFFull) != 0ull;
            half_wave_minmax(mn, mx);
            if (in && (lane_id() & 31) == 0) store_node(&fd.bvh[dst + c / 32], mn, mx, any ? src + c : VREN_BVH_INVALID_NODE);
        }
        __syncthreads();
        src = dst;
        size /= 32;
    }
}

} // namespace

int launch_lights(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w)
{
    if (w.light_bits == 4) lights_kernel<4><<<F, kLT, 0, s>>>(d_frames);
    else lights_kernel<8><<<F, kLT, 0, s>>>(d_frames);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    return VREN_OK;
}

// total_leaf_slots: the grid in leaf slots, the frames' live slots (live_leaf_base numbering) for kLeavesLive
int launch_bvh(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, uint32_t total_leaf_slots,
               bool leaves_from_lights, const ChainWork* w, uint32_t span)
{
    bvh_leaf_kernel<<<total_leaf_slots / 256, 256, 0, s>>>(d_frames, F, leaves_from_lights, span);
    VREN_LAUNCH_CHECK(ctx);
    if (span == kLeavesPadding) return VREN_OK;
    if (w) chain_mark(*w, s);
    bvh_upper_kernel<<<F, 1024, 0, s>>>(d_frames);
    VREN_LAUNCH_CHECK(ctx);
    if (w) chain_mark(*w, s);
    return VREN_OK;
}

static bool is_pow32(uint32_t n)
{
    if (n < 32) return false;
    while (n > 1) {
        if (n % 32) return false;
        n /= 32;
    }
    return true;
}

extern "C" {

VREN_EXPORT int vren_build_bvh(vren_ctx* ctx, vren_stream_t stream, vren_bvh_node* nodes, uint32_t padded_leaf_count)
{
    if (!ctx) return VREN_EINVAL;
    if (!nodes || !is_pow32(padded_leaf_count))
        return vren_fail(ctx, VREN_EINVAL, "build_bvh: leaf count must be a power of 32 and >= 32 (build_bvh.cpp:45)");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "build_bvh: context has no device");
    hipStream_t s = (hipStream_t)stream;
    FrameDev f;
    memset(&f, 0, sizeof(f));
    f.bvh = nodes;
    f.bvh_padded = padded_leaf_count;
    FrameDev* slot = vren_next_frame_slot(ctx);
    int st = vren_upload_frames(ctx, s, slot, &f, 1);
    if (st) return st;
    uint32_t slots = (padded_leaf_count + VREN_LEAF_ALIGN - 1) / VREN_LEAF_ALIGN * VREN_LEAF_ALIGN;
    return launch_bvh(ctx, s, slot, 1, slots, false, nullptr, kLeavesAll);
}

VREN_EXPORT int vren_construct_point_light_bvh(vren_ctx* ctx, vren_stream_t stream, const float* light_positions,
                                               const float* point_lights, uint32_t light_count,
                                               const vren_camera* camera, const vren_cluster_cfg* cfg,
                                               float* view_positions, uint32_t* light_codes, vren_bvh_node* bvh)
{
    if (!ctx) return VREN_EINVAL;
    int st = vren_validate_camera(ctx, camera);
    if (!st) st = vren_validate_cfg(ctx, cfg);
    if (st) return st;
    if (light_count == 0) return VREN_OK; // the reference skips the stage without lights (clustered_shading.cpp:957)
    if (!light_positions || !point_lights || !view_positions || !light_codes || !bvh)
        return vren_fail(ctx, VREN_EINVAL, "construct_point_light_bvh: null buffer");
    if (light_count > (1u << 20))
        return vren_fail(ctx, VREN_EINVAL, "construct_point_light_bvh: at most 2^20 lights (config.hpp:18)");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "construct_point_light_bvh: context has no device");
    hipStream_t s = (hipStream_t)stream;
    FrameDev f;
    memset(&f, 0, sizeof(f));
    vren_frame_setup(&f, 32, 32, camera, cfg);
    f.light_pos = light_positions;
    f.lights = point_lights;
    f.L = light_count;
    f.view_pos = view_positions;
    f.codes = light_codes;
    f.bvh = bvh;
    f.bvh_padded = vren_calc_bvh_padded_leaf_count(light_count);
    f.bvh_root = vren_calc_bvh_root_index(light_count);
    f.bvh_levels = vren_calc_bvh_level_count(light_count);
    f.bvh_length = vren_calc_bvh_buffer_length(light_count);
    if (light_count > kLdsCap) {
        st = vren_ensure_scratch(ctx, (size_t)light_count * 8, s);
        if (st) return st;
        f.codes_tmp = static_cast<uint32_t*>(ctx->scratch);
    }
    FrameDev* slot = vren_next_frame_slot(ctx);
    st = vren_upload_frames(ctx, s, slot, &f, 1);
    if (st) return st;
    ChainWork w;
    memset(&w, 0, sizeof(w));
    st = launch_lights(ctx, s, slot, 1, w);
    if (st) return st;
    uint32_t slots = (f.bvh_padded + VREN_LEAF_ALIGN - 1) / VREN_LEAF_ALIGN * VREN_LEAF_ALIGN;
    return launch_bvh(ctx, s, slot, 1, slots, true, nullptr, kLeavesAll);
}

} // extern "C"
