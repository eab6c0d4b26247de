// reduce.hip -- reduce<T, Op> (reduce.hpp:15-49, reduce.cpp:32-140, reduce.comp:50-87) for gfx950.
//
// The reference output is the whole in-place Blelloch up-sweep tree of the pow2-padded input:
//   tree[i] = OP over the aligned block of 2^k elements ending at i, k = number of trailing 1 bits of i,
// combined as a balanced binary tree (each node = OP(right half, left half), reduce.comp:76).
// Here one pass handles TILE consecutive elements of a (possibly strided) view per workgroup: every thread
// up-sweeps ITEMS consecutive elements in registers, waves finish the tile with shuffles, and the tile's
// elements are written back in place of the ones read. The tile totals (at stride TILE) are up-swept by the
// next pass over the strided view, and so on: ceil(log2(p) / log2(TILE)) passes instead of the reference's
// one dispatch per 10 levels with 1024-element LDS trees, and every pass but the first touches only
// p / TILE^k elements.
#include "common.hpp"

namespace {

struct OpAddU {
    __device__ __host__ static uint32_t identity() { return 0u; }
    __device__ static uint32_t apply(uint32_t right, uint32_t left) { return right + left; }
};
struct OpMinU {
    __device__ __host__ static uint32_t identity() { return 0xFFFFFFFFu; }
    __device__ static uint32_t apply(uint32_t right, uint32_t left) { return left < right ? left : right; }
};
struct OpMaxU {
    __device__ __host__ static uint32_t identity() { return 0u; }
    __device__ static uint32_t apply(uint32_t right, uint32_t left) { return right < left ? left : right; }
};

// GLSL min(x, y) = y < x ? y : x applied as OP(s[b], s[a]) with x = right, y = left
struct OpAddV {
    __device__ static float4 identity() { return make_float4(0.f, 0.f, 0.f, 0.f); }
    __device__ static float4 apply(float4 r, float4 l) { return make_float4(r.x + l.x, r.y + l.y, r.z + l.z, r.w + l.w); }
};
struct OpMinV {
    __device__ static float4 identity() { return make_float4(1e35f, 1e35f, 1e35f, 1e35f); }
    __device__ static float m(float x, float y) { return y < x ? y : x; }
    __device__ static float4 apply(float4 r, float4 l) { return make_float4(m(r.x, l.x), m(r.y, l.y), m(r.z, l.z), m(r.w, l.w)); }
};
struct OpMaxV {
    __device__ static float4 identity() { return make_float4(-1e35f, -1e35f, -1e35f, -1e35f); }
    __device__ static float m(float x, float y) { return x < y ? y : x; }
    __device__ static float4 apply(float4 r, float4 l) { return make_float4(m(r.x, l.x), m(r.y, l.y), m(r.z, l.z), m(r.w, l.w)); }
};

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, unsigned d) { return shfl_up(v, d); }
template <>
__device__ __forceinline__ float4 shfl_up_t<float4>(float4 v, unsigned d)
{
    return make_float4(shfl_up(v.x, d), shfl_up(v.y, d), shfl_up(v.z, d), shfl_up(v.w, d));
}

constexpr int kThreads = 256;

// One tile of the view: elements e = tile * TILE + t, stored at base[offset + e * stride] (e < count),
// identity beyond `valid`. Writes back every element with index < count.
template <typename T, typename Op, int ITEMS>
__global__ __launch_bounds__(kThreads) void upsweep_tile_kernel(const T* in, T* out,
                                                                uint64_t in_seg_stride, uint64_t out_seg_stride,
                                                                uint64_t valid, uint64_t count, uint64_t offset,
                                                                uint64_t stride)
{
    constexpr int TILE = kThreads * ITEMS;
    __shared__ T s_wave[kThreads / VREN_WAVE];
    const uint32_t tid = threadIdx.x;
    const uint64_t seg = blockIdx.y;
    const T* src = in + seg * in_seg_stride;
    T* dst = out + seg * out_seg_stride;
    const uint64_t tile_start = (uint64_t)blockIdx.x * TILE;
    const uint64_t first = tile_start + (uint64_t)tid * ITEMS;

    T v[ITEMS];
    if (stride == 1 && offset == 0 && first + ITEMS <= valid && (sizeof(T) * ITEMS) % 16 == 0 &&
        (reinterpret_cast<uintptr_t>(src + first) & 15) == 0) {
        // vectorised contiguous load
        const uint4* p = reinterpret_cast<const uint4*>(src + first);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(T) * ITEMS / 16); k++) {
            uint4 q = p[k];
            memcpy(reinterpret_cast<char*>(v) + 16 * k, &q, 16);
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            uint64_t e = first + k;
            v[k] = e < valid ? src[offset + e * stride] : Op::identity();
        }
    }
    // levels inside the thread's ITEMS elements
#pragma unroll
    for (int l = 1; l < ITEMS; l <<= 1) {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            if ((k & (2 * l - 1)) == 2 * l - 1) v[k] = Op::apply(v[k], v[k - l]);
        }
    }
    // levels across threads: the node value of thread t is v[ITEMS - 1]
    T x = v[ITEMS - 1];
    const uint32_t lane = lane_id();
#pragma unroll
    for (unsigned l = 1; l < VREN_WAVE; l <<= 1) {
        T o = shfl_up_t(x, l);
        if ((lane & (2 * l - 1)) == 2 * l - 1) x = Op::apply(x, o);
    }
    const uint32_t wave = tid / VREN_WAVE;
    if (lane == VREN_WAVE - 1) s_wave[wave] = x;
    __syncthreads();
    if (lane == VREN_WAVE - 1) {
        // levels across the waves of the block (kThreads / 64 = 4 -> 2 levels)
        T w[kThreads / VREN_WAVE];
#pragma unroll
        for (int i = 0; i < kThreads / VREN_WAVE; i++) w[i] = s_wave[i];
#pragma unroll
        for (int l = 1; l < kThreads / VREN_WAVE; l <<= 1) {
#pragma unroll
            for (int i = kThreads / VREN_WAVE - 1; i >= 0; i--) {
                if ((i & (2 * l - 1)) == 2 * l - 1) w[i] = Op::apply(w[i], w[i - l]);
            }
        }
        x = w[wave];
    }
    v[ITEMS - 1] = x;
    if (stride == 1 && offset == 0 && first + ITEMS <= count && (sizeof(T) * ITEMS) % 16 == 0 &&
        (reinterpret_cast<uintptr_t>(dst + first) & 15) == 0) {
        uint4* p = reinterpret_cast<uint4*>(dst + first);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(T) * ITEMS / 16); k++) {
            uint4 q;
            memcpy(&q, reinterpret_cast<const char*>(v) + 16 * k, 16);
            p[k] = q;
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            uint64_t e = first + k;
            if (e < count) dst[offset + e * stride] = v[k];
        }
    }
}

template <typename T, typename Op, int ITEMS>
int reduce_tree(vren_ctx* ctx, hipStream_t s, const T* in, uint32_t n, T* out, uint32_t blocks)
{
    constexpr uint64_t TILE = (uint64_t)kThreads * ITEMS;
    const uint64_t p = next_pow2_u32(n);
    // pass 0: in (n valid, segment stride n) -> out (p elements, segment stride p)
    uint64_t count = p, offset = 0, stride = 1;
    bool first = true;
    while (true) {
        uint64_t tiles = ceil_div_u64(count, TILE);
        dim3 grid((uint32_t)tiles, blocks);
        if (first) {
            upsweep_tile_kernel<T, Op, ITEMS><<<grid, kThreads, 0, s>>>(in, out, n, p, n, count, 0, 1);
        } else {
            upsweep_tile_kernel<T, Op, ITEMS><<<grid, kThreads, 0, s>>>(out, out, p, p, count, count, offset, stride);
        }
        VREN_LAUNCH_CHECK(ctx);
        if (count <= TILE) break;
        // next view: the tile totals
        offset = offset + (TILE - 1) * stride;
        stride = stride * TILE;
        count = count / TILE;
        first = false;
    }
    return VREN_OK;
}

// ---- scalar total ------------------------------------------------------------------------------------------

template <typename Op>
__global__ __launch_bounds__(kThreads) void total_partial_kernel(const uint32_t* __restrict__ in, uint64_t n,
                                                                 uint32_t* __restrict__ partials)
{
    __shared__ uint32_t s[kThreads / VREN_WAVE];
    uint32_t acc = Op::identity();
    const uint64_t n4 = (reinterpret_cast<uintptr_t>(in) & 15) == 0 ? n / 4 : 0;
    const uint4* in4 = reinterpret_cast<const uint4*>(in);
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * kThreads) {
        uint4 q = ntload4(&in4[i]);
        acc = Op::apply(acc, q.x);
        acc = Op::apply(acc, q.y);
        acc = Op::apply(acc, q.z);
        acc = Op::apply(acc, q.w);
    }
    for (uint64_t i = n4 * 4 + (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kThreads)
        acc = Op::apply(acc, in[i]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc = Op::apply(acc, shfl_xor(acc, m));
    if (lane_id() == 0) s[threadIdx.x / VREN_WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = s[0];
        for (int w = 1; w < kThreads / VREN_WAVE; w++) r = Op::apply(r, s[w]);
        partials[blockIdx.x] = r;
    }
}

template <typename Op>
__global__ __launch_bounds__(kThreads) void total_final_kernel(const uint32_t* __restrict__ partials, uint32_t count,
                                                               uint32_t* __restrict__ out)
{
    __shared__ uint32_t s[kThreads / VREN_WAVE];
    uint32_t acc = Op::identity();
    for (uint32_t i = threadIdx.x; i < count; i += kThreads) acc = Op::apply(acc, partials[i]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc = Op::apply(acc, shfl_xor(acc, m));
    if (lane_id() == 0) s[threadIdx.x / VREN_WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = s[0];
        for (int w = 1; w < kThreads / VREN_WAVE; w++) r = Op::apply(r, s[w]);
        *out = r;
    }
}

template <typename Op>
int reduce_total(vren_ctx* ctx, hipStream_t s, const uint32_t* in, uint64_t n, uint32_t* out)
{
    uint32_t blocks = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * 8, std::max<uint64_t>(1, ceil_div_u64(n, 4096)));
    int st = vren_ensure_scratch(ctx, blocks * sizeof(uint32_t));
    if (st) return st;
    uint32_t* partials = static_cast<uint32_t*>(ctx->scratch);
    total_partial_kernel<Op><<<blocks, kThreads, 0, s>>>(in, n, partials);
    VREN_LAUNCH_CHECK(ctx);
    total_final_kernel<Op><<<1, kThreads, 0, s>>>(partials, blocks, out);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

} // namespace

extern "C" {

VREN_EXPORT int vren_reduce_u32(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in, uint32_t n,
                                uint32_t* out_tree, int op, uint32_t blocks)
{
    if (!ctx) return VREN_EINVAL;
    if (!in || !out_tree || n == 0 || blocks == 0 || blocks > 65535 || n > (1u << 31))
        return vren_fail(ctx, VREN_EINVAL, "reduce_u32: null buffer, n == 0, n > 2^31 or blocks not in [1, 65535]");
    if (in == out_tree && blocks > 1 && next_pow2_u32(n) != n)
        return vren_fail(ctx, VREN_EINVAL, "reduce_u32: in-place multi-block reduce needs a power-of-2 length");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "reduce_u32: context has no device");
    hipStream_t s = (hipStream_t)stream;
    switch (op) {
    case VREN_REDUCE_ADD: return reduce_tree<uint32_t, OpAddU, 16>(ctx, s, in, n, out_tree, blocks);
    case VREN_REDUCE_MIN: return reduce_tree<uint32_t, OpMinU, 16>(ctx, s, in, n, out_tree, blocks);
    case VREN_REDUCE_MAX: return reduce_tree<uint32_t, OpMaxU, 16>(ctx, s, in, n, out_tree, blocks);
    default: return vren_fail(ctx, VREN_EINVAL, "reduce_u32: unknown op %d", op);
    }
}

VREN_EXPORT int vren_reduce_vec4(vren_ctx* ctx, vren_stream_t stream, const float* in, uint32_t n, float* out_tree,
                                 int op, uint32_t blocks)
{
    if (!ctx) return VREN_EINVAL;
    if (!in || !out_tree || n == 0 || blocks == 0 || blocks > 65535 || n > (1u << 31))
        return vren_fail(ctx, VREN_EINVAL, "reduce_vec4: null buffer, n == 0, n > 2^31 or blocks not in [1, 65535]");
    if (in == out_tree && blocks > 1 && next_pow2_u32(n) != n)
        return vren_fail(ctx, VREN_EINVAL, "reduce_vec4: in-place multi-block reduce needs a power-of-2 length");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "reduce_vec4: context has no device");
    hipStream_t s = (hipStream_t)stream;
    const float4* vin = reinterpret_cast<const float4*>(in);
    float4* vout = reinterpret_cast<float4*>(out_tree);
    switch (op) {
    case VREN_REDUCE_ADD: return reduce_tree<float4, OpAddV, 4>(ctx, s, vin, n, vout, blocks);
    case VREN_REDUCE_MIN: return reduce_tree<float4, OpMinV, 4>(ctx, s, vin, n, vout, blocks);
    case VREN_REDUCE_MAX: return reduce_tree<float4, OpMaxV, 4>(ctx, s, vin, n, vout, blocks);
    default: return vren_fail(ctx, VREN_EINVAL, "reduce_vec4: unknown op %d", op);
    }
}

VREN_EXPORT int vren_reduce_total_u32(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in, uint64_t n,
                                      uint32_t* out, int op)
{
    if (!ctx) return VREN_EINVAL;
    if (!in || !out || n == 0) return vren_fail(ctx, VREN_EINVAL, "reduce_total_u32: null buffer or n == 0");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "reduce_total_u32: context has no device");
    hipStream_t s = (hipStream_t)stream;
    switch (op) {
    case VREN_REDUCE_ADD: return reduce_total<OpAddU>(ctx, s, in, n, out);
    case VREN_REDUCE_MIN: return reduce_total<OpMinU>(ctx, s, in, n, out);
    case VREN_REDUCE_MAX: return reduce_total<OpMaxU>(ctx, s, in, n, out);
    default: return vren_fail(ctx, VREN_EINVAL, "reduce_total_u32: unknown op %d", op);
    }
}

} // extern "C"
