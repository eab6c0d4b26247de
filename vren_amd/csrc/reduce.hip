// reduce.hip -- reduce<T, Op> (reduce.hpp:15-49, reduce.cpp:32-140, reduce.comp:50-87) for gfx950.
//
// The reference output is the whole in-place Blelloch up-sweep tree of the pow2-padded input:
//   tree[i] = OP over the aligned block of 2^k elements ending at i, k = number of trailing 1 bits of i,
// combined as a balanced binary tree (each node = OP(right half, left half), reduce.comp:76).
// Here one pass handles TILE consecutive elements of a (possibly strided) view per workgroup, held in registers
// with coalesced 16-byte loads and stores (the tree levels are formed in registers, across lanes and across the
// tile's wave runs), and the tile's elements are written back in place of the ones read. The tile totals (at stride TILE) are up-swept by the
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

constexpr int kThreads = 256; // scalar total

// Up-sweep tile: kRows rows of kTreeThreads x E elements, E = 16 / sizeof(T), so that every thread loads and stores
// E consecutive elements of each row with one 16-byte access (coalesced across the wave). The tree levels inside
// a thread's E elements are formed in registers, the next six across the lanes with shuffles (a wave covers an
// aligned run of 64 E elements of a row), and the last six across the tile's 64 (row, wave) runs by one wave.
constexpr int kTreeThreads = 512;
constexpr int kTreeWaves = kTreeThreads / VREN_WAVE;
constexpr int kRows = 8;
static_assert(kRows * kTreeWaves == VREN_WAVE, "one run per lane of the finishing wave");

template <typename T>
constexpr uint64_t tree_tile() { return (uint64_t)kTreeThreads * (16 / sizeof(T)) * kRows; }

// One tile of the view: elements e = tile * TILE + i, read from in[in_offset + e * in_stride] (identity at and
// beyond `valid`) and written to out[offset + e * stride] (e < count). With `totals`, the tile's total (the node
// at its last element) is also written to totals[segment * gridDim.x + tile], a compact array that the next pass
// reads instead of the totals' strided places in the tree.
template <typename T, typename Op>
__global__ __launch_bounds__(kTreeThreads) void upsweep_tile_kernel(const T* in, T* out,
                                                                    uint64_t in_seg_stride, uint64_t out_seg_stride,
                                                                    uint64_t valid, uint64_t count,
                                                                    uint64_t in_offset, uint64_t in_stride,
                                                                    uint64_t offset, uint64_t stride,
                                                                    T* __restrict__ totals)
{
    constexpr int E = 16 / sizeof(T);
    constexpr uint64_t TILE = tree_tile<T>();
    static_assert(E * sizeof(T) == 16, "16-byte groups");
    __shared__ T s_run[kRows * kTreeWaves];
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    const uint64_t seg = blockIdx.y;
    const T* src = in + seg * in_seg_stride;
    T* dst = out + seg * out_seg_stride;
    const uint64_t tile_start = (uint64_t)blockIdx.x * TILE;
    const bool vin = in_stride == 1 && in_offset == 0 && tile_start + TILE <= valid &&
                     (reinterpret_cast<uintptr_t>(src + tile_start) & 15) == 0;
    const bool vout = stride == 1 && offset == 0 && tile_start + TILE <= count &&
                      (reinterpret_cast<uintptr_t>(dst + tile_start) & 15) == 0;

    T v[kRows][E];
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        const uint64_t first = tile_start + ((uint64_t)j * kTreeThreads + tid) * E;
        if (vin) {
            const uint4 q = *reinterpret_cast<const uint4*>(src + first);
            memcpy(&v[j][0], &q, 16);
        } else {
#pragma unroll
            for (int c = 0; c < E; c++) {
                const uint64_t e = first + c;
                v[j][c] = e < valid ? src[in_offset + e * in_stride] : Op::identity();
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        // levels inside the thread's E elements
#pragma unroll
        for (int l = 1; l < E; l <<= 1) {
#pragma unroll
            for (int c = 0; c < E; c++)
                if ((c & (2 * l - 1)) == 2 * l - 1) v[j][c] = Op::apply(v[j][c], v[j][c - l]);
        }
        // levels across the lanes: the node of lane i is its last element
        T x = v[j][E - 1];
#pragma unroll
        for (unsigned l = 1; l < VREN_WAVE; l <<= 1) {
            T o = shfl_up_t(x, l);
            if ((lane & (2 * l - 1)) == 2 * l - 1) x = Op::apply(x, o);
        }
        v[j][E - 1] = x;
        if (lane == VREN_WAVE - 1) s_run[j * kTreeWaves + wave] = x;
    }
    __syncthreads();
    if (wave == 0) {
        // levels across the 64 runs, in element order (row-major over (row, wave))
        T x = s_run[lane];
#pragma unroll
        for (unsigned l = 1; l < VREN_WAVE; l <<= 1) {
            T o = shfl_up_t(x, l);
            if ((lane & (2 * l - 1)) == 2 * l - 1) x = Op::apply(x, o);
        }
        s_run[lane] = x;
    }
    __syncthreads();
    if (totals && tid == 0) totals[seg * gridDim.x + blockIdx.x] = s_run[kRows * kTreeWaves - 1];
    if (lane == VREN_WAVE - 1) {
#pragma unroll
        for (int j = 0; j < kRows; j++) v[j][E - 1] = s_run[j * kTreeWaves + wave];
    }
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        const uint64_t first = tile_start + ((uint64_t)j * kTreeThreads + tid) * E;
        if (vout) {
            uint4 q;
            memcpy(&q, &v[j][0], 16);
            *reinterpret_cast<uint4*>(dst + first) = q;
        } else {
#pragma unroll
            for (int c = 0; c < E; c++) {
                const uint64_t e = first + c;
                if (e < count) dst[offset + e * stride] = v[j][c];
            }
        }
    }
}

template <typename T, typename Op>
int reduce_tree(vren_ctx* ctx, hipStream_t s, const T* in, uint32_t n, T* out, uint32_t blocks)
{
    constexpr uint64_t TILE = tree_tile<T>();
    const uint64_t p = next_pow2_u32(n);
    // compact tile totals of each pass but the last (p / TILE, then p / TILE^2, ... per segment), double-buffered
    const uint64_t t1 = p > TILE ? p / TILE : 0, t2 = t1 > TILE ? t1 / TILE : 0;
    const size_t a_bytes = (size_t)((t1 * blocks * sizeof(T) + 255) & ~(uint64_t)255);
    if (t1) {
        int st = vren_ensure_scratch(ctx, a_bytes + t2 * blocks * sizeof(T), s);
        if (st) return st;
    }
    T* buf[2] = {static_cast<T*>(ctx->scratch), reinterpret_cast<T*>(static_cast<char*>(ctx->scratch) + a_bytes)};
    // pass 0: in (n valid, segment stride n) -> out (p elements, segment stride p); pass k > 0: the compact
    // totals of pass k - 1 -> their places in the tree (the last element of each tile of the previous view)
    uint64_t count = p, offset = 0, stride = 1;
    for (int pass = 0;; pass++) {
        const uint64_t tiles = ceil_div_u64(count, TILE);
        const dim3 grid((uint32_t)tiles, blocks);
        T* totals = count > TILE ? buf[pass & 1] : nullptr;
        if (pass == 0) {
            upsweep_tile_kernel<T, Op><<<grid, kTreeThreads, 0, s>>>(in, out, n, p, n, count, 0, 1, 0, 1, totals);
        } else {
            upsweep_tile_kernel<T, Op><<<grid, kTreeThreads, 0, s>>>(buf[(pass - 1) & 1], out, count, p, count, count,
                                                                     0, 1, offset, stride, totals);
        }
        VREN_LAUNCH_CHECK(ctx);
        if (count <= TILE) break;
        // next view: the tile totals
        offset = offset + (TILE - 1) * stride;
        stride = stride * TILE;
        count = count / TILE;
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
    int st = vren_ensure_scratch(ctx, blocks * sizeof(uint32_t), s);
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
    case VREN_REDUCE_ADD: return reduce_tree<uint32_t, OpAddU>(ctx, s, in, n, out_tree, blocks);
    case VREN_REDUCE_MIN: return reduce_tree<uint32_t, OpMinU>(ctx, s, in, n, out_tree, blocks);
    case VREN_REDUCE_MAX: return reduce_tree<uint32_t, OpMaxU>(ctx, s, in, n, out_tree, blocks);
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
    case VREN_REDUCE_ADD: return reduce_tree<float4, OpAddV>(ctx, s, vin, n, vout, blocks);
    case VREN_REDUCE_MIN: return reduce_tree<float4, OpMinV>(ctx, s, vin, n, vout, blocks);
    case VREN_REDUCE_MAX: return reduce_tree<float4, OpMaxV>(ctx, s, vin, n, vout, blocks);
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
