// downsweep.hip -- blelloch_scan::downsweep (blelloch_scan.hpp:31-39, blelloch_scan.cpp:57-139,
// blelloch_scan_downsweep.comp:33-125): the second half of the reference's two-phase scan, public because its radix
// sort calls it on up-swept digit counts (radix_sort.cpp:281-289).
//
// The reference walks the tree level by level: one global dispatch per level above 10, then a 1024-thread LDS pass.
// Level l visits the pairs (a, b) = (k 2^(l+1) + 2^l - 1, a + 2^l) and sets d[a] <- d[b], d[b] <- d[a] + d[b], with
// the block's last element first cleared when clear_last. Position a is the last element of a 2^l-element node and
// of no larger one, so no level above l writes it: every output is a closed form of the input,
//     out[i] = P + sum over the set bits l of i of d[((i >> (l + 1)) << (l + 1)) + 2^l - 1],
// P being the root's prefix. That is what is computed here, with each element read once and written once:
//   1. (blocks longer than one tile) one thread per 4096-element tile evaluates the closed form at the tile level,
//      over the tiles' last elements, into a prefix array;
//   2. one workgroup per tile holds 16 consecutive elements per thread in registers; the thread's chunk prefix is the
//      closed form over the tile's chunk-last elements (staged in LDS, at most 8 reads), its 16 outputs the closed
//      form over its own registers.
// Root prefix P: 0 when clear_last; otherwise the block's last element, except for blocks shorter than 1024, where
// the reference's 1024-wide LDS pass pads the block with zeros and the padding's prefix, 0, reaches the block's
// root (so such a block is always scanned from 0).
#include "common.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 16;                  // elements per thread
constexpr int kTile = kThreads * kChunk;    // 4096 elements per workgroup

// closed form at the tile level: prefix of tile k of each block
__global__ void downsweep_tile_prefix_kernel(const uint32_t* __restrict__ data, uint32_t length, uint32_t tiles,
                                             uint32_t log_tiles, uint64_t n_total, int root_cleared,
                                             uint32_t* __restrict__ prefix)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const uint64_t seg = g / tiles;
    const uint32_t k = (uint32_t)(g - seg * tiles);
    const uint32_t* base = data + seg * length;
    uint32_t p = root_cleared ? 0u : base[length - 1];
    for (uint32_t l = 0; l < log_tiles; l++) {
        if ((k >> l) & 1u) {
            const uint64_t j = ((uint64_t)(k >> (l + 1)) << (l + 1)) + (1ull << l) - 1; // left sibling's last tile
            p += base[j * kTile + kTile - 1];
        }
    }
    prefix[g] = p;
}

// out[j] of a 16-element chunk from its prefix p and its input e (the closed form over 4 levels, unrolled)
__device__ __forceinline__ void chunk_downsweep(const uint32_t (&e)[kChunk], uint32_t p, uint32_t (&o)[kChunk])
{
#pragma unroll
    for (int j = 0; j < kChunk; j++) {
        uint32_t v = p;
#pragma unroll
        for (int l = 0; l < 4; l++)
            if ((j >> l) & 1) v += e[((j >> (l + 1)) << (l + 1)) + (1 << l) - 1];
        o[j] = v;
    }
}

__global__ __launch_bounds__(kThreads) void downsweep_tile_kernel(uint32_t* __restrict__ data, uint32_t length,
                                                                  uint32_t tile_len, uint32_t tiles,
                                                                  const uint32_t* __restrict__ prefix,
                                                                  int root_cleared)
{
    __shared__ uint32_t s_last[kThreads];
    const uint32_t tid = threadIdx.x;
    const uint64_t g = blockIdx.x;
    const uint64_t seg = g / tiles;
    const uint32_t t = (uint32_t)(g - seg * tiles);
    uint32_t* tile = data + seg * length + (uint64_t)t * tile_len;
    const uint32_t root = prefix ? prefix[g] : (root_cleared ? 0u : tile[tile_len - 1]);

    if (tile_len < kChunk) { // blocks of 1, 2, 4 or 8 elements: one element per thread
        uint32_t v = 0;
        if (tid < tile_len) {
            v = root;
            for (uint32_t l = 0; (1u << l) < tile_len; l++)
                if ((tid >> l) & 1u) v += tile[((tid >> (l + 1)) << (l + 1)) + (1u << l) - 1];
        }
        __syncthreads(); // every read of the block precedes every write
        if (tid < tile_len) tile[tid] = v;
        return;
    }

    const uint32_t chunks = tile_len / kChunk;
    uint32_t e[kChunk];
    if (tid < chunks) {
        uint32_t* c = tile + tid * kChunk;
        if ((reinterpret_cast<uintptr_t>(c) & 15) == 0) {
#pragma unroll
            for (int q = 0; q < kChunk / 4; q++) {
                uint4 v = ntload4(c + 4 * q);
                e[4 * q] = v.x;
                e[4 * q + 1] = v.y;
                e[4 * q + 2] = v.z;
                e[4 * q + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kChunk; j++) e[j] = c[j];
        }
        s_last[tid] = e[kChunk - 1];
    }
    __syncthreads();
    if (tid >= chunks) return;
    // the chunk's prefix: closed form over the chunk-last elements (the tile's root is replaced by its prefix and
    // never read as a left sibling)
    uint32_t p = root;
    for (uint32_t l = 0; (1u << l) < chunks; l++)
        if ((tid >> l) & 1u) p += s_last[((tid >> (l + 1)) << (l + 1)) + (1u << l) - 1];
    uint32_t o[kChunk];
    chunk_downsweep(e, p, o);
    uint32_t* c = tile + tid * kChunk;
    if ((reinterpret_cast<uintptr_t>(c) & 15) == 0) {
#pragma unroll
        for (int q = 0; q < kChunk / 4; q++)
            ntstore4(c + 4 * q, make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]));
    } else {
#pragma unroll
        for (int j = 0; j < kChunk; j++) c[j] = o[j];
    }
}

} // namespace

extern "C" VREN_EXPORT int vren_blelloch_downsweep_u32(vren_ctx* ctx, vren_stream_t stream, uint32_t* data,
                                                       uint32_t length, uint32_t blocks, int clear_last)
{
    if (!ctx) return VREN_EINVAL;
    if (!data || length == 0 || (length & (length - 1)) != 0 || blocks == 0)
        return vren_fail(ctx, VREN_EINVAL,
                         "blelloch_downsweep_u32: null buffer, length not a power of 2 (blelloch_scan.cpp:68) or "
                         "blocks == 0");
    const uint32_t tile_len = length < (uint32_t)kTile ? length : (uint32_t)kTile;
    const uint32_t tiles = length / tile_len;
    const uint64_t n_tiles = (uint64_t)tiles * blocks;
    if (n_tiles >= (1ull << 31)) return vren_fail(ctx, VREN_EINVAL, "blelloch_downsweep_u32: input too large");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "blelloch_downsweep_u32: context has no device");
    const int root_cleared = clear_last || length < 1024;
    hipStream_t s = (hipStream_t)stream;
    uint32_t* prefix = nullptr;
    if (tiles > 1) {
        int st = vren_ensure_scratch(ctx, n_tiles * sizeof(uint32_t), s);
        if (st) return st;
        prefix = static_cast<uint32_t*>(ctx->scratch);
        const uint32_t log_tiles = ilog2_u64(tiles);
        downsweep_tile_prefix_kernel<<<(uint32_t)ceil_div_u64(n_tiles, 256), 256, 0, s>>>(
            data, length, tiles, log_tiles, n_tiles, root_cleared, prefix);
        VREN_LAUNCH_CHECK(ctx);
    }
    downsweep_tile_kernel<<<(uint32_t)n_tiles, kThreads, 0, s>>>(data, length, tile_len, tiles, prefix,
                                                                 root_cleared);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}
