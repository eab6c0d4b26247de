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
//      over the tiles' last elements, into a prefix array (all its reads issued before the sum, no dependent chain);
//   2. one workgroup per tile holds it in registers as rows of 1024 elements, 4 consecutive elements per thread (one
//      16-byte access per row, consecutive threads on consecutive addresses); a thread's 4-element group prefix is
//      the closed form over the tile's group-last elements (staged in LDS), its 4 outputs the closed form over its
//      own registers.
// Root prefix P: 0 when clear_last; otherwise the block's last element, except for blocks shorter than 1024, where
// the reference's 1024-wide LDS pass pads the block with zeros and the padding's prefix, 0, reaches the block's
// root (so such a block is always scanned from 0).
#include "common.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kGroup = 4;                  // consecutive elements per thread and row
constexpr int kRowLen = kThreads * kGroup; // 1024
constexpr int kRows = 4;
constexpr int kTile = kRowLen * kRows;     // 4096 elements per workgroup
constexpr int kMaxLogTiles = 32;

typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));

// closed form at the tile level: prefix of tile k of each block. Every term is loaded (a term whose bit is clear
// loads the block's last element and is dropped), so the loads are independent and issue together.
__global__ void downsweep_tile_prefix_kernel(const uint32_t* __restrict__ data, uint32_t length, uint32_t tiles,
                                             uint32_t log_tiles, uint64_t n_total, int root_cleared,
                                             uint32_t* __restrict__ prefix)
{
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const uint64_t seg = g / tiles;
    const uint32_t k = (uint32_t)(g - seg * tiles);
    const uint32_t* base = data + seg * length;
    uint32_t t[kMaxLogTiles];
#pragma unroll
    for (uint32_t l = 0; l < kMaxLogTiles; l++) {
        const bool on = l < log_tiles && ((k >> l) & 1u);
        // left sibling's last tile at level l
        const uint64_t j = ((uint64_t)(k >> (l + 1)) << (l + 1)) + (1ull << l) - 1;
        t[l] = base[on ? j * kTile + kTile - 1 : (uint64_t)length - 1];
        if (!on) t[l] = 0;
        if (l + 1 >= log_tiles) break;
    }
    uint32_t p = root_cleared ? 0u : base[length - 1];
#pragma unroll
    for (uint32_t l = 0; l < kMaxLogTiles; l++) {
        p += t[l];
        if (l + 1 >= log_tiles) break;
    }
    prefix[g] = p;
}

__global__ __launch_bounds__(kThreads) void downsweep_tile_kernel(uint32_t* __restrict__ data, uint32_t length,
                                                                  uint32_t tile_len, uint32_t tiles,
                                                                  const uint32_t* __restrict__ prefix,
                                                                  int root_cleared)
{
    __shared__ uint32_t s_last[kTile / kGroup];
    const uint32_t tid = threadIdx.x;
    const uint64_t g = blockIdx.x;
    const uint64_t seg = g / tiles;
    const uint32_t t = (uint32_t)(g - seg * tiles);
    uint32_t* tile = data + seg * length + (uint64_t)t * tile_len;
    const uint32_t root = prefix ? prefix[g] : (root_cleared ? 0u : tile[tile_len - 1]);

    if (tile_len < kGroup) { // blocks of 1 or 2 elements: one element per thread
        uint32_t v = 0;
        if (tid < tile_len) {
            v = root;
            if (tid == 1) v += tile[0];
        }
        __syncthreads(); // every read of the block precedes every write
        if (tid < tile_len) tile[tid] = v;
        return;
    }

    // rows of row_len elements, row_len / 4 threads per row
    const uint32_t row_len = tile_len < (uint32_t)kRowLen ? tile_len : (uint32_t)kRowLen;
    const uint32_t rows = tile_len / row_len, row_groups = row_len / kGroup;
    const bool active = tid < row_groups;
    uint32_t e[kRows][kGroup];
#pragma unroll
    for (uint32_t r = 0; r < kRows; r++) {
        if (r < rows && active) {
            const u32x4_a4 v = __builtin_nontemporal_load((const u32x4_a4*)(tile + r * row_len + kGroup * tid));
            e[r][0] = v.x;
            e[r][1] = v.y;
            e[r][2] = v.z;
            e[r][3] = v.w;
            s_last[r * row_groups + tid] = v.w;
        }
    }
    __syncthreads();
    if (!active) return;
    const uint32_t groups = rows * row_groups;
#pragma unroll
    for (uint32_t r = 0; r < kRows; r++) {
        if (r >= rows) break;
        // the group's prefix: closed form over the group-last elements (the tile's root is replaced by its prefix
        // and never read as a left sibling)
        const uint32_t G = r * row_groups + tid;
        uint32_t p = root;
        for (uint32_t l = 0; (1u << l) < groups; l++)
            if ((G >> l) & 1u) p += s_last[((G >> (l + 1)) << (l + 1)) + (1u << l) - 1];
        u32x4_a4 o;
        o.x = p;
        o.y = p + e[r][0];
        o.z = p + e[r][1];
        o.w = p + e[r][1] + e[r][2];
        __builtin_nontemporal_store(o, (u32x4_a4*)(tile + r * row_len + kGroup * tid));
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
