// scan.hip -- blelloch_scan (blelloch_scan.hpp:9-49, blelloch_scan.cpp:57-166) as a single-pass exclusive scan.
//
// The reference runs an in-place up-sweep (reduce) and then one down-sweep dispatch per level above 10 plus an
// LDS down-sweep, i.e. ~2 log2(N) / 10 + log2(N) - 10 strided passes over the buffer. Here each element is read
// once and written once: tiles of 4096 uint32 are taken from a ticket counter, scanned in registers/LDS, and
// chained with a decoupled look-back over 64-bit (state | value) status words published at agent scope, so the
// chain works across the 8 XCDs' separate L2s. Result: exclusive prefix sums mod 2^32, segment by segment.
#include "scan_device.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;

__global__ void scan_init_kernel(uint64_t* status, uint64_t n_status, uint32_t* counter)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n_status; i += (uint64_t)gridDim.x * blockDim.x) status[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

__global__ __launch_bounds__(kThreads) void scan_kernel(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t tiles_per_seg,
                                                        uint64_t* status, uint32_t* counter)
{
    __shared__ uint32_t s_stage[kTile];
    __shared__ uint32_t s_wave[kThreads / VREN_WAVE];
    __shared__ uint32_t s_ticket;
    __shared__ uint32_t s_prefix;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_ticket = atomic_add_agent(counter, 1u);
    __syncthreads();
    const uint32_t g = s_ticket;
    const uint32_t seg = g / tiles_per_seg;
    const uint32_t t = g - seg * tiles_per_seg;
    const uint64_t tile_start = (uint64_t)t * kTile;
    const uint64_t seg_base = (uint64_t)seg * n;
    const uint32_t cnt = (uint32_t)min<uint64_t>((uint64_t)kTile, n - tile_start);
    const uint32_t* src = in + seg_base + tile_start;
    uint32_t* dst = out + seg_base + tile_start;
    const bool vec = cnt == kTile && (reinterpret_cast<uintptr_t>(src) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(dst) & 15) == 0;

    // striped, coalesced load into LDS
    if (vec) {
#pragma unroll
        for (int j = 0; j < kItems / 4; j++) {
            uint32_t q = j * kThreads + tid;
            uint4 v = ntload4(reinterpret_cast<const uint4*>(src) + q);
            s_stage[4 * q + 0] = v.x;
            s_stage[4 * q + 1] = v.y;
            s_stage[4 * q + 2] = v.z;
            s_stage[4 * q + 3] = v.w;
        }
    } else {
        for (uint32_t i = tid; i < (uint32_t)kTile; i += kThreads) s_stage[i] = i < cnt ? src[i] : 0u;
    }
    __syncthreads();
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        v[k] = s_stage[tid * kItems + k];
        sum += v[k];
    }
    uint32_t block_total;
    uint32_t excl = block_exclusive_sum<kThreads>(sum, s_wave, &block_total);
    if (tid < VREN_WAVE) {
        uint32_t prefix = wave_lookback(status + (uint64_t)seg * tiles_per_seg, t, block_total);
        if (tid == 0) s_prefix = prefix;
    }
    __syncthreads();
    uint32_t run = s_prefix + excl;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        s_stage[tid * kItems + k] = run;
        run += v[k];
    }
    __syncthreads();
    if (vec) {
#pragma unroll
        for (int j = 0; j < kItems / 4; j++) {
            uint32_t q = j * kThreads + tid;
            uint4 w = make_uint4(s_stage[4 * q + 0], s_stage[4 * q + 1], s_stage[4 * q + 2], s_stage[4 * q + 3]);
            ntstore4(reinterpret_cast<uint4*>(dst) + q, w);
        }
    } else {
        for (uint32_t i = tid; i < cnt; i += kThreads) dst[i] = s_stage[i];
    }
}

} // namespace

size_t vren_scan_work_bytes(uint64_t n, uint32_t blocks)
{
    return 256 + ceil_div_u64(n, kTile) * blocks * sizeof(uint64_t);
}

// exclusive scan with caller-provided look-back workspace (vren_scan_work_bytes)
int vren_scan_with_work(vren_ctx* ctx, hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t blocks,
                        void* work)
{
    uint64_t tps = ceil_div_u64(n, kTile);
    uint64_t tiles = tps * blocks;
    uint32_t* counter = static_cast<uint32_t*>(work);
    uint64_t* status = reinterpret_cast<uint64_t*>(static_cast<char*>(work) + 256);
    uint32_t init_blocks = (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(1, ceil_div_u64(tiles, 256)));
    scan_init_kernel<<<init_blocks, 256, 0, s>>>(status, tiles, counter);
    VREN_LAUNCH_CHECK(ctx);
    scan_kernel<<<(uint32_t)tiles, kThreads, 0, s>>>(in, out, n, (uint32_t)tps, status, counter);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}

extern "C" VREN_EXPORT int vren_exclusive_scan_u32(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in,
                                                   uint32_t* out, uint64_t n, uint32_t blocks)
{
    if (!ctx) return VREN_EINVAL;
    if (!in || !out || n == 0 || blocks == 0)
        return vren_fail(ctx, VREN_EINVAL, "exclusive_scan_u32: null buffer, n == 0 or blocks == 0");
    if (ceil_div_u64(n, kTile) * blocks >= (1ull << 31))
        return vren_fail(ctx, VREN_EINVAL, "exclusive_scan_u32: input too large");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "exclusive_scan_u32: context has no device");
    int st = vren_ensure_scratch(ctx, vren_scan_work_bytes(n, blocks));
    if (st) return st;
    return vren_scan_with_work(ctx, (hipStream_t)stream, in, out, n, blocks, ctx->scratch);
}
