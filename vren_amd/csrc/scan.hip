// scan.hip -- blelloch_scan (blelloch_scan.hpp:9-49, blelloch_scan.cpp:57-166) as a single-pass exclusive scan.
//
// The reference runs an in-place up-sweep (reduce) and then one down-sweep dispatch per level above 10 plus an
// LDS down-sweep, i.e. ~2 log2(N) / 10 + log2(N) - 10 strided passes over the buffer. Here each element is read
// once and written once: tiles of 16384 uint32 are taken from a ticket counter, scanned in registers, and
// chained with a decoupled look-back over 64-bit (state | value) status words published at agent scope, so the
// chain works across the 8 XCDs' separate L2s. Result: exclusive prefix sums mod 2^32, segment by segment.
#include "scan_device.hpp"

namespace {

// A tile is kRows rows of kThreads x 4 elements; thread `tid` holds 4 consecutive elements of every row, loaded
// and stored as one 16-byte access per row (coalesced across the wave). The tile's 64 (row, wave) segments of 256
// elements are scanned by one wave from their totals, so the data never leaves registers. Tiles of 16384 keys keep
// the look-back chain short (4096 tiles at 2^26): the inclusive prefix moves back one 64-tile window per
// agent-scope round trip, which for 4-KiB tiles was the bound.
constexpr int kThreads = 512;
constexpr int kWaves = kThreads / VREN_WAVE;
constexpr int kRows = 8;
constexpr int kTile = kThreads * 4 * kRows;
static_assert(kRows * kWaves == VREN_WAVE, "one segment per lane of the scanning wave");

__global__ void scan_init_kernel(uint64_t* status, uint64_t n_status, uint32_t* counter)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n_status; i += (uint64_t)gridDim.x * blockDim.x) status[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

__global__ __launch_bounds__(kThreads) void scan_kernel(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t tiles_per_seg,
                                                        uint64_t* status, uint32_t* counter)
{
    __shared__ uint32_t s_part[kRows * kWaves]; // (row, wave) segment totals, then their exclusive prefixes
    __shared__ uint32_t s_wsum[kWaves];          // each wave's total over its rows
    __shared__ uint32_t s_ticket;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = wave_id();
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

    uint4 v[kRows];
    if (vec) {
#pragma unroll
        for (int j = 0; j < kRows; j++) v[j] = ntload4(reinterpret_cast<const uint4*>(src) + j * kThreads + tid);
    } else {
#pragma unroll
        for (int j = 0; j < kRows; j++) {
            const uint32_t e = (j * kThreads + tid) * 4;
            v[j].x = e + 0 < cnt ? src[e + 0] : 0u;
            v[j].y = e + 1 < cnt ? src[e + 1] : 0u;
            v[j].z = e + 2 < cnt ? src[e + 2] : 0u;
            v[j].w = e + 3 < cnt ? src[e + 3] : 0u;
        }
    }
    // the tile's total first, published before the row scans, so that the next tiles' walks find this tile's
    // aggregate while it still scans its rows
    {
        uint32_t tsum = 0;
#pragma unroll
        for (int j = 0; j < kRows; j++) tsum += (v[j].x + v[j].y) + (v[j].z + v[j].w);
        tsum = wave_sum(tsum);
        if (lane == 0) s_wsum[wave] = tsum;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t agg = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) agg += s_wsum[w];
        wave_lookback_publish(status + (uint64_t)seg * tiles_per_seg, t, agg);
    }
    uint32_t excl[kRows]; // exclusive prefix of the thread's 4 elements within its wave's segment of each row
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        const uint32_t sum = (v[j].x + v[j].y) + (v[j].z + v[j].w);
        const uint32_t incl = wave_inclusive_sum(sum);
        excl[j] = incl - sum;
        if (lane == VREN_WAVE - 1) s_part[j * kWaves + wave] = incl;
    }
    __syncthreads();
    if (wave == 0) {
        const uint32_t x = s_part[lane];
        const uint32_t xi = wave_inclusive_sum(x);
        const uint32_t agg = __builtin_amdgcn_readlane(xi, VREN_WAVE - 1);
        const uint32_t prefix = wave_lookback_walk(status + (uint64_t)seg * tiles_per_seg, t, agg);
        s_part[lane] = prefix + xi - x;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        uint32_t run = s_part[j * kWaves + wave] + excl[j];
        uint4 o;
        o.x = run;
        run += v[j].x;
        o.y = run;
        run += v[j].y;
        o.z = run;
        run += v[j].z;
        o.w = run;
        if (vec) {
            ntstore4(reinterpret_cast<uint4*>(dst) + j * kThreads + tid, o);
        } else {
            const uint32_t e = (j * kThreads + tid) * 4;
            if (e + 0 < cnt) dst[e + 0] = o.x;
            if (e + 1 < cnt) dst[e + 1] = o.y;
            if (e + 2 < cnt) dst[e + 2] = o.z;
            if (e + 3 < cnt) dst[e + 3] = o.w;
        }
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
    int st = vren_ensure_scratch(ctx, vren_scan_work_bytes(n, blocks), (hipStream_t)stream);
    if (st) return st;
    return vren_scan_with_work(ctx, (hipStream_t)stream, in, out, n, blocks, ctx->scratch);
}

// ---- allocation-free variant -------------------------------------------------------------------------------------
// Callers that cannot allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc) get the same scan over a fixed workspace: a file-scope device array, of which every GPU holds its own
// copy at its own address until the process ends. Its first 256 bytes are the tile ticket counter, the rest one
// 64-bit look-back status word per tile, for at most kNoAllocTiles tiles over all segments (2^30 keys). The host
// never passes the array's name to a kernel (on the host it is only a stand-in for the symbol): vren_create, with
// its device current, resolves this device's address once into ctx->noalloc_scan_work, and the entry point passes
// that. Every call on one GPU shares the array; scan_init_kernel clears the counter and the status words at the
// start of each call, so calls in sequence on one stream are independent, and calls that overlap (other streams,
// other contexts on the same GPU) race.
constexpr uint64_t kNoAllocTiles = 1ull << 16;
constexpr size_t kNoAllocBytes = 256 + kNoAllocTiles * sizeof(uint64_t);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

int vren_noalloc_scan_setup(vren_ctx* ctx)
{
    void* p = nullptr;
    VREN_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ctx->noalloc_scan_work = p;
    return VREN_OK;
}

extern "C" VREN_EXPORT int vren_exclusive_scan_u32_noalloc(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in,
                                                           uint32_t* out, uint64_t n, uint32_t blocks)
{
    if (!ctx) return VREN_EINVAL;
    if (!in || !out || n == 0 || blocks == 0)
        return vren_fail(ctx, VREN_EINVAL, "exclusive_scan_u32_noalloc: null buffer, n == 0 or blocks == 0");
    if (n > kNoAllocTiles * kTile || ceil_div_u64(n, kTile) * blocks > kNoAllocTiles)
        return vren_fail(ctx, VREN_EINVAL,
                         "exclusive_scan_u32_noalloc: %llu x %u keys need more than the fixed workspace's %llu tiles "
                         "of %d keys",
                         (unsigned long long)n, blocks, (unsigned long long)kNoAllocTiles, kTile);
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "exclusive_scan_u32_noalloc: context has no device");
    if (!ctx->noalloc_scan_work)
        return vren_fail(ctx, VREN_EHIP, "exclusive_scan_u32_noalloc: workspace address not resolved");
    return vren_scan_with_work(ctx, (hipStream_t)stream, in, out, n, blocks, ctx->noalloc_scan_work);
}
