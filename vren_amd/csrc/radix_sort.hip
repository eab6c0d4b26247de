// radix_sort.hip -- radix_sort (radix_sort.hpp:9-53, radix_sort.cpp:149-338) as a onesweep LSD sort for gfx950.
//
// The reference runs, per 4-bit digit, six dispatches (fill, local count, 16-block reduce, global offset,
// 16-block down-sweep, reorder) over N/1024 workgroups, i.e. 8 passes x 3 reads/writes of the keys and ~48
// dispatches, and its reorder assumes a 32-wide subgroup (radix_sort_reorder.comp:15-17). Here:
//   1. one histogram kernel reads the keys once and counts every pass's digit (LDS atomics, then agent-scope
//      global atomics);
//   2. one kernel per pass: tiles taken from a ticket counter are ranked stably in LDS with wave64 ballot
//      matching, each tile's per-digit counts are chained across tiles by a per-digit decoupled look-back, and
//      the tile is scattered through an LDS staging buffer so global writes go out in per-digit runs.
// Per pass each key (and value) is read once and written once. Digit width, keys per thread and waves per
// workgroup are template parameters chosen by the autotuner (autotune.hip).
#include <algorithm>
#include <cstring>

#include "radix_device.hpp"
#include "radix_sort.hpp"
#include "scan_device.hpp"

namespace {

// predecessor states read per round trip of the onesweep look-back
constexpr int kLookback = 8;

// status word of (tile, digit): [31:30] state, [29:0] count
#define RS_AGG (1u << 30)
#define RS_INCL (2u << 30)
#define RS_MASK ((1u << 30) - 1u)

__global__ void rs_init_kernel(uint32_t* hist, uint32_t n_hist, uint32_t* counters, uint32_t n_counters)
{
    for (uint32_t i = threadIdx.x; i < n_hist; i += blockDim.x) hist[i] = 0u;
    for (uint32_t i = threadIdx.x; i < n_counters; i += blockDim.x) counters[i] = 0u;
}

// one workgroup of 1024 threads per CU: the partial histograms are merged into the global one with agent-scope
// atomics that all land in a few cache lines, so the number of workgroups is kept small
constexpr uint32_t kHistThreads = 1024;

// PAIRS: the keys are the first words of interleaved (key, value) pairs (key i at keys[2 i])
template <int BITS, bool PAIRS = false>
__global__ __launch_bounds__(kHistThreads) void rs_hist_kernel(const uint32_t* __restrict__ keys, uint64_t n,
                                                               uint32_t begin_bit, uint32_t end_bit, uint32_t passes,
                                                               uint32_t* __restrict__ hist,
                                                               uint32_t* __restrict__ clear, uint64_t clear_words)
{
    constexpr uint32_t KPV = PAIRS ? 2 : 4; // keys per 16-byte load
    constexpr int RADIX = 1 << BITS;
    __shared__ uint32_t s_h[8 * RADIX];
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += kHistThreads) s_h[i] = 0u;
    // clear the first pass's look-back statuses while we are here
    for (uint64_t i = (uint64_t)blockIdx.x * kHistThreads + threadIdx.x; i < clear_words;
         i += (uint64_t)gridDim.x * kHistThreads)
        clear[i] = 0u;
    __syncthreads();
    // the pass loop runs outside the keys held in registers, so its shift, width and bounds (wave-uniform) are
    // computed once per batch of keys rather than once per key
    auto count_keys = [&](const uint32_t* k, int m) {
        for (uint32_t p = 0; p < passes; p++) {
            const uint32_t sh = begin_bit + p * BITS;
            const uint32_t w = min((uint32_t)BITS, end_bit - sh);
            uint32_t* h = s_h + p * RADIX;
            for (int u = 0; u < m; u++) atomicAdd(&h[__builtin_amdgcn_ubfe(k[u], sh, w)], 1u);
        }
    };
    const bool aligned = (reinterpret_cast<uintptr_t>(keys) & 15) == 0;
    const uint64_t n4 = aligned ? n / KPV : 0;
    const uint4* k4 = reinterpret_cast<const uint4*>(keys);
    const uint64_t stride = (uint64_t)gridDim.x * kHistThreads;
    uint64_t i = (uint64_t)blockIdx.x * kHistThreads + threadIdx.x;
    // eight 128-bit loads in flight per thread before the LDS counting
    constexpr int U = 8;
    for (; i + (U - 1) * stride < n4; i += U * stride) {
        uint4 q[U];
#pragma unroll
        for (int u = 0; u < U; u++) q[u] = ntload4(k4 + i + u * stride);
        uint32_t k[U * KPV];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (PAIRS) {
                k[2 * u] = q[u].x;
                k[2 * u + 1] = q[u].z;
            } else {
                k[4 * u] = q[u].x;
                k[4 * u + 1] = q[u].y;
                k[4 * u + 2] = q[u].z;
                k[4 * u + 3] = q[u].w;
            }
        }
        for (uint32_t p = 0; p < passes; p++) {
            const uint32_t sh = begin_bit + p * BITS;
            const uint32_t w = min((uint32_t)BITS, end_bit - sh);
            uint32_t* h = s_h + p * RADIX;
#pragma unroll
            for (int u = 0; u < U * (int)KPV; u++) atomicAdd(&h[__builtin_amdgcn_ubfe(k[u], sh, w)], 1u);
        }
    }
    for (; i < n4; i += stride) {
        const uint4 q = k4[i];
        const uint32_t k[4] = {q.x, PAIRS ? q.z : q.y, q.z, q.w};
        count_keys(k, (int)KPV);
    }
    for (uint64_t j = n4 * KPV + (uint64_t)blockIdx.x * kHistThreads + threadIdx.x; j < n; j += stride) {
        const uint32_t k = keys[PAIRS ? 2 * j : j];
        count_keys(&k, 1);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < passes * RADIX; i += kHistThreads) {
        uint32_t c = s_h[i];
        if (c) atomic_add_agent(&hist[i], c);
    }
}

// exclusive scan of each pass's digit histogram in place: the digit's first output position. One workgroup.
template <int BITS>
__global__ __launch_bounds__(256) void rs_hexscan_kernel(uint32_t* hist, uint32_t passes)
{
    constexpr int RADIX = 1 << BITS;
    __shared__ uint32_t s_wave[4];
    for (uint32_t p = 0; p < passes; p++) {
        const uint32_t v = threadIdx.x < (uint32_t)RADIX ? hist[p * RADIX + threadIdx.x] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_sum<256>(v, s_wave, &total);
        if (threadIdx.x < (uint32_t)RADIX) hist[p * RADIX + threadIdx.x] = ex;
    }
}

// Reduce-then-scan form, pass part 1: the digit counts of every tile (tiles as in the scatter pass), written
// digit-major (count of digit d in tile t at d * tiles + t) so that one exclusive scan of the whole array gives
// every (digit, tile) its first output position. Per-wave LDS counters keep the atomics uncontended.
template <int BITS, int KPT, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void rs_upsweep_kernel(const uint32_t* __restrict__ keys, uint64_t n,
                                                                uint32_t shift, uint32_t width,
                                                                uint32_t* __restrict__ tile_hist, uint32_t tiles)
{
    constexpr int THREADS = WAVES * 64;
    constexpr int RADIX = 1 << BITS;
    constexpr int TILE = THREADS * KPT;
    __shared__ uint32_t s_h[WAVES * RADIX];
    const uint32_t tid = threadIdx.x, wave = tid / VREN_WAVE;
    for (uint32_t i = tid; i < (uint32_t)(WAVES * RADIX); i += THREADS) s_h[i] = 0u;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint64_t tile_start = (uint64_t)tile * TILE;
    const uint32_t tile_count = (uint32_t)min<uint64_t>((uint64_t)TILE, n - tile_start);
    const uint32_t mask = (1u << width) - 1u;
    uint32_t* h = s_h + wave * RADIX;
    const uint32_t* src = keys + tile_start;
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const uint32_t n4 = tile_count / 4;
        for (uint32_t i = tid; i < n4; i += THREADS) {
            const uint4 q = ntload4(reinterpret_cast<const uint4*>(src) + i);
            atomicAdd(&h[(q.x >> shift) & mask], 1u);
            atomicAdd(&h[(q.y >> shift) & mask], 1u);
            atomicAdd(&h[(q.z >> shift) & mask], 1u);
            atomicAdd(&h[(q.w >> shift) & mask], 1u);
        }
        for (uint32_t i = n4 * 4 + tid; i < tile_count; i += THREADS) atomicAdd(&h[(src[i] >> shift) & mask], 1u);
    } else {
        for (uint32_t i = tid; i < tile_count; i += THREADS) atomicAdd(&h[(src[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (uint32_t d = tid; d < (uint32_t)RADIX; d += THREADS) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < WAVES; w++) c += s_h[w * RADIX + d];
        tile_hist[(uint64_t)d * tiles + tile] = c;
    }
}

// keys and values are staged through LDS one after the other when staging both at once would leave room for
// fewer than two workgroups per CU
template <int KPT, int WAVES, bool VALS>
constexpr bool rs_sequential_staging()
{
    return VALS && 2 * WAVES * 64 * KPT * 4 > 80 * 1024;
}

// RTS = false: onesweep (tiles from a ticket counter, per-digit decoupled look-back, digit_base = the digit's
// first output position). RTS = true: reduce-then-scan scatter pass (tile = workgroup id; digit_base = the scanned
// digit-major tile counts of rs_upsweep_kernel), no tickets and no waiting.
// IN_PAIRS: keys_in holds interleaved (key, value) pairs and vals_in is unused. OUT_PAIRS: the pass writes
// interleaved pairs to keys_out and vals_out is unused. Both need VALS.
// Bucket starts written by a final pass over interleaved pairs (OUT_PAIRS and BSTARTS): bucket(k) = (k >> shift) &
// mask. The staged keys of a tile are sorted by the whole bucket (the pass digit is the bucket's top bits and the
// earlier pass left each digit run sorted by the rest), so a staged key whose staged predecessor has the same pass
// digit has that key as its predecessor in the output too, and writes offsets[c] = its position for the buckets c in
// (bucket(predecessor), bucket(key)]. The first key of each tile's digit run and the buckets past the last key are
// left to bucket_run_starts_kernel, which reads them from the output.
struct BucketStarts {
    uint32_t* offsets;
    uint32_t shift, mask;
};

template <int BITS, int KPT, int WAVES, bool VALS, bool RTS = false, bool IN_PAIRS = false, bool OUT_PAIRS = false,
          bool BSTARTS = false>
__global__ __launch_bounds__(WAVES * 64) void rs_onesweep_kernel(const uint32_t* __restrict__ keys_in,
                                                                 const uint32_t* __restrict__ vals_in,
                                                                 uint32_t* __restrict__ keys_out,
                                                                 uint32_t* __restrict__ vals_out, uint64_t n,
                                                                 uint32_t shift, uint32_t width,
                                                                 const uint32_t* __restrict__ digit_base,
                                                                 uint32_t* status, uint32_t* counter,
                                                                 uint32_t* __restrict__ clear_next,
                                                                 const BucketStarts bs)
{
    constexpr int THREADS = WAVES * 64;
    constexpr int RADIX = 1 << BITS;
    constexpr int TILE = THREADS * KPT;
    constexpr bool SEQ = rs_sequential_staging<KPT, WAVES, VALS>();
    static_assert(THREADS >= RADIX, "one thread per digit");
    static_assert(VALS || !(IN_PAIRS || OUT_PAIRS), "interleaved pairs carry values");
    static_assert(OUT_PAIRS || !BSTARTS, "bucket starts come from a pass that writes pairs");
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[(VALS && !SEQ) ? TILE : 1];
    __shared__ uint32_t s_whist[WAVES * RADIX];
    __shared__ uint32_t s_goff[RADIX];
    __shared__ uint32_t s_scan[THREADS / VREN_WAVE];
    __shared__ uint32_t s_ticket;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid / VREN_WAVE;
    if (!RTS && tid == 0) s_ticket = atomic_add_agent(counter, 1u);
    for (uint32_t i = tid; i < (uint32_t)(WAVES * RADIX); i += THREADS) s_whist[i] = 0u;
    __syncthreads();
    const uint32_t tile = RTS ? blockIdx.x : s_ticket;
    const uint64_t tile_start = (uint64_t)tile * TILE;
    const uint32_t tile_count = (uint32_t)min<uint64_t>((uint64_t)TILE, n - tile_start);
    if (!RTS && clear_next) {
        for (uint32_t i = tid; i < (uint32_t)RADIX; i += THREADS) clear_next[(uint64_t)tile * RADIX + i] = 0u;
    }

    // warp-striped load; positions past the end of a partial tile hold 0xFFFFFFFF, whose digit is the largest, so
    // they rank after every real key of the tile and the real keys keep the staged positions [0, tile_count)
    const uint32_t wbase = wave * KPT * VREN_WAVE;
    const uint32_t* kin = keys_in + tile_start + wbase + lane;
    const uint32_t* vin = vals_in + tile_start + wbase + lane;
    uint32_t key[KPT], val[VALS ? KPT : 1], digit[KPT], local[KPT];
    if (IN_PAIRS) {
        const vu2* pin = reinterpret_cast<const vu2*>(keys_in) + tile_start + wbase + lane;
        const bool full_tile = tile_count == (uint32_t)TILE;
#pragma unroll
        for (int j = 0; j < KPT; j++) {
            vu2 q = {0xFFFFFFFFu, 0u};
            if (full_tile || wbase + j * VREN_WAVE + lane < tile_count) q = __builtin_nontemporal_load(pin + j * VREN_WAVE);
            key[j] = q.x;
            val[j] = q.y;
        }
    } else if (tile_count == (uint32_t)TILE) {
#pragma unroll
        for (int j = 0; j < KPT; j++) key[j] = __builtin_nontemporal_load(kin + j * VREN_WAVE);
    } else {
#pragma unroll
        for (int j = 0; j < KPT; j++)
            key[j] = wbase + j * VREN_WAVE + lane < tile_count ? __builtin_nontemporal_load(kin + j * VREN_WAVE)
                                                               : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int j = 0; j < KPT; j++) digit[j] = __builtin_amdgcn_ubfe(key[j], shift, width);
    rank_wave_keys<BITS, KPT, false>(digit, wbase, tile_count, &s_whist[wave * RADIX], local);
    // the values are read after the ranking: their loads overlap the counts, the look-back and the key stores
    if (VALS && !IN_PAIRS) {
#pragma unroll
        for (int j = 0; j < KPT; j++)
            val[j] = wbase + j * VREN_WAVE + lane < tile_count ? __builtin_nontemporal_load(vin + j * VREN_WAVE) : 0u;
    }
    __syncthreads();
    // per digit: tile count, then the exclusive prefix across waves offset by the digit's tile-local start, so a
    // key's staged position is s_whist[wave][d] + local
    uint32_t dcount = 0;
    uint32_t wc[WAVES];
    if (tid < (uint32_t)RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            wc[w] = s_whist[w * RADIX + tid];
            dcount += wc[w];
        }
    }
    // onesweep: the tile's per-digit counts (the padding of a partial tile included: only the last tile has
    // padding, and no tile looks back at the last one) are published before the block scan
    if (!RTS && tid < (uint32_t)RADIX)
        st_agent_u32(status + (uint64_t)tile * RADIX + tid, (tile == 0 ? RS_INCL : RS_AGG) | dcount);
    uint32_t total;
    uint32_t dstart = block_exclusive_sum<THREADS>(dcount, s_scan, &total);
    uint32_t prefix = 0;
    if (tid < (uint32_t)RADIX) {
        uint32_t run = dstart;
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            s_whist[w * RADIX + tid] = run;
            run += wc[w];
        }
        if (!RTS && tile != 0) {
            // windowed decoupled look-back on this digit: the states of the kLookback nearest unread predecessors
            // are loaded at once and their counts summed up to the first inclusive one; an unpublished state is
            // waited for and the walk resumes there (tile 0 is always inclusive, so the walk ends at or before it)
            int64_t t = (int64_t)tile - 1;
            for (;;) {
                uint32_t st[kLookback];
#pragma unroll
                for (int i = 0; i < kLookback; i++)
                    st[i] = t - i >= 0 ? ld_agent_u32(status + (uint64_t)(t - i) * RADIX + tid) : RS_INCL;
                int i = 0;
                bool done = false;
                for (; i < kLookback; i++) {
                    if ((st[i] & ~RS_MASK) == 0u) break;
                    prefix += st[i] & RS_MASK;
                    if (st[i] & RS_INCL) {
                        done = true;
                        break;
                    }
                }
                if (done) break;
                t -= i;
                if (i < kLookback) __builtin_amdgcn_s_sleep(1);
            }
            st_agent_u32(status + (uint64_t)tile * RADIX + tid, RS_INCL | (prefix + dcount));
        }
        // global destination of staged position i = s_goff[digit] + i: RTS from the scanned digit-major tile
        // counts, onesweep from the pass histogram's exclusive scan (rs_hexscan_kernel) plus the look-back prefix
        if (RTS) {
            const uint32_t tiles = (uint32_t)((n + TILE - 1) / TILE);
            s_goff[tid] = digit_base[(uint64_t)tid * tiles + tile] - dstart;
        } else {
            s_goff[tid] = digit_base[tid] + prefix - dstart;
        }
    }
    __syncthreads();
    // scatter into LDS in tile-local sorted order (every staged position is < TILE, padding included)
#pragma unroll
    for (int j = 0; j < KPT; j++) {
        local[j] += s_whist[wave * RADIX + digit[j]];
        s_keys[local[j]] = key[j];
        if (VALS && !SEQ) s_vals[local[j]] = val[j];
    }
    __syncthreads();
    // staged position i = tid + j * THREADS goes to s_goff[digit] + i (< n: the digit's count over the tiles up to
    // this one never exceeds its count in the pass histogram). A full tile reads all its staged keys before the
    // first store, so the LDS reads overlap.
    const bool full = tile_count == (uint32_t)TILE;
    if constexpr (OUT_PAIRS) {
        // keys and values of staged position i leave together as one 8-byte pair (positions past the end of a
        // partial tile hold padding and are not stored)
        uint32_t k[KPT], o[KPT], v[KPT];
#pragma unroll
        for (int j = 0; j < KPT; j++) {
            k[j] = s_keys[tid + j * THREADS];
            o[j] = s_goff[__builtin_amdgcn_ubfe(k[j], shift, width)] + tid + j * THREADS;
        }
        if constexpr (BSTARTS) {
            // the staged predecessor, read before the staging is reused for values
#pragma unroll
            for (int j = 0; j < KPT; j++) {
                const uint32_t i = tid + j * THREADS;
                if (i == 0 || !(full || i < tile_count)) continue;
                const uint32_t kp = s_keys[i - 1];
                if (__builtin_amdgcn_ubfe(kp, shift, width) != __builtin_amdgcn_ubfe(k[j], shift, width)) continue;
                const uint32_t b1 = (k[j] >> bs.shift) & bs.mask;
                for (uint32_t c = ((kp >> bs.shift) & bs.mask) + 1u; c <= b1; c++) bs.offsets[c] = o[j];
            }
        }
        if (SEQ) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KPT; j++) s_keys[local[j]] = val[j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KPT; j++) v[j] = s_keys[tid + j * THREADS];
        } else {
#pragma unroll
            for (int j = 0; j < KPT; j++) v[j] = s_vals[tid + j * THREADS];
        }
        vu2* pout = reinterpret_cast<vu2*>(keys_out);
#pragma unroll
        for (int j = 0; j < KPT; j++)
            if (full || tid + j * THREADS < tile_count) pout[o[j]] = vu2{k[j], v[j]};
        return;
    }
    uint32_t o[KPT];
    if (full) {
        uint32_t k[KPT];
#pragma unroll
        for (int j = 0; j < KPT; j++) k[j] = s_keys[tid + j * THREADS];
#pragma unroll
        for (int j = 0; j < KPT; j++) o[j] = s_goff[__builtin_amdgcn_ubfe(k[j], shift, width)] + tid + j * THREADS;
        if (VALS && !SEQ) {
            uint32_t v[KPT];
#pragma unroll
            for (int j = 0; j < KPT; j++) v[j] = s_vals[tid + j * THREADS];
#pragma unroll
            for (int j = 0; j < KPT; j++) {
                keys_out[o[j]] = k[j];
                vals_out[o[j]] = v[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < KPT; j++) keys_out[o[j]] = k[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < KPT; j++) {
            const uint32_t i = tid + j * THREADS;
            o[j] = 0u;
            if (i < tile_count) {
                const uint32_t k = s_keys[i];
                o[j] = s_goff[__builtin_amdgcn_ubfe(k, shift, width)] + i;
                keys_out[o[j]] = k;
                if (VALS && !SEQ) vals_out[o[j]] = s_vals[i];
            }
        }
    }
    if (SEQ) {
        // the values through the same staging buffer, to the destinations kept from the keys
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KPT; j++) s_keys[local[j]] = val[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KPT; j++) {
            const uint32_t i = tid + j * THREADS;
            if (full || i < tile_count) vals_out[o[j]] = s_keys[i];
        }
    }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
    size_t tmp_keys, tmp_vals, hist, counters, status_a, status_b, total;
};

Layout layout_for(uint64_t n, bool vals, uint32_t min_tile, uint32_t max_radix)
{
    Layout l;
    uint64_t tiles = ceil_div_u64(std::max<uint64_t>(n, 1), min_tile);
    size_t off = 0;
    l.tmp_keys = off;
    off += align256(n * 4);
    l.tmp_vals = off;
    off += vals ? align256(n * 4) : 0;
    l.hist = off;
    off += align256(8 * 256 * 4);
    l.counters = off;
    off += 256;
    l.status_a = off;
    off += align256(tiles * max_radix * 4);
    l.status_b = off;
    off += align256(tiles * max_radix * 4);
    l.total = off;
    return l;
}

// smallest tile among the instantiated configurations (4 waves x 8 keys per thread)
constexpr uint32_t kMinTile = 4 * 64 * 8;

// reduce-then-scan LSD sort: per pass, tile digit counts -> one exclusive scan -> scatter (the tile-count array
// and the scan workspace reuse the onesweep status regions of the scratch)
template <int BITS, int KPT, int WAVES>
int run_sort_rts(vren_ctx* ctx, hipStream_t s, uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t begin_bit,
                 uint32_t end_bit, char* scratch, const Layout& L)
{
    constexpr int RADIX = 1 << BITS;
    constexpr int TILE = WAVES * 64 * KPT;
    const uint32_t passes = (end_bit - begin_bit + BITS - 1) / BITS;
    const uint32_t tiles = (uint32_t)ceil_div_u64(n, TILE);
    uint32_t* tile_hist = reinterpret_cast<uint32_t*>(scratch + L.status_a);
    void* scan_work = scratch + L.status_b;
    uint32_t* tk = reinterpret_cast<uint32_t*>(scratch + L.tmp_keys);
    uint32_t* tv = vals ? reinterpret_cast<uint32_t*>(scratch + L.tmp_vals) : nullptr;
    uint32_t *ksrc = keys, *vsrc = vals, *kdst = tk, *vdst = tv;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t shift = begin_bit + p * BITS;
        uint32_t width = std::min<uint32_t>(BITS, end_bit - shift);
        rs_upsweep_kernel<BITS, KPT, WAVES><<<tiles, WAVES * 64, 0, s>>>(ksrc, n, shift, width, tile_hist, tiles);
        VREN_LAUNCH_CHECK(ctx);
        int st = vren_scan_with_work(ctx, s, tile_hist, tile_hist, (uint64_t)RADIX * tiles, 1, scan_work);
        if (st) return st;
        if (vals) {
            rs_onesweep_kernel<BITS, KPT, WAVES, true, true><<<tiles, WAVES * 64, 0, s>>>(
                ksrc, vsrc, kdst, vdst, n, shift, width, tile_hist, nullptr, nullptr, nullptr, BucketStarts{});
        } else {
            rs_onesweep_kernel<BITS, KPT, WAVES, false, true><<<tiles, WAVES * 64, 0, s>>>(
                ksrc, nullptr, kdst, nullptr, n, shift, width, tile_hist, nullptr, nullptr, nullptr, BucketStarts{});
        }
        VREN_LAUNCH_CHECK(ctx);
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
    }
    if (ksrc != keys) {
        VREN_HIP(ctx, hipMemcpyAsync(keys, ksrc, n * 4, hipMemcpyDeviceToDevice, s));
        if (vals) VREN_HIP(ctx, hipMemcpyAsync(vals, vsrc, n * 4, hipMemcpyDeviceToDevice, s));
    }
    return VREN_OK;
}

template <int BITS, int KPT, int WAVES>
int run_sort(vren_ctx* ctx, hipStream_t s, uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t begin_bit,
             uint32_t end_bit, char* scratch, const Layout& L)
{
    constexpr int RADIX = 1 << BITS;
    constexpr int TILE = WAVES * 64 * KPT;
    const uint32_t passes = (end_bit - begin_bit + BITS - 1) / BITS;
    const uint64_t tiles = ceil_div_u64(n, TILE);
    uint32_t* hist = reinterpret_cast<uint32_t*>(scratch + L.hist);
    uint32_t* counters = reinterpret_cast<uint32_t*>(scratch + L.counters);
    uint32_t* st[2] = {reinterpret_cast<uint32_t*>(scratch + L.status_a),
                       reinterpret_cast<uint32_t*>(scratch + L.status_b)};
    uint32_t* tk = reinterpret_cast<uint32_t*>(scratch + L.tmp_keys);
    uint32_t* tv = vals ? reinterpret_cast<uint32_t*>(scratch + L.tmp_vals) : nullptr;

    rs_init_kernel<<<1, 256, 0, s>>>(hist, passes * RADIX, counters, passes);
    VREN_LAUNCH_CHECK(ctx);
    uint32_t hb = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus, std::max<uint64_t>(1, ceil_div_u64(n, 16384)));
    rs_hist_kernel<BITS><<<hb, kHistThreads, 0, s>>>(keys, n, begin_bit, end_bit, passes, hist, st[0], tiles * RADIX);
    VREN_LAUNCH_CHECK(ctx);
    rs_hexscan_kernel<BITS><<<1, 256, 0, s>>>(hist, passes);
    VREN_LAUNCH_CHECK(ctx);
    uint32_t *ksrc = keys, *vsrc = vals, *kdst = tk, *vdst = tv;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t shift = begin_bit + p * BITS;
        uint32_t width = std::min<uint32_t>(BITS, end_bit - shift);
        uint32_t* clear_next = p + 1 < passes ? st[(p + 1) & 1] : nullptr;
        if (vals) {
            rs_onesweep_kernel<BITS, KPT, WAVES, true><<<(uint32_t)tiles, WAVES * 64, 0, s>>>(
                ksrc, vsrc, kdst, vdst, n, shift, width, hist + p * RADIX, st[p & 1], counters + p, clear_next,
                BucketStarts{});
        } else {
            rs_onesweep_kernel<BITS, KPT, WAVES, false><<<(uint32_t)tiles, WAVES * 64, 0, s>>>(
                ksrc, nullptr, kdst, nullptr, n, shift, width, hist + p * RADIX, st[p & 1], counters + p, clear_next,
                BucketStarts{});
        }
        VREN_LAUNCH_CHECK(ctx);
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
    }
    if (ksrc != keys) {
        VREN_HIP(ctx, hipMemcpyAsync(keys, ksrc, n * 4, hipMemcpyDeviceToDevice, s));
        if (vals) VREN_HIP(ctx, hipMemcpyAsync(vals, vsrc, n * 4, hipMemcpyDeviceToDevice, s));
    }
    return VREN_OK;
}

// The bucket starts that the final pass leaves out: the first key of every tile's digit run (its output predecessor
// belongs to another tile or digit) and the buckets past the last key. Run (t, d) starts at digit_base[d] plus the
// inclusive count of digit d over tiles < t, which the pass's look-back states hold; one thread per (tile, digit).
__global__ __launch_bounds__(256) void bucket_run_starts_kernel(const uint32_t* __restrict__ pairs, uint64_t n,
                                                                const uint32_t* __restrict__ status,
                                                                const uint32_t* __restrict__ digit_base, uint32_t tiles,
                                                                uint32_t radix, const BucketStarts bs,
                                                                uint32_t n_buckets)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (uint64_t)tiles * radix) return;
    const uint32_t t = (uint32_t)(g / radix), d = (uint32_t)(g - (uint64_t)t * radix);
    const uint32_t before = t == 0 ? 0u : status[(uint64_t)(t - 1) * radix + d] & RS_MASK;
    const uint32_t upto = status[(uint64_t)t * radix + d] & RS_MASK;
    if (upto > before) {
        const uint64_t q = (uint64_t)digit_base[d] + before;
        if (q < n) { // a run of real keys (the last tile's padding counts but is not in the output)
            const uint32_t bq = (pairs[2 * q] >> bs.shift) & bs.mask;
            const uint32_t first = q == 0 ? 0u : ((pairs[2 * q - 2] >> bs.shift) & bs.mask) + 1u;
            for (uint32_t c = first; c <= bq; c++) bs.offsets[c] = (uint32_t)q;
        }
    }
    if (g == 0) {
        const uint32_t bl = (pairs[2 * (n - 1)] >> bs.shift) & bs.mask;
        for (uint32_t c = bl + 1u; c < n_buckets; c++) bs.offsets[c] = (uint32_t)n;
    }
}

// stable sort of interleaved (key, value) pairs on bits [begin_bit, end_bit), at most two digit passes: the first
// pass reads the pairs and the last one writes them (a single pass does both); between two passes the keys and
// values sit split in the scratch. Per pair: 8 B histogram read + 16 B per pass. With bucket_offsets, the final
// pass and bucket_run_starts_kernel also write the start of every bucket (k >> begin_bit) & (n_buckets - 1).
template <int BITS, int KPT, int WAVES>
int run_sort_pairs(vren_ctx* ctx, hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t begin_bit,
                   uint32_t end_bit, char* scratch, const Layout& L, uint32_t* bucket_offsets, uint32_t n_buckets)
{
    constexpr int RADIX = 1 << BITS;
    constexpr int TILE = WAVES * 64 * KPT;
    const uint32_t passes = (end_bit - begin_bit + BITS - 1) / BITS;
    if (passes > 2) return vren_fail(ctx, VREN_EINVAL, "radix_sort: interleaved pairs take at most %d bits", 2 * BITS);
    const uint64_t tiles = ceil_div_u64(n, TILE);
    uint32_t* hist = reinterpret_cast<uint32_t*>(scratch + L.hist);
    uint32_t* counters = reinterpret_cast<uint32_t*>(scratch + L.counters);
    uint32_t* st[2] = {reinterpret_cast<uint32_t*>(scratch + L.status_a),
                       reinterpret_cast<uint32_t*>(scratch + L.status_b)};
    uint32_t* tk = reinterpret_cast<uint32_t*>(scratch + L.tmp_keys);
    uint32_t* tv = reinterpret_cast<uint32_t*>(scratch + L.tmp_vals);

    rs_init_kernel<<<1, 256, 0, s>>>(hist, passes * RADIX, counters, passes);
    VREN_LAUNCH_CHECK(ctx);
    uint32_t hb = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus, std::max<uint64_t>(1, ceil_div_u64(n, 16384)));
    rs_hist_kernel<BITS, true><<<hb, kHistThreads, 0, s>>>(in, n, begin_bit, end_bit, passes, hist, st[0], tiles * RADIX);
    VREN_LAUNCH_CHECK(ctx);
    rs_hexscan_kernel<BITS><<<1, 256, 0, s>>>(hist, passes);
    VREN_LAUNCH_CHECK(ctx);
    const BucketStarts bs{bucket_offsets, begin_bit, n_buckets ? n_buckets - 1u : 0u};
    if (passes == 1) {
        if (bucket_offsets)
            rs_onesweep_kernel<BITS, KPT, WAVES, true, false, true, true, true><<<(uint32_t)tiles, WAVES * 64, 0, s>>>(
                in, nullptr, out, nullptr, n, begin_bit, end_bit - begin_bit, hist, st[0], counters, nullptr, bs);
        else
            rs_onesweep_kernel<BITS, KPT, WAVES, true, false, true, true><<<(uint32_t)tiles, WAVES * 64, 0, s>>>(
                in, nullptr, out, nullptr, n, begin_bit, end_bit - begin_bit, hist, st[0], counters, nullptr, bs);
        VREN_LAUNCH_CHECK(ctx);
    } else {
        const uint32_t shift1 = begin_bit + BITS;
        rs_onesweep_kernel<BITS, KPT, WAVES, true, false, true, false><<<(uint32_t)tiles, WAVES * 64, 0, s>>>(
            in, nullptr, tk, tv, n, begin_bit, BITS, hist, st[0], counters, st[1], BucketStarts{});
        VREN_LAUNCH_CHECK(ctx);
        if (bucket_offsets)
            rs_onesweep_kernel<BITS, KPT, WAVES, true, false, false, true, true><<<(uint32_t)tiles, WAVES * 64, 0, s>>>(
                tk, tv, out, nullptr, n, shift1, end_bit - shift1, hist + RADIX, st[1], counters + 1, nullptr, bs);
        else
            rs_onesweep_kernel<BITS, KPT, WAVES, true, false, false, true><<<(uint32_t)tiles, WAVES * 64, 0, s>>>(
                tk, tv, out, nullptr, n, shift1, end_bit - shift1, hist + RADIX, st[1], counters + 1, nullptr, bs);
        VREN_LAUNCH_CHECK(ctx);
    }
    if (bucket_offsets) {
        const uint64_t threads = tiles * RADIX;
        bucket_run_starts_kernel<<<(uint32_t)ceil_div_u64(threads, 256), 256, 0, s>>>(
            out, n, st[(passes - 1) & 1], hist + (uint64_t)(passes - 1) * RADIX, (uint32_t)tiles, RADIX, bs, n_buckets);
        VREN_LAUNCH_CHECK(ctx);
    }
    return VREN_OK;
}

typedef int (*sort_fn)(vren_ctx*, hipStream_t, uint32_t*, uint32_t*, uint64_t, uint32_t, uint32_t, char*, const Layout&);

struct SortConfig {
    int bits, kpt, waves;
    sort_fn fn;
    int algorithm; // 0 onesweep with look-back, 1 reduce-then-scan
};

#define RS_CFG(B, K, W) {B, K, W, &run_sort<B, K, W>, 0}
#define RS_CFG_RTS(B, K, W) {B, K, W, &run_sort_rts<B, K, W>, 1}
const SortConfig kConfigs[] = {
    RS_CFG(8, 8, 4),  RS_CFG(8, 15, 4), RS_CFG(8, 23, 4), RS_CFG(8, 8, 8),  RS_CFG(8, 15, 8),
    RS_CFG(8, 23, 8), RS_CFG(8, 8, 16), RS_CFG(8, 15, 16), RS_CFG(8, 23, 16), RS_CFG(6, 8, 4),  RS_CFG(6, 15, 4), RS_CFG(6, 23, 4),
    RS_CFG(6, 8, 8),  RS_CFG(6, 15, 8), RS_CFG(6, 23, 8), RS_CFG(6, 8, 16), RS_CFG(4, 8, 4),
    RS_CFG(4, 15, 4), RS_CFG(4, 23, 4), RS_CFG(4, 8, 8),  RS_CFG(4, 15, 8), RS_CFG(4, 23, 8),
    RS_CFG(4, 8, 16),
    RS_CFG_RTS(8, 15, 8), RS_CFG_RTS(8, 23, 8), RS_CFG_RTS(8, 8, 16), RS_CFG_RTS(8, 15, 16), RS_CFG_RTS(8, 23, 4),
};

} // namespace

int vren_radix_config_count() { return (int)(sizeof(kConfigs) / sizeof(kConfigs[0])); }

void vren_radix_config_info(int i, int* bits, int* kpt, int* waves, int* algorithm)
{
    *bits = kConfigs[i].bits;
    *kpt = kConfigs[i].kpt;
    *waves = kConfigs[i].waves;
    if (algorithm) *algorithm = kConfigs[i].algorithm;
}

int vren_radix_sort_with_config(vren_ctx* ctx, hipStream_t s, int cfg, uint32_t* keys, uint32_t* vals, uint64_t n,
                                uint32_t begin_bit, uint32_t end_bit, void* scratch, size_t scratch_bytes)
{
    Layout L = layout_for(n, vals != nullptr, kMinTile, 256);
    if (scratch_bytes < L.total)
        return vren_fail(ctx, VREN_EINVAL, "radix_sort: scratch of %zu bytes, %zu required", scratch_bytes, L.total);
    if (begin_bit >= end_bit) return VREN_OK;
    return kConfigs[cfg].fn(ctx, s, keys, vals, n, begin_bit, end_bit, static_cast<char*>(scratch), L);
}

int vren_radix_sort_pairs_interleaved(vren_ctx* ctx, hipStream_t s, const uint32_t* in_pairs, uint32_t* out_pairs,
                                      uint64_t n, uint32_t begin_bit, uint32_t end_bit, void* scratch,
                                      size_t scratch_bytes, uint32_t* bucket_offsets, uint32_t n_buckets)
{
    Layout L = layout_for(n, true, kMinTile, 256);
    if (scratch_bytes < L.total)
        return vren_fail(ctx, VREN_EINVAL, "radix_sort: scratch of %zu bytes, %zu required", scratch_bytes, L.total);
    if (n == 0) return VREN_OK;
    if (begin_bit >= end_bit) {
        if (bucket_offsets) return vren_fail(ctx, VREN_EINVAL, "radix_sort: bucket starts need a bit range");
        VREN_HIP(ctx, hipMemcpyAsync(out_pairs, in_pairs, n * 8, hipMemcpyDeviceToDevice, s));
        return VREN_OK;
    }
    // the default configuration's tile (8-bit digits, 23 keys per thread, 16 waves); below 2^21 pairs, 11 keys per
    // thread (more tiles to spread over the CUs: 2^20 pairs in 54 instead of 73 us for 65536 buckets)
    if (n < (1u << 21))
        return run_sort_pairs<8, 11, 16>(ctx, s, in_pairs, out_pairs, n, begin_bit, end_bit,
                                         static_cast<char*>(scratch), L, bucket_offsets, n_buckets);
    return run_sort_pairs<8, 23, 16>(ctx, s, in_pairs, out_pairs, n, begin_bit, end_bit, static_cast<char*>(scratch), L,
                                     bucket_offsets, n_buckets);
}

extern "C" {

VREN_EXPORT size_t vren_radix_sort_scratch_bytes(uint64_t n, int with_values)
{
    return layout_for(n, with_values != 0, kMinTile, 256).total;
}

VREN_EXPORT int vren_radix_sort(vren_ctx* ctx, vren_stream_t stream, uint32_t* keys, uint32_t* vals, uint64_t n,
                                uint32_t begin_bit, uint32_t end_bit, void* scratch, size_t scratch_bytes)
{
    if (!ctx) return VREN_EINVAL;
    if (!keys || n == 0 || !scratch) return vren_fail(ctx, VREN_EINVAL, "radix_sort: null buffer or n == 0");
    if (n >= (1ull << 30)) return vren_fail(ctx, VREN_EINVAL, "radix_sort: n must be below 2^30");
    if (begin_bit > end_bit || end_bit > 32) return vren_fail(ctx, VREN_EINVAL, "radix_sort: bad bit range");
    if (scratch_bytes < vren_radix_sort_scratch_bytes(n, vals != nullptr))
        return vren_fail(ctx, VREN_EINVAL, "radix_sort: scratch of %zu bytes, %zu required", scratch_bytes,
                         vren_radix_sort_scratch_bytes(n, vals != nullptr));
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "radix_sort: context has no device");
    int cfg = vren_radix_pick_config(ctx, (hipStream_t)stream, n, vals != nullptr, end_bit - begin_bit);
    return vren_radix_sort_with_config(ctx, (hipStream_t)stream, cfg, keys, vals, n, begin_bit, end_bit, scratch,
                                       scratch_bytes);
}

} // extern "C"
