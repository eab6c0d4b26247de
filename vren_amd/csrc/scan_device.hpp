// scan_device.hpp -- block-level scan pieces and the single-pass decoupled look-back (Merrill & Garland 2016)
// shared by the exclusive scan primitive, cluster compaction and the light-offset scan.
#pragma once

#include "common.hpp"

// Exclusive prefix of `v` over the whole block (THREADS a multiple of 64). s_wave: THREADS / 64 words.
// Returns the exclusive prefix; *total receives the block sum. Contains __syncthreads().
template <int THREADS>
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* s_wave, uint32_t* total)
{
    constexpr int W = THREADS / VREN_WAVE;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / VREN_WAVE;
    uint32_t incl = wave_inclusive_sum(v);
    if (lane == VREN_WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
        uint32_t x = s_wave[w];
        if (w < (int)wave) before += x;
        sum += x;
    }
    *total = sum;
    __syncthreads();
    return before + incl - v;
}

// Decoupled look-back in two steps, for a caller with work to do between them: wave_lookback_publish (any one
// lane) publishes tile t's aggregate (tile 0: its inclusive value); wave_lookback_walk (ONE full wave) returns the
// exclusive prefix of the segment's tiles before t and publishes t's inclusive value. `status` points at the first
// status word of the segment; status words must be VREN_LB_EMPTY before the launch. Tile 0 of a segment never
// waits. Forward progress: tiles are taken from a ticket counter, so every predecessor is already running.
__device__ __forceinline__ void wave_lookback_publish(uint64_t* status, uint32_t t, uint32_t agg)
{
    st_agent_u64(&status[t], (t == 0 ? VREN_LB_INCL : VREN_LB_AGG) | agg);
}

__device__ __forceinline__ uint32_t wave_lookback_walk(uint64_t* status, uint32_t t, uint32_t agg)
{
    const uint32_t lane = lane_id();
    if (t == 0) return 0u;
    uint32_t prefix = 0;
    int64_t base = (int64_t)t - 1;
    while (true) {
        int64_t idx = base - (int64_t)lane;
        uint64_t s = VREN_LB_INCL; // before the segment start: an inclusive zero
        if (idx >= 0) {
            s = ld_agent_u64(&status[idx]);
            while (VREN_LB_STATE(s) == VREN_LB_EMPTY) {
                __builtin_amdgcn_s_sleep(1);
                s = ld_agent_u64(&status[idx]);
            }
        }
        uint64_t incl_mask = ballot(VREN_LB_STATE(s) == VREN_LB_INCL);
        uint32_t v = (uint32_t)s;
        if (incl_mask) {
            uint32_t first = (uint32_t)__builtin_ctzll(incl_mask);
            prefix += wave_sum(lane <= first ? v : 0u);
            break;
        }
        prefix += wave_sum(v);
        base -= VREN_WAVE;
    }
    if (lane == 0) st_agent_u64(&status[t], VREN_LB_INCL | (uint32_t)(prefix + agg));
    return prefix;
}

// both steps at once (nothing between publishing the aggregate and walking)
__device__ __forceinline__ uint32_t wave_lookback(uint64_t* status, uint32_t t, uint32_t agg)
{
    if (lane_id() == 0) wave_lookback_publish(status, t, agg);
    return wave_lookback_walk(status, t, agg);
}
