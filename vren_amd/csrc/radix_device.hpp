// radix_device.hpp -- the stable tile ranking shared by the onesweep radix sort and the per-frame light sort.
//
// Keys are held warp-striped: key j of lane l of wave w sits at tile position (w * KPT + j) * 64 + l, so the
// order (w, j, l) is the input order and ranking in that order is stable. A key's digit peers inside the wave
// come from BITS ballots (wave64 match), its rank among them from mbcnt, and the wave's running per-digit
// count lives in LDS (one row of RADIX counters per wave). After ranking, per-digit counts are turned into
// exclusive prefixes across waves, giving every key its tile-local position:
//   digit_start[d] + wave_prefix[w][d] + local[j].
#pragma once

#include "common.hpp"

template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid)
{
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t m = ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// s_whist: WAVES * RADIX counters, zeroed before the call (and a __syncthreads() after zeroing).
// Writes local[j] = rank of key j among the same-digit keys of this wave that precede it.
template <int BITS, int KPT>
__device__ __forceinline__ void rank_wave_keys(const uint32_t (&digit)[KPT], uint32_t valid_count_base,
                                               uint32_t tile_count, uint32_t* s_wave_row, uint32_t (&local)[KPT])
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int j = 0; j < KPT; j++) {
        uint32_t pos = valid_count_base + (uint32_t)j * VREN_WAVE + lane;
        bool valid = pos < tile_count;
        uint32_t d = digit[j];
        uint64_t peers = match_digit<BITS>(d, valid);
        uint32_t before = 0;
        if (valid) before = s_wave_row[d];
        uint32_t rank = mbcnt(peers);
        // the highest peer lane publishes the new running count
        if (valid && (peers >> lane) == 1ull) s_wave_row[d] = before + rank + 1u;
        local[j] = before + rank;
        __builtin_amdgcn_wave_barrier();
    }
}
