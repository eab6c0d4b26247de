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

// One digit bit of the wave match: bc = the lane's bit b broadcast to a word (v_bfe_i32), VCC = its ballot, and
// per 32-lane half peers &= ~(ballot ^ bc) in one v_bitop3_b32 (truth table over S0 = 0xF0, S1 = 0xCC,
// S2 = 0xAA: 0xF0 & ~(0xCC ^ 0xAA) = 0x90). A VALU write of VCC needs two wait states before a VALU reads it as
// an operand (s_nop 1).
template <int B>
__device__ __forceinline__ void match_bit(uint32_t d, uint32_t& lo, uint32_t& hi)
{
    uint32_t bc;
    asm("v_bfe_i32 %2, %3, %4, 1\n\t"
        "v_cmp_ne_u32_e32 vcc, 0, %2\n\t"
        "s_nop 1\n\t"
        "v_bitop3_b32 %0, %0, vcc_lo, %2 bitop3:0x90\n\t"
        "v_bitop3_b32 %1, %1, vcc_hi, %2 bitop3:0x90"
        : "+v"(lo), "+v"(hi), "=&v"(bc)
        : "v"(d), "i"(B)
        : "vcc");
}

// Bit 0 of a match that starts from every lane: peers = ~(ballot ^ bc) (truth table ~(0xCC ^ 0xAA) = 0x99, S0
// unused).
__device__ __forceinline__ void match_bit0_all(uint32_t d, uint32_t& lo, uint32_t& hi)
{
    uint32_t bc;
    asm("v_bfe_i32 %2, %3, 0, 1\n\t"
        "v_cmp_ne_u32_e32 vcc, 0, %2\n\t"
        "s_nop 1\n\t"
        "v_bitop3_b32 %0, %2, vcc_lo, %2 bitop3:0x99\n\t"
        "v_bitop3_b32 %1, %2, vcc_hi, %2 bitop3:0x99"
        : "=&v"(lo), "=&v"(hi), "=&v"(bc)
        : "v"(d)
        : "vcc");
}

template <int B, int BITS>
__device__ __forceinline__ void match_bits(uint32_t d, uint32_t& lo, uint32_t& hi)
{
    if constexpr (B < BITS) {
        match_bit<B>(d, lo, hi);
        match_bits<B + 1, BITS>(d, lo, hi);
    }
}

// Lanes of the wave holding the same BITS-bit digit as this lane, four VALU operations per digit bit.
// ALL = true: every lane of the wave takes part (lo:hi are outputs only); ALL = false: only the lanes set in lo:hi
// on entry.
template <int BITS, bool ALL = false>
__device__ __forceinline__ void match_digit(uint32_t d, uint32_t& lo, uint32_t& hi)
{
    if constexpr (ALL) {
        match_bit0_all(d, lo, hi);
        match_bits<1, BITS>(d, lo, hi);
    } else {
        match_bits<0, BITS>(d, lo, hi);
    }
}

// s_whist: WAVES * RADIX counters, zeroed before the call (and a __syncthreads() after zeroing).
// Writes local[j] = rank of key j among the same-digit keys of this wave that precede it.
// CHECK = true: positions at or past tile_count are excluded (neither ranked nor counted). CHECK = false: every
// lane takes part; a caller padding a partial tile gives its padding the largest digit, so that the padding ranks
// after every real key of the tile and the real keys keep positions [0, tile_count).
template <int BITS, int KPT, bool CHECK = true>
__device__ __forceinline__ void rank_wave_keys(const uint32_t (&digit)[KPT], uint32_t valid_count_base,
                                               uint32_t tile_count, uint32_t* s_wave_row, uint32_t (&local)[KPT])
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int j = 0; j < KPT; j++) {
        const uint32_t d = digit[j];
        bool valid = true;
        uint32_t lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;
        if (CHECK) {
            valid = valid_count_base + (uint32_t)j * VREN_WAVE + lane < tile_count;
            const uint64_t v = ballot(valid);
            lo = (uint32_t)v;
            hi = (uint32_t)(v >> 32);
        }
        const uint32_t before = valid ? s_wave_row[d] : 0u;
        match_digit<BITS, !CHECK>(d, lo, hi);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
        const uint64_t peers = ((uint64_t)hi << 32) | lo;
        // the highest peer lane publishes the new running count
        if (valid && (peers >> lane) == 1ull) s_wave_row[d] = before + rank + 1u;
        local[j] = before + rank;
        __builtin_amdgcn_wave_barrier();
    }
}
