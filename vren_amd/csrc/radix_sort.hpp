// radix_sort.hpp -- internal interface between the onesweep sort and its autotuner.
#pragma once

#include "common.hpp"

int vren_radix_config_count();
void vren_radix_config_info(int i, int* bits, int* kpt, int* waves, int* algorithm);
int vren_radix_sort_with_config(vren_ctx* ctx, hipStream_t s, int cfg, uint32_t* keys, uint32_t* vals, uint64_t n,
                                uint32_t begin_bit, uint32_t end_bit, void* scratch, size_t scratch_bytes);
// stable sort of interleaved (key, value) pairs (8-byte aligned) from in_pairs into out_pairs on bits
// [begin_bit, end_bit), at most 16 bits; scratch as for vren_radix_sort with values. With bucket_offsets (n > 0),
// also writes the first output position of every bucket (k >> begin_bit) & (n_buckets - 1), n past the last key's.
int vren_radix_sort_pairs_interleaved(vren_ctx* ctx, hipStream_t s, const uint32_t* in_pairs, uint32_t* out_pairs,
                                      uint64_t n, uint32_t begin_bit, uint32_t end_bit, void* scratch,
                                      size_t scratch_bytes, uint32_t* bucket_offsets = nullptr,
                                      uint32_t n_buckets = 0);
// configuration for this problem: forced, tuned (cached), or the default
int vren_radix_pick_config(vren_ctx* ctx, hipStream_t s, uint64_t n, bool vals, uint32_t key_bits);

// exclusive scan with a caller-provided look-back workspace of vren_scan_work_bytes(n, blocks)
size_t vren_scan_work_bytes(uint64_t n, uint32_t blocks);
int vren_scan_with_work(vren_ctx* ctx, hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t blocks,
                        void* work);
