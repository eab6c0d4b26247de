// bucket_sort.hip -- bucket_sort (bucket_sort.hpp:9-45, bucket_sort.cpp:72-161, bucket_sort_{count,write}.comp).
//
// The reference counts (key & 0xFFFF) with global atomics into 65 536 counters stored after the output pairs,
// scans them with a fixed 2^16 Blelloch scan, and scatters with atomicAdd slot allocation, so the order inside a
// bucket is nondeterministic. Here the bucket of a pair is the bit field [shift, shift + log2(n_buckets)) of its
// key, so a stable LSD radix sort restricted to those bits, carrying the pair's value, is a stable counting sort
// of the pairs -- one valid outcome of the reference's sort, and the canonical one:
//   sort     the interleaved pairs on the bucket bits with the onesweep radix sort (one or two 8-bit passes; the
//            first reads the pairs, the last writes them; no atomics, no gather);
//   offsets  bucket b starts at the first sorted position whose bucket is >= b, so each position i writes the
//            starts of the buckets in (bucket(i - 1), bucket(i)], and the last position those after it (each start
//            is written exactly once). The final sort pass does this for every key whose predecessor it staged
//            itself, and a small kernel over the tiles' digit runs for the first key of each run, so the sorted
//            pairs are not read again.
#include "radix_sort.hpp"

namespace {

__global__ void bucket_fill_kernel(uint32_t* offsets, uint32_t n_buckets, uint32_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_buckets; i += gridDim.x * blockDim.x) offsets[i] = v;
}

} // namespace

extern "C" VREN_EXPORT int vren_bucket_sort(vren_ctx* ctx, vren_stream_t stream, const uint32_t* in_pairs, uint32_t n,
                                            uint32_t* out_pairs, uint32_t* bucket_offsets, uint32_t n_buckets,
                                            uint32_t key_shift)
{
    if (!ctx) return VREN_EINVAL;
    if (!in_pairs || !out_pairs || !bucket_offsets)
        return vren_fail(ctx, VREN_EINVAL, "bucket_sort: null buffer");
    if (n_buckets == 0 || n_buckets > 65536 || (n_buckets & (n_buckets - 1)) != 0 || key_shift > 31)
        return vren_fail(ctx, VREN_EINVAL, "bucket_sort: n_buckets must be a power of 2 in [1, 65536], shift < 32");
    if (n >= (1u << 30)) return vren_fail(ctx, VREN_EINVAL, "bucket_sort: n must be below 2^30");
    if (in_pairs == out_pairs) return vren_fail(ctx, VREN_EINVAL, "bucket_sort: input and output must differ");
    if (((uintptr_t)in_pairs | (uintptr_t)out_pairs) & 7)
        return vren_fail(ctx, VREN_EINVAL, "bucket_sort: pair buffers must be 8-byte aligned");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "bucket_sort: context has no device");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) {
        bucket_fill_kernel<<<std::max<uint32_t>(1, std::min<uint32_t>(256, (n_buckets + 255) / 256)), 256, 0, s>>>(
            bucket_offsets, n_buckets, 0u);
        VREN_LAUNCH_CHECK(ctx);
        return VREN_OK;
    }
    const uint32_t bits = ilog2_u64(n_buckets);
    if (bits == 0) { // one bucket: the pairs in order, starting at 0
        VREN_HIP(ctx, hipMemcpyAsync(out_pairs, in_pairs, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
        bucket_fill_kernel<<<1, 256, 0, s>>>(bucket_offsets, 1u, 0u);
        VREN_LAUNCH_CHECK(ctx);
        return VREN_OK;
    }
    // bits of the key at or above 32 are zero in the bucket number and need no sorting
    const uint32_t end_bit = std::min<uint32_t>(32u, key_shift + bits);
    const size_t rs_bytes = vren_radix_sort_scratch_bytes(n, 1);
    int st = vren_ensure_scratch(ctx, rs_bytes, s);
    if (st) return st;
    return vren_radix_sort_pairs_interleaved(ctx, s, in_pairs, out_pairs, n, key_shift, end_bit, ctx->scratch, rs_bytes,
                                             bucket_offsets, n_buckets);
}
