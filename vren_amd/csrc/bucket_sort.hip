// bucket_sort.hip -- bucket_sort (bucket_sort.hpp:9-45, bucket_sort.cpp:72-161, bucket_sort_{count,write}.comp).
//
// The reference counts (key & 0xFFFF) with global atomics into 65 536 counters stored after the output pairs,
// scans them with a fixed 2^16 Blelloch scan, and scatters with atomicAdd slot allocation, so the order inside a
// bucket is nondeterministic. Here the bucket of a pair is the bit field [shift, shift + log2(n_buckets)) of its
// key, so a stable LSD radix sort restricted to those bits, carrying the pair's value, is a stable counting sort
// of the pairs -- one valid outcome of the reference's sort, and the canonical one:
//   split   the (key, value) pairs into a key array and a value array;
//   sort    the keys on the bucket bits with the onesweep radix sort, values following (no atomics, no gather);
//   merge   the sorted arrays back into pairs, and in the same pass the bucket starts: bucket b starts at the first
//           sorted position whose bucket is >= b, so each position i writes the starts of the buckets in
//           (bucket(i - 1), bucket(i)], and the last position those after it (each start is written exactly once).
#include "radix_sort.hpp"

namespace {

__global__ __launch_bounds__(256) void bucket_split_kernel(const uint2* __restrict__ in, uint32_t n,
                                                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint2 q = in[i];
        keys[i] = q.x;
        vals[i] = q.y;
    }
}

__global__ __launch_bounds__(256) void bucket_merge_kernel(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals, uint32_t n,
                                                           uint32_t shift, uint32_t mask, uint32_t n_buckets,
                                                           uint2* __restrict__ out, uint32_t* __restrict__ offsets)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        out[i] = make_uint2(k, vals[i]);
        const uint32_t b = (k >> shift) & mask;
        // buckets (prev, b] start here; prev = -1 before the first position
        const uint32_t first = i == 0 ? 0u : ((keys[i - 1] >> shift) & mask) + 1u;
        for (uint32_t c = first; c <= b; c++) offsets[c] = i;
        if (i == n - 1)
            for (uint32_t c = b + 1; c < n_buckets; c++) offsets[c] = n;
    }
}

__global__ void bucket_fill_kernel(uint32_t* offsets, uint32_t n_buckets, uint32_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_buckets; i += gridDim.x * blockDim.x) offsets[i] = v;
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

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
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "bucket_sort: context has no device");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) {
        bucket_fill_kernel<<<std::max<uint32_t>(1, std::min<uint32_t>(256, (n_buckets + 255) / 256)), 256, 0, s>>>(
            bucket_offsets, n_buckets, 0u);
        VREN_LAUNCH_CHECK(ctx);
        return VREN_OK;
    }
    const uint32_t bits = ilog2_u64(n_buckets);
    // bits of the key at or above 32 are zero in the bucket number and need no sorting
    const uint32_t end_bit = std::min<uint32_t>(32u, key_shift + bits);
    const size_t rs_bytes = vren_radix_sort_scratch_bytes(n, 1);
    const size_t off_keys = 0, off_vals = a256((size_t)n * 4), off_rs = off_vals + a256((size_t)n * 4);
    int st = vren_ensure_scratch(ctx, off_rs + rs_bytes);
    if (st) return st;
    char* base = static_cast<char*>(ctx->scratch);
    uint32_t* keys = reinterpret_cast<uint32_t*>(base + off_keys);
    uint32_t* vals = reinterpret_cast<uint32_t*>(base + off_vals);
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)ctx->num_cus * 16, (n + 255) / 256));
    bucket_split_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint2*>(in_pairs), n, keys, vals);
    VREN_LAUNCH_CHECK(ctx);
    if (end_bit > key_shift) {
        int cfg = vren_radix_pick_config(ctx, s, n, true, end_bit - key_shift);
        st = vren_radix_sort_with_config(ctx, s, cfg, keys, vals, n, key_shift, end_bit, base + off_rs, rs_bytes);
        if (st) return st;
    }
    bucket_merge_kernel<<<blocks, 256, 0, s>>>(keys, vals, n, key_shift, n_buckets - 1, n_buckets,
                                                reinterpret_cast<uint2*>(out_pairs), bucket_offsets);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}
