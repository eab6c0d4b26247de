// bucket_sort.hip -- bucket_sort (bucket_sort.hpp:9-45, bucket_sort.cpp:72-161, bucket_sort_{count,write}.comp).
//
// The reference counts (key & 0xFFFF) with global atomics into 65 536 counters stored after the output pairs,
// scans them with a fixed 2^16 Blelloch scan, and scatters with atomicAdd slot allocation, so the order inside a
// bucket is nondeterministic. Here the bucket id b = (key >> shift) & (n_buckets - 1) of every pair is sorted with
// the stable onesweep radix sort (log2(n_buckets) bits) carrying the pair's index, the pairs are gathered in that
// order, and the bucket starts are the exclusive scan of the bucket histogram: a stable counting sort, i.e. one
// valid outcome of the reference's sort, and the canonical one.
#include "radix_sort.hpp"

namespace {

__global__ void bucket_clear_kernel(uint32_t* counts, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) counts[i] = 0u;
}

__global__ void bucket_keys_kernel(const uint2* __restrict__ in, uint32_t n, uint32_t shift, uint32_t mask,
                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t* counts)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t b = (in[i].x >> shift) & mask;
        keys[i] = b;
        vals[i] = i;
        atomicAdd(&counts[b], 1u);
    }
}

__global__ void bucket_gather_kernel(const uint2* __restrict__ in, const uint32_t* __restrict__ perm, uint32_t n,
                                     uint2* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[perm[i]];
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
    const uint32_t bits = ilog2_u64(n_buckets);
    const uint64_t nn = n ? n : 1;
    const size_t rs_bytes = vren_radix_sort_scratch_bytes(nn, 1);
    const size_t scan_bytes = vren_scan_work_bytes(n_buckets, 1);
    size_t off_keys = 0, off_vals = a256(nn * 4), off_counts = off_vals + a256(nn * 4);
    size_t off_scan = off_counts + a256((size_t)n_buckets * 4), off_rs = off_scan + a256(scan_bytes);
    int st = vren_ensure_scratch(ctx, off_rs + rs_bytes);
    if (st) return st;
    char* base = static_cast<char*>(ctx->scratch);
    uint32_t* keys = reinterpret_cast<uint32_t*>(base + off_keys);
    uint32_t* vals = reinterpret_cast<uint32_t*>(base + off_vals);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + off_counts);
    bucket_clear_kernel<<<std::max<uint32_t>(1, std::min<uint32_t>(256, (n_buckets + 255) / 256)), 256, 0, s>>>(counts, n_buckets);
    VREN_LAUNCH_CHECK(ctx);
    if (n > 0) {
        const uint32_t blocks = std::min<uint32_t>(4096, (n + 255) / 256);
        bucket_keys_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint2*>(in_pairs), n, key_shift, n_buckets - 1,
                                                 keys, vals, counts);
        VREN_LAUNCH_CHECK(ctx);
        if (bits > 0) {
            int cfg = vren_radix_pick_config(ctx, s, n, true, bits);
            st = vren_radix_sort_with_config(ctx, s, cfg, keys, vals, n, 0, bits, base + off_rs, rs_bytes);
            if (st) return st;
        }
        bucket_gather_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint2*>(in_pairs), vals, n,
                                                   reinterpret_cast<uint2*>(out_pairs));
        VREN_LAUNCH_CHECK(ctx);
    }
    return vren_scan_with_work(ctx, s, counts, bucket_offsets, n_buckets, 1, base + off_scan);
}
