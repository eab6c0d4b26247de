// autotune.hip -- project-written autotuner for the onesweep radix sort (the chain's shapes are tuned by
// vren_batch_autotune in batch.hip; both kinds of choice are saved to and loaded from the same file here).
//
// The reference fixes the sort at 4-bit digits, 1024-thread workgroups and one key per thread
// (radix_sort.hpp:15-16). Here the digit width (4, 6, 8 bits), keys per thread (8, 15, 23) and waves per
// workgroup (4, 8, 16) are template parameters; every combination that fits in LDS is compiled, and the tuner
// times each one on fresh pseudo-random keys of the caller's size class with HIP events, keeping the fastest.
// Choices are cached per context, keyed by (log2 n, with values, key bits), and can be saved to / loaded from
// a small text file ("bucket config" per line). All configurations produce identical output (stable sort).
#include <cstdlib>
#include <cstring>
#include <vector>

#include "radix_sort.hpp"

namespace {

// defaults when nothing is tuned or forced: 8-bit digits, 23 keys per thread, 16 waves, onesweep (the fastest on
// MI355X at 2^26 key-value pairs)
constexpr int kDefaultBits = 8, kDefaultKpt = 23, kDefaultWaves = 16, kDefaultAlgorithm = 0;

uint64_t bucket_key(uint64_t n, bool vals, uint32_t key_bits)
{
    return ((uint64_t)ilog2_u64(n) << 8) | ((uint64_t)(vals ? 1 : 0) << 7) | key_bits;
}

int default_config()
{
    for (int i = 0; i < vren_radix_config_count(); i++) {
        int b, k, w, a;
        vren_radix_config_info(i, &b, &k, &w, &a);
        if (b == kDefaultBits && k == kDefaultKpt && w == kDefaultWaves && a == kDefaultAlgorithm) return i;
    }
    return 0;
}

__global__ void fill_random_kernel(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t seed)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        // splitmix-style integer hash of (seed, i)
        uint64_t z = (i + ((uint64_t)seed << 40)) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        keys[i] = (uint32_t)z;
        if (vals) vals[i] = (uint32_t)i;
    }
}

} // namespace

int vren_radix_pick_config(vren_ctx* ctx, hipStream_t s, uint64_t n, bool vals, uint32_t key_bits)
{
    if (ctx->radix_forced >= 0) return ctx->radix_forced;
    auto it = ctx->radix_tuned.find(bucket_key(n, vals, key_bits));
    if (it != ctx->radix_tuned.end()) return it->second;
    (void)s;
    return default_config();
}

extern "C" {

VREN_EXPORT int vren_radix_sort_config_count(void) { return vren_radix_config_count(); }

VREN_EXPORT int vren_radix_sort_config_info(int config, uint32_t* radix_bits, uint32_t* keys_per_thread,
                                            uint32_t* waves_per_block, uint32_t* algorithm)
{
    if (config < 0 || config >= vren_radix_config_count()) return VREN_EINVAL;
    int b, k, w, a;
    vren_radix_config_info(config, &b, &k, &w, &a);
    if (radix_bits) *radix_bits = (uint32_t)b;
    if (keys_per_thread) *keys_per_thread = (uint32_t)k;
    if (waves_per_block) *waves_per_block = (uint32_t)w;
    if (algorithm) *algorithm = (uint32_t)a;
    return VREN_OK;
}

VREN_EXPORT int vren_radix_sort_force_config(vren_ctx* ctx, int config)
{
    if (!ctx || config < -1 || config >= vren_radix_config_count()) return VREN_EINVAL;
    ctx->radix_forced = config;
    return VREN_OK;
}

VREN_EXPORT int vren_autotune_radix_sort(vren_ctx* ctx, vren_stream_t stream, uint64_t n, int with_values,
                                         uint32_t key_bits, int* best_config, float* best_ms)
{
    if (!ctx || n == 0 || n >= (1ull << 30) || key_bits == 0 || key_bits > 32) return VREN_EINVAL;
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "autotune: context has no device");
    hipStream_t s = (hipStream_t)stream;
    size_t scratch_bytes = vren_radix_sort_scratch_bytes(n, with_values);
    uint32_t *keys = nullptr, *vals = nullptr;
    void* scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = VREN_OK;
    int best = -1;
    float best_t = 0.f;
    auto cleanup = [&]() {
        vren_teardown(hipStreamSynchronize(s));
        if (keys) vren_teardown(hipFree(keys));
        if (vals) vren_teardown(hipFree(vals));
        if (scratch) vren_teardown(hipFree(scratch));
        if (e0) vren_teardown(hipEventDestroy(e0));
        if (e1) vren_teardown(hipEventDestroy(e1));
    };
    if (hipMalloc(&keys, n * 4) != hipSuccess || (with_values && hipMalloc(&vals, n * 4) != hipSuccess) ||
        hipMalloc(&scratch, scratch_bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return vren_fail(ctx, VREN_ENOMEM, "autotune: allocation of %llu keys failed", (unsigned long long)n);
    }
    const int reps = 3;
    for (int c = 0; c < vren_radix_config_count() && status == VREN_OK; c++) {
        float acc = 0.f;
        for (int r = 0; r <= reps && status == VREN_OK; r++) {
            fill_random_kernel<<<1024, 256, 0, s>>>(keys, vals, n, (uint32_t)(c * 16 + r));
            if (hipEventRecord(e0, s) != hipSuccess) {
                status = vren_fail(ctx, VREN_EHIP, "autotune: event record failed");
                break;
            }
            status = vren_radix_sort_with_config(ctx, s, c, keys, vals, n, 0, key_bits, scratch, scratch_bytes);
            float ms = 0.f;
            if (status == VREN_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                status = vren_fail(ctx, VREN_EHIP, "autotune: timing the configuration failed");
            if (r > 0) acc += ms; // first run warms the configuration up
        }
        float t = acc / reps;
        if (status == VREN_OK && (best < 0 || t < best_t)) {
            best = c;
            best_t = t;
        }
    }
    cleanup();
    if (status != VREN_OK) return status;
    ctx->radix_tuned[bucket_key(n, with_values != 0, key_bits)] = best;
    if (best_config) *best_config = best;
    if (best_ms) *best_ms = best_t;
    return VREN_OK;
}

VREN_EXPORT int vren_autotune_save(vren_ctx* ctx, const char* path)
{
    if (!ctx || !path) return VREN_EINVAL;
    FILE* f = fopen(path, "w");
    if (!f) return vren_fail(ctx, VREN_EINVAL, "autotune_save: cannot open %s", path);
    for (auto& kv : ctx->radix_tuned) {
        int b, k, w, a;
        vren_radix_config_info(kv.second, &b, &k, &w, &a);
        fprintf(f, "radix %llu %d %d %d %d\n", (unsigned long long)kv.first, b, k, w, a);
    }
    for (auto& kv : ctx->chain_tuned)
        fprintf(f, "chain %llu %u %u %u %u\n", (unsigned long long)kv.first, kv.second.shade_pixels_per_thread,
                kv.second.assign_waves_per_cu, kv.second.light_sort_bits, kv.second.light_stage_concurrent);
    fclose(f);
    return VREN_OK;
}

VREN_EXPORT int vren_autotune_load(vren_ctx* ctx, const char* path)
{
    if (!ctx || !path) return VREN_EINVAL;
    FILE* f = fopen(path, "r");
    if (!f) return vren_fail(ctx, VREN_EINVAL, "autotune_load: cannot open %s", path);
    char line[256];
    while (fgets(line, sizeof(line), f)) {
        unsigned long long key;
        int b, k, w, a;
        unsigned p, q, r, c = 0;
        if (sscanf(line, " radix %llu %d %d %d %d", &key, &b, &k, &w, &a) == 5) {
            for (int i = 0; i < vren_radix_config_count(); i++) {
                int b2, k2, w2, a2;
                vren_radix_config_info(i, &b2, &k2, &w2, &a2);
                if (b == b2 && k == k2 && w == w2 && a == a2) ctx->radix_tuned[(uint64_t)key] = i;
            }
        } else if (sscanf(line, " chain %llu %u %u %u %u", &key, &p, &q, &r, &c) >= 4) {
            // only shapes a batch accepts (the same rules as vren_batch_set_shape)
            if ((p == 1 || p == 2) && (q == 8 || q == 16 || q == 24 || q == 32) && (r == 4 || r == 8) && c <= 1)
                ctx->chain_tuned[(uint64_t)key] = vren_chain_shape{p, q, r, c};
        }
    }
    fclose(f);
    return VREN_OK;
}

} // extern "C"
