// clusters.hip -- find_unique_cluster_list (clustered_shading.cpp:363-448, find_unique_clusters.comp:46-121).
//
// One 256-thread workgroup per 32x32 tile (flattened over every frame of a batch); each thread derives the
// 32-bit cluster keys of 4 adjacent pixels (depth slice from the D32 depth, cube-face normal code from the
// RGBA16F normal). The reference sorts the tile's 1024 keys with a 55-step LDS bitonic network and 20 barriers;
// here each wave extracts its unique keys in ascending order by repeated wave64 minima (one iteration per
// distinct key -- a tile of a real scene holds a handful), wave 0 merges the four sorted wave lists the same
// way, and the tile's unique count is chained to the previous tiles of the same frame by a decoupled look-back. This replaces the reference's atomicAdd allocation (whose
// order across tiles is nondeterministic, find_unique_clusters.comp:100-103) by the canonical row-major tile
// order, and needs no separate scan dispatch. Each pixel finds its key among the tile's unique keys by binary
// search and stores the frame-wide cluster index.
//
// Defined behaviour where the reference has none: pixels of partial tiles sample (x mod W, y mod H) like its
// REPEAT sampler (gbuffer.cpp:28-34) and do not store a reference; a view depth below near gives slice 0; keys
// past the key capacity are not stored and raise the capacity flag.
#include "chain.hpp"
#include "scan_device.hpp"

namespace {

__device__ __forceinline__ uint32_t discretize_normal(const float n[3])
{
    // clustered_shading.glsl:7-43
    if (n[0] == 0.0f && n[1] == 0.0f && n[2] == 0.0f) return 0xFFFFFFFFu;
    float min_t = 1e35f;
    uint32_t axis = 0, face_idx = 0;
    for (uint32_t i = 0; i < 3; i++) {
        float sgn = n[i] > 0.0f ? 1.0f : (n[i] < 0.0f ? -1.0f : 0.0f);
        float t = sgn / n[i];
        if (t < min_t) {
            min_t = t;
            axis = i;
            face_idx = (n[i] > 0.0f ? 1u : 0u) * 3u + i;
        }
    }
    float p[3] = {n[0] * min_t, n[1] * min_t, n[2] * min_t};
    float u = p[(axis + 1) % 3];
    float v = p[(axis + 2) % 3];
    uint32_t du = vren_f2u_sat(vren_floorf((u + 1.0f) / 2.0f * 3.0f));
    uint32_t dv = vren_f2u_sat(vren_floorf((v + 1.0f) / 2.0f * 3.0f));
    return (face_idx * 9u + du * 3u + dv) & 0x3Fu;
}

__device__ __forceinline__ uint32_t face_of_normal(const float n[3])
{
    if (n[0] == 0.0f && n[1] == 0.0f && n[2] == 0.0f) return 7u;
    float min_t = 1e35f;
    uint32_t face_idx = 0;
    for (uint32_t i = 0; i < 3; i++) {
        float sgn = n[i] > 0.0f ? 1.0f : (n[i] < 0.0f ? -1.0f : 0.0f);
        float t = sgn / n[i];
        if (t < min_t) {
            min_t = t;
            face_idx = (n[i] > 0.0f ? 1u : 0u) * 3u + i;
        }
    }
    return face_idx;
}

__device__ __forceinline__ uint32_t pixel_key(const FrameDev& fd, uint32_t ti, uint32_t tj, uint32_t gx, uint32_t gy,
                                              float depth, uint2 nb)
{
    // find_unique_clusters.comp:48-65
    float v[4] = {(float)gx / (float)fd.W, (float)gy / (float)fd.H, depth, 1.0f};
    v[1] = 1.0f - v[1];
    v[0] = v[0] * 2.0f - 1.0f;
    v[1] = v[1] * 2.0f - 1.0f;
    float r[4];
    vren_mat4_mul_vec4(fd.inv_proj, v, r);
    const float z = r[2] / r[3];
    const float n[3] = {vren_half_to_float((uint16_t)(nb.x & 0xFFFFu)), vren_half_to_float((uint16_t)(nb.x >> 16)),
                        vren_half_to_float((uint16_t)(nb.y & 0xFFFFu))};
    if (fd.key_mode == VREN_KEYS_SLICED) {
        float s = vren_floorf(vren_logf(z / fd.near_plane) / fd.log_fn * (float)fd.depth_slices);
        uint32_t k = vren_f2u_sat(s);
        if (k >= fd.depth_slices) k = fd.depth_slices - 1;
        return (ti & 0xFFu) | ((tj & 0xFFu) << 8) | ((k & 0xFFu) << 16) | (face_of_normal(n) << 24);
    }
    const uint32_t k = vren_f2u_sat(vren_floorf(vren_logf(z / fd.near_plane) / fd.log_a));
    return (ti & 0xFFu) | ((tj & 0xFFu) << 8) | ((k & 0x3FFu) << 16) | (discretize_normal(n) << 26);
}

__global__ void chain_init_kernel(uint64_t* status, uint32_t n_status, uint32_t* tickets)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_status; i += gridDim.x * blockDim.x)
        status[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < 4) tickets[threadIdx.x] = 0u;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t o = shfl_xor(v, m);
        v = o < v ? o : v;
    }
    return v;
}

constexpr uint64_t kNone = 1ull << 32; // larger than every 32-bit key

// One 256-thread workgroup per 32x32 tile; thread t owns the 4 horizontally adjacent pixels
// (4 * (t % 8) + q, t / 8), q = 0..3, so depth and normal loads are 16- and 32-byte vectors.
__global__ __launch_bounds__(256) void clusters_kernel(const FrameDev* frames, uint32_t F, uint64_t* tile_status,
                                                       uint32_t* ticket)
{
    __shared__ uint32_t s_wlist[4][256];
    __shared__ uint32_t s_wn[4];
    __shared__ uint32_t s_unique[1024];
    __shared__ uint32_t s_t, s_count, s_prefix;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid / VREN_WAVE;
    if (tid == 0) s_t = atomic_add_agent(ticket, 1u);
    __syncthreads();
    const uint32_t g = s_t;
    const uint32_t f = find_frame_wave(frames, F, g, [](const FrameDev& d) { return d.tile_base; });
    const FrameDev& fd = frames[f];
    const uint32_t t = g - fd.tile_base;
    const uint32_t ti = t % fd.Sx, tj = t / fd.Sx;
    const uint32_t gy = tj * 32 + (tid >> 3);
    const uint32_t gx0 = ti * 32 + (tid & 7) * 4;
    // the thread's 4 pixels: one 16-byte depth load and two 16-byte normal loads when they lie inside the frame
    // on one aligned row; otherwise per pixel, wrapping (x mod W, y mod H) like the reference's REPEAT sampler
    float dep[4];
    uint2 nrm[4];
    const uint64_t p0 = (uint64_t)gy * fd.W + gx0;
    if (gy < fd.H && gx0 + 3 < fd.W && (p0 & 3) == 0 && ((reinterpret_cast<uintptr_t>(fd.depth) | reinterpret_cast<uintptr_t>(fd.normals)) & 15) == 0) {
        const float4 d4 = *reinterpret_cast<const float4*>(fd.depth + p0);
        const uint4 n01 = *reinterpret_cast<const uint4*>(fd.normals + 4 * p0);
        const uint4 n23 = *reinterpret_cast<const uint4*>(fd.normals + 4 * p0 + 8);
        dep[0] = d4.x;
        dep[1] = d4.y;
        dep[2] = d4.z;
        dep[3] = d4.w;
        nrm[0] = make_uint2(n01.x, n01.y);
        nrm[1] = make_uint2(n01.z, n01.w);
        nrm[2] = make_uint2(n23.x, n23.y);
        nrm[3] = make_uint2(n23.z, n23.w);
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t p = (uint64_t)(gy % fd.H) * fd.W + ((gx0 + q) % fd.W);
            dep[q] = fd.depth[p];
            nrm[q] = reinterpret_cast<const uint2*>(fd.normals)[p];
        }
    }
    uint32_t key[4];
#pragma unroll
    for (int q = 0; q < 4; q++) key[q] = pixel_key(fd, ti, tj, gx0 + q, gy, dep[q], nrm[q]);

    // unique keys of the wave's 256 pixels, extracted in ascending order by repeated wave minima
    uint32_t rem = 0xFu;
    uint32_t n = 0;
    while (true) {
        uint64_t mine = kNone;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if ((rem >> q) & 1u) mine = (uint64_t)key[q] < mine ? (uint64_t)key[q] : mine;
        const uint64_t m = wave_min_u64(mine);
        if (m == kNone) break;
        if (lane == 0) s_wlist[wave][n] = (uint32_t)m;
        n++;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (key[q] == (uint32_t)m) rem &= ~(1u << q);
    }
    if (lane == 0) s_wn[wave] = n;
    __syncthreads();
    // merge the four sorted wave lists into the tile's sorted unique list (wave 0)
    if (wave == 0) {
        const uint32_t n0 = s_wn[0], n1 = s_wn[1], n2 = s_wn[2], n3 = s_wn[3];
        const uint32_t total = n0 + n1 + n2 + n3;
        uint32_t ent[16];
        uint32_t emask = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t e = lane + 64u * i;
            ent[i] = 0u;
            if (e < total) {
                uint32_t w = 0, idx = e;
                if (idx >= n0) { idx -= n0; w = 1; if (idx >= n1) { idx -= n1; w = 2; if (idx >= n2) { idx -= n2; w = 3; } } }
                ent[i] = s_wlist[w][idx];
                emask |= 1u << i;
            }
        }
        uint32_t u = 0;
        while (true) {
            uint64_t mine = kNone;
#pragma unroll
            for (int i = 0; i < 16; i++)
                if ((emask >> i) & 1u) mine = (uint64_t)ent[i] < mine ? (uint64_t)ent[i] : mine;
            const uint64_t m = wave_min_u64(mine);
            if (m == kNone) break;
            if (lane == 0) s_unique[u] = (uint32_t)m;
            u++;
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (ent[i] == (uint32_t)m) emask &= ~(1u << i);
        }
        // decoupled look-back over the frame's tiles (row-major order)
        const uint32_t prefix = wave_lookback(tile_status + fd.tile_base, t, u);
        if (lane == 0) {
            s_count = u;
            s_prefix = prefix;
        }
    }
    __syncthreads();
    const uint32_t count = s_count, prefix = s_prefix;
    for (uint32_t i = tid; i < count; i += 256) {
        const uint32_t idx = prefix + i;
        if (idx < fd.key_cap) fd.keys[idx] = s_unique[i];
    }
    if (tid == 0 && t == fd.Sx * fd.Sy - 1) {
        const uint32_t total = prefix + count;
        *fd.cluster_count = total;
        if (total > fd.key_cap) atomic_or_agent(fd.flags, VREN_FLAG_KEY_CAPACITY);
    }
    if (gy < fd.H) {
        uint32_t ref[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t lo = 0, hi = count;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_unique[mid] <= key[q]) lo = mid;
                else hi = mid;
            }
            ref[q] = prefix + lo;
        }
        uint32_t* row = fd.refs + (uint64_t)gy * fd.W;
        if (gx0 + 3 < fd.W && ((reinterpret_cast<uintptr_t>(row + gx0) & 15) == 0)) {
            *reinterpret_cast<uint4*>(row + gx0) = make_uint4(ref[0], ref[1], ref[2], ref[3]);
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (gx0 + q < fd.W) row[gx0 + q] = ref[q];
        }
    }
}

} // namespace

void chain_work_bind(ChainWork* w, void* mem, uint32_t total_tiles, uint32_t total_offs_tiles)
{
    w->tickets = static_cast<uint32_t*>(mem);
    w->tile_status = reinterpret_cast<uint64_t*>(static_cast<char*>(mem) + 256);
    w->offs_status = w->tile_status + total_tiles;
    w->total_tiles = total_tiles;
    w->total_offs_tiles = total_offs_tiles;
}

int launch_chain_init(vren_ctx* ctx, hipStream_t s, const ChainWork& w)
{
    uint32_t n = w.total_tiles + w.total_offs_tiles;
    uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(256, (n + 255) / 256));
    chain_init_kernel<<<blocks, 256, 0, s>>>(w.tile_status, n, w.tickets);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    return VREN_OK;
}

int launch_clusters(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w)
{
    clusters_kernel<<<w.total_tiles, 256, 0, s>>>(d_frames, F, w.tile_status, w.tickets + 0);
    VREN_LAUNCH_CHECK(ctx);
    chain_mark(w, s);
    return VREN_OK;
}

extern "C" VREN_EXPORT int vren_find_unique_cluster_list(vren_ctx* ctx, vren_stream_t stream, uint32_t width,
                                                         uint32_t height, const vren_camera* camera,
                                                         const vren_cluster_cfg* cfg, const float* depth,
                                                         const uint16_t* normals, uint32_t* cluster_keys,
                                                         uint32_t key_capacity, uint32_t* cluster_count,
                                                         uint32_t* cluster_refs)
{
    if (!ctx) return VREN_EINVAL;
    int st = vren_validate_camera(ctx, camera);
    if (!st) st = vren_validate_cfg(ctx, cfg);
    if (st) return st;
    if (width == 0 || height == 0 || width > 8192 || height > 8192)
        return vren_fail(ctx, VREN_EINVAL, "find_unique_cluster_list: screen must be in [1, 8192]^2 (8-bit tile ids)");
    if (!depth || !normals || !cluster_keys || !cluster_count || !cluster_refs || key_capacity == 0)
        return vren_fail(ctx, VREN_EINVAL, "find_unique_cluster_list: null buffer or zero key capacity");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "find_unique_cluster_list: context has no device");
    hipStream_t s = (hipStream_t)stream;
    FrameDev f;
    memset(&f, 0, sizeof(f));
    vren_frame_setup(&f, width, height, camera, cfg);
    f.depth = depth;
    f.normals = normals;
    f.keys = cluster_keys;
    f.key_cap = key_capacity;
    f.cluster_count = cluster_count;
    f.refs = cluster_refs;
    f.flags = ctx->d_flags;
    ChainWork w;
    memset(&w, 0, sizeof(w));
    uint32_t tiles = f.Sx * f.Sy;
    st = vren_ensure_scratch(ctx, chain_work_bytes(tiles, 0));
    if (st) return st;
    chain_work_bind(&w, ctx->scratch, tiles, 0);
    FrameDev* slot = vren_next_frame_slot(ctx);
    st = vren_upload_frames(ctx, s, slot, &f, 1);
    if (!st) st = launch_chain_init(ctx, s, w);
    if (!st) st = launch_clusters(ctx, s, slot, 1, w);
    return st;
}
