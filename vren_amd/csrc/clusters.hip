// clusters.hip -- find_unique_cluster_list (clustered_shading.cpp:363-448, find_unique_clusters.comp:46-121).
//
// One wave per 32x32 tile (tiles flattened over every frame of a batch); each lane derives the 32-bit cluster keys
// of 16 pixels (depth slice from the D32 depth, cube-face normal code from the RGBA16F normal). The reference
// sorts the tile's 1024 keys with a 55-step LDS bitonic network and 20 barriers; here the wave extracts the tile's
// unique keys in ascending order by repeated wave64 minima (one round per distinct key -- a tile of a real scene
// holds a handful), and the unique counts of a group of 8 tiles are chained to the previous groups of the same frame
// by a decoupled look-back. This replaces the reference's atomicAdd allocation (whose order across tiles is nondeterministic,
// find_unique_clusters.comp:100-103) by the canonical row-major tile order, and needs no separate scan dispatch.
// Each pixel keeps the index of its key in the tile's list from the round that emitted the key and stores the
// frame-wide cluster index.
//
// Defined behaviour where the reference has none: pixels of partial tiles sample (x mod W, y mod H) like its
// REPEAT sampler (gbuffer.cpp:28-34) and do not store a reference; a view depth below near gives slice 0; keys
// past the key capacity are not stored and raise the capacity flag.
#include "chain.hpp"
#include "scan_device.hpp"

namespace {

// The dominant axis of a normal: the reference picks the smallest t_i = sign(n_i) / n_i (first index on ties;
// a zero component gives NaN and is never picked), clustered_shading.glsl:7-43. The components are f16 values
// converted exactly, so distinct magnitudes differ by at least 2^-11 relative and their float reciprocals keep
// the reverse order: the smallest t_i is the largest |n_i| (first index on ties), found without dividing.
// Returns the axis, or 3 when no component qualifies (the reference then keeps t = 1e35 and axis 0).
__device__ __forceinline__ uint32_t dominant_axis(const float n[3])
{
    uint32_t axis = 3;
    float best = 0.0f;
#pragma unroll
    for (uint32_t i = 0; i < 3; i++) {
        const float a = fabsf(n[i]);
        if (a > best) {
            best = a;
            axis = i;
        }
    }
    return axis;
}

// 1 / a, correctly rounded, for a > 0 an f16 value widened to float: v_rcp_f32 (within 1 ulp) and one FMA Newton
// step. A divisor with an 11-bit significand keeps 1 / a at least 2^-36 (relative) away from every rounding
// boundary of float, far beyond the step's ~2^-46 error; the CPU tests check every finite f16 divisor against
// every reciprocal within 1 ulp. a = +inf gives +0, as the division does (the Newton step would give NaN).
__device__ __forceinline__ float rcp_f16_exact(float a)
{
    const float y = __builtin_amdgcn_rcpf(a);
    const float r = __builtin_fmaf(y, __builtin_fmaf(-a, y, 1.0f), y);
    return a < __builtin_inff() ? r : 0.0f;
}

__device__ __forceinline__ uint32_t discretize_normal(const float n[3])
{
    // clustered_shading.glsl:7-43
    if (n[0] == 0.0f && n[1] == 0.0f && n[2] == 0.0f) return 0xFFFFFFFFu;
    const uint32_t ax = dominant_axis(n);
    float min_t = 1e35f;
    uint32_t axis = 0, face_idx = 0;
    if (ax < 3) {
        axis = ax;
        // sign(n) / n = 1 / |n|, |n[ax]| > 0 being the largest magnitude (finite or +inf)
        min_t = rcp_f16_exact(fabsf(n[ax]));
        face_idx = (n[ax] > 0.0f ? 1u : 0u) * 3u + ax;
    }
    float p[3] = {n[0] * min_t, n[1] * min_t, n[2] * min_t};
    float u = p[(axis + 1) % 3];
    float v = p[(axis + 2) % 3];
    uint32_t du = vren_third_index(u);
    uint32_t dv = vren_third_index(v);
    return (face_idx * 9u + du * 3u + dv) & 0x3Fu;
}

__device__ __forceinline__ uint32_t face_of_normal(const float n[3])
{
    if (n[0] == 0.0f && n[1] == 0.0f && n[2] == 0.0f) return 7u;
    const uint32_t ax = dominant_axis(n);
    return ax < 3 ? (n[ax] > 0.0f ? 1u : 0u) * 3u + ax : 0u;
}

// The depth slice of a view depth z by the reference's formula (find_unique_clusters.comp:58-62): log mode
// f2u_sat(floor(log(z / near) / log_a)) before the 10-bit mask, sliced mode the same over [near, far] in
// depth_slices steps, clamped to the last slice. A non-positive or NaN z gives 0 (log of a non-positive ratio is
// NaN or -inf, which f2u_sat maps to 0).
__device__ __forceinline__ uint32_t slice_exact(const FrameDev& fd, float z)
{
    if (fd.key_mode == VREN_KEYS_SLICED) {
        uint32_t k = vren_f2u_sat(vren_floorf(vren_logf(z / fd.near_plane) / fd.log_fn * (float)fd.depth_slices));
        return k >= fd.depth_slices ? fd.depth_slices - 1 : k;
    }
    return vren_f2u_sat(vren_floorf(vren_logf(z / fd.near_plane) / fd.log_a));
}

// Slices kept in the threshold table (thr[1 .. kt - 1] are looked up; z >= thr[kt] takes the formula): every
// sliced-mode slice, log-mode slices below 1023 (the key keeps k & 0x3FF, so k >= 1023 goes through the formula)
__device__ __forceinline__ uint32_t slice_table_kt(const FrameDev& fd)
{
    return fd.key_mode == VREN_KEYS_SLICED ? fd.depth_slices - 1 : 1023u;
}

// thr[k] = the smallest positive float z with slice_exact(z) >= k, for k = 1 .. kt (+inf when no finite z reaches
// k), thr[0] = 0. slice_exact is non-decreasing in z: z / near and the division by log_a (or log_fn) are correctly
// rounded divisions by positive constants, floor and f2u_sat are monotone, and vren_logf is non-decreasing over
// every non-negative float (checked exhaustively by the oracle tests). So for 0 < z < thr[kt],
// slice_exact(z) = #{k in 1 .. kt - 1 : thr[k] <= z}, and positive floats order like their bit patterns, so each
// threshold is a binary search over bits (slice_exact(+inf) >= kt in both modes).
__global__ __launch_bounds__(1024) void slice_table_kernel(const FrameDev* __restrict__ frames, float* thr_all)
{
    const FrameDev& fd = frames[blockIdx.x];
    float* thr = thr_all + (uint64_t)blockIdx.x * VREN_SLICE_TABLE;
    const uint32_t kt = slice_table_kt(fd);
    const uint32_t k = threadIdx.x;
    if (k > kt) return;
    if (k == 0) {
        thr[0] = 0.0f;
        return;
    }
    // bisection over float bit patterns between lo (slice < k) and hi (slice >= k); it starts from +-1e-4 around
    // the real threshold near * exp(k log_a) (log mode) or near * exp(k log_fn / S) (sliced), in double, far wider
    // than the few ulps the float formula's rounding moves it, and from (0, +inf] when that bracket does not hold
    const double e = fd.key_mode == VREN_KEYS_SLICED ? (double)k * (double)fd.log_fn / (double)fd.depth_slices
                                                     : (double)k * (double)fd.log_a;
    const double zk = (double)fd.near_plane * exp(e);
    uint32_t lo = 0u, hi = 0x7F800000u; // slice_exact(+0) = 0 < k <= slice_exact(+inf)
    if (zk < 3.0e38) {
        const uint32_t blo = __float_as_uint((float)(zk * (1.0 - 1e-4)));
        const uint32_t bhi = __float_as_uint((float)(zk * (1.0 + 1e-4)));
        if (slice_exact(fd, __uint_as_float(blo)) < k && slice_exact(fd, __uint_as_float(bhi)) >= k) {
            lo = blo;
            hi = bhi;
        }
    }
    while (hi - lo > 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (slice_exact(fd, __uint_as_float(mid)) >= k) hi = mid;
        else lo = mid;
    }
    thr[k] = __uint_as_float(hi);
}

// view depth of a pixel (find_unique_clusters.comp:48-56)
__device__ __forceinline__ float pixel_view_z(const FrameDev& fd, uint32_t gx, uint32_t gy, float depth)
{
    float r2, r3;
    if (fd.proj_zw_xy_zero) {
        // the z and w rows of inverse(P) have zero x and y coefficients: ((0 x + 0 y) + m2 depth) + m3 reduces to
        // m2 depth + m3 exactly (a zero result may only change sign, and z = 0 gives the same slice either way)
        r2 = fd.inv_proj[2 * 4 + 2] * depth + fd.inv_proj[3 * 4 + 2];
        r3 = fd.inv_proj[2 * 4 + 3] * depth + fd.inv_proj[3 * 4 + 3];
    } else {
        float v[4] = {(float)gx / (float)fd.W, (float)gy / (float)fd.H, depth, 1.0f};
        v[1] = 1.0f - v[1];
        v[0] = v[0] * 2.0f - 1.0f;
        v[1] = v[1] * 2.0f - 1.0f;
        float r[4];
        vren_mat4_mul_vec4(fd.inv_proj, v, r);
        r2 = r[2];
        r3 = r[3];
    }
    return r2 / r3;
}

// the cluster key from its slice and normal (find_unique_clusters.comp:58-65)
// f16 -> f32 on the conversion unit. A key depends on a NaN component only through comparisons and through
// f2u_sat(floor(NaN)) = 0, never on its payload, so the conversion's NaN payload need not be the oracle's.
__device__ __forceinline__ float key_h2f(uint32_t h)
{
    const uint16_t b = (uint16_t)h;
    _Float16 x;
    __builtin_memcpy(&x, &b, 2);
    return (float)x;
}

__device__ __forceinline__ uint32_t assemble_key(const FrameDev& fd, uint32_t ti, uint32_t tj, uint32_t k, uint2 nb)
{
    const float n[3] = {key_h2f(nb.x & 0xFFFFu), key_h2f(nb.x >> 16), key_h2f(nb.y & 0xFFFFu)};
    if (fd.key_mode == VREN_KEYS_SLICED)
        return (ti & 0xFFu) | ((tj & 0xFFu) << 8) | ((k & 0xFFu) << 16) | (face_of_normal(n) << 24);
    return (ti & 0xFFu) | ((tj & 0xFFu) << 8) | ((k & 0x3FFu) << 16) | (discretize_normal(n) << 26);
}

__global__ void chain_init_kernel(uint64_t* status, uint32_t n_status, uint32_t* tickets)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_status; i += gridDim.x * blockDim.x)
        status[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < 4) tickets[threadIdx.x] = 0u;
}

constexpr uint32_t kClusterWaves = 8;   // waves per workgroup, one 32x32 tile per wave (16 measured slower)

// The depths and normals of lane `lane` of a tile's rows 2 h and 2 h + 1 (pixels i = 4 j + q with j = 2 h + jj,
// stored at [4 jj + q]): pixel (x0 + q, tj * 32 + 8 j + lane / 8). A tile inside the frame with 16-byte aligned
// rows (uniform across the wave) takes one 16-byte depth load and two 16-byte normal loads per row; other tiles
// load per pixel, wrapping (x mod W, y mod H) like the reference's REPEAT sampler. Every load of the half is issued
// before any is used; a tile's two halves are loaded one after the other, which keeps the register peak of a wave
// below 64 (eight waves per SIMD).
__device__ __forceinline__ void load_tile_half(const FrameDev& fd, uint32_t ti, uint32_t tj, uint32_t lane, uint32_t h,
                                               float dep[8], uint2 nrm[8])
{
    const uint32_t x0 = ti * 32 + (lane & 7) * 4;
    const bool interior = (ti + 1) * 32 <= fd.W && (tj + 1) * 32 <= fd.H && (fd.W & 3) == 0 &&
                          ((reinterpret_cast<uintptr_t>(fd.depth) | reinterpret_cast<uintptr_t>(fd.normals)) & 15) == 0;
    if (interior) {
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const uint64_t p0 = (uint64_t)(tj * 32 + (2 * h + jj) * 8 + (lane >> 3)) * fd.W + x0;
            const vf4 d4 = *(gvf4*)(fd.depth + p0);
            const vu4 n01 = *(gvu4*)(fd.normals + 4 * p0);
            const vu4 n23 = *(gvu4*)(fd.normals + 4 * p0 + 8);
            dep[4 * jj + 0] = d4.x;
            dep[4 * jj + 1] = d4.y;
            dep[4 * jj + 2] = d4.z;
            dep[4 * jj + 3] = d4.w;
            nrm[4 * jj + 0] = make_uint2(n01.x, n01.y);
            nrm[4 * jj + 1] = make_uint2(n01.z, n01.w);
            nrm[4 * jj + 2] = make_uint2(n23.x, n23.y);
            nrm[4 * jj + 3] = make_uint2(n23.z, n23.w);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t x = x0 + (i & 3), y = tj * 32 + (2 * h + (i >> 2)) * 8 + (lane >> 3);
            const uint64_t p = (uint64_t)(y % fd.H) * fd.W + (x % fd.W);
            dep[i] = *(gfloat*)(fd.depth + p);
            const vu2 n = *(gvu2*)(fd.normals + 4 * p);
            nrm[i] = make_uint2(n.x, n.y);
        }
    }
}

// Tiles are flattened over every frame of the table (tile_base, a multiple of VREN_TILE_GROUP). A workgroup takes
// ONE ticket for a group of kClusterWaves consecutive tiles of one frame, one per wave -- a ticket on a single
// counter costs ~11 ns of serialised atomic throughput, so one per tile would cap the kernel at ~90 M tiles/s.
// The group's unique counts are summed in LDS and chained to the frame's previous groups by one decoupled
// look-back per workgroup (per tile, the walk's round trips of ~1.4 us covered 64 tiles each and could not keep up
// with the tiles finishing). The walk waits only on groups of lower tickets, whose workgroups are already running,
// so every wait ends. Lane l owns the pixels (4 (l % 8) + q, 8 j + l / 8), q, j = 0..3, of its tile.
__global__ __launch_bounds__(kClusterWaves * 64, 8) void clusters_kernel(const FrameDev* __restrict__ frames,
                                                                       const uint32_t* __restrict__ tile_frame,
                                                                       uint32_t total_tiles, uint64_t* tile_status,
                                                                       uint32_t* ticket, const float* __restrict__ slice_thr,
                                                                       uint2* __restrict__ tile_cl)
{
    static_assert(kClusterWaves == VREN_TILE_GROUP, "one workgroup per tile group");
    static_assert(VREN_SLICE_TABLE == 2 * kClusterWaves * 64, "two table entries per thread");
    __shared__ uint32_t s_unique[kClusterWaves][1024];
    __shared__ float s_thr[VREN_SLICE_TABLE];
    __shared__ uint32_t s_u[kClusterWaves];
    __shared__ uint32_t s_t, s_prefix;
    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    if (threadIdx.x == 0) s_t = atomic_add_agent(ticket, 1u);
    __syncthreads();
    const uint32_t g0 = s_t * kClusterWaves;
    if (g0 >= total_tiles) return; // the whole workgroup
    // a workgroup's tiles are one group of its frame (frame tile ranges start at multiples of VREN_TILE_GROUP), so
    // one slice table serves them all
    const FrameDev& fd = frames[frame_of_tile(tile_frame, g0)];
    const uint32_t kt = slice_table_kt(fd);
    {
        // both halves of the table are requested before either is stored
        const uint32_t i0 = threadIdx.x, i1 = threadIdx.x + kClusterWaves * 64;
        const float* tb = slice_thr + (uint64_t)frame_of_tile(tile_frame, g0) * VREN_SLICE_TABLE;
        const float a0 = i0 <= kt ? tb[i0] : 0.0f, a1 = i1 <= kt ? tb[i1] : 0.0f;
        s_thr[i0] = a0;
        s_thr[i1] = a1;
    }
    __syncthreads();
    uint32_t* uq = s_unique[wave];
    const uint32_t g = g0 + wave;
    const uint32_t t0 = g0 - fd.tile_base; // the group's first tile in its frame
    const uint32_t t = t0 + wave;
    const uint32_t tiles = fd.Sx * fd.Sy;
    const bool live = g < total_tiles && t < tiles; // not past the table nor on a frame's padding
    const uint32_t ti = t % fd.Sx, tj = t / fd.Sx;
    const uint32_t x0 = ti * 32 + (lane & 7) * 4;
    uint32_t key[16];
    uint32_t u = 0;
    if (live) {
        const float* thr = s_thr;
        const float thr_kt = thr[kt];
        // slice estimate t = log2(z) c1 + c0 ~ ln(z / near) / log_a (log mode) or ln(z / near) S / log_fn (sliced):
        // its error is orders of magnitude below one slice, so the slice is the estimate's floor k0 moved by at
        // most one, decided by thr[k0] and thr[k0 + 1]; thr[k0 - 1] and thr[k0 + 2] confirm the move, and a pixel
        // whose confirmation fails (or beyond the table) takes the formula
        const float c1 = fd.key_mode == VREN_KEYS_SLICED ? 0.69314718f * (float)fd.depth_slices / fd.log_fn
                                                          : 0.69314718f / fd.log_a;
        const float c0 = -__builtin_amdgcn_logf(fd.near_plane) * c1;
        const float emax = (float)(kt > 0u ? kt - 1u : 0u); // k0 <= kt - 1, so every index below is <= kt
        uint32_t pend = 0u;
#pragma unroll
        for (uint32_t h = 0; h < 2; h++) {
        float dep[8];
        uint2 nrm[8];
        load_tile_half(fd, ti, tj, lane, h, dep, nrm);
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int j = 2 * h + jj;
            const uint32_t y = tj * 32 + j * 8 + (lane >> 3);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = 4 * j + q;
                const float z = pixel_view_z(fd, x0 + q, y, dep[4 * jj + q]);
                uint32_t k = 0u;
                if (z > 0.0f) {
                    const float e = fminf(fmaxf(__builtin_amdgcn_logf(z) * c1 + c0, 0.0f), emax);
                    const uint32_t k0 = (uint32_t)e;
                    const float vm = thr[k0 > 0u ? k0 - 1u : 0u], v0 = thr[k0], v1 = thr[k0 + 1u];
                    const float v2 = thr[min(k0 + 2u, kt)];
                    const bool dn = z < v0, up = z >= v1;
                    k = k0 - (dn ? 1u : 0u) + (up ? 1u : 0u);
                    const bool bad = !(z < thr_kt) || (dn && (k0 == 0u || z < vm)) || (up && z >= v2);
                    if (bad) pend |= 1u << i;
                }
                key[i] = assemble_key(fd, ti, tj, k, nrm[4 * jj + q]);
            }
        }
        // the second half's loads stay after the first half's keys (the register peak of one half)
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("" ::: "memory");
        }
        // the formula for the flagged pixels, their depth and normal read again (rare: beyond the table)
        if (pend) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                if ((pend >> i) & 1u) {
                    const uint32_t x = x0 + (i & 3), y = tj * 32 + (i >> 2) * 8 + (lane >> 3);
                    const uint64_t p = (uint64_t)(y % fd.H) * fd.W + (x % fd.W);
                    const float z = pixel_view_z(fd, x, y, fd.depth[p]);
                    key[i] = assemble_key(fd, ti, tj, slice_exact(fd, z), reinterpret_cast<const uint2*>(fd.normals)[p]);
                }
            }
        }
        // the tile's unique keys in ascending order, by repeated wave minima (one round per distinct key). Every
        // key of the tile has the tile's ids in its low 16 bits, so the keys compare by their high halves; a
        // pixel whose key has been emitted takes 0x10000 + the key's index in the list, above every high half
        // still pending, and keeps it as its index (no search afterwards)
        const uint32_t low = key[0] & 0xFFFFu;
#pragma unroll
        for (int q = 0; q < 16; q++) key[q] >>= 16;
        for (;;) {
            uint32_t mine = key[0];
#pragma unroll
            for (int q = 1; q < 16; q++) mine = min(mine, key[q]);
            const uint32_t m = wave_min_u32(mine);
            if (m > 0xFFFFu) break; // every key emitted (wave-uniform)
            if (lane == 0) uq[u] = (m << 16) | low;
#pragma unroll
            for (int q = 0; q < 16; q++) key[q] = key[q] == m ? 0x10000u + u : key[q];
            u++;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // the group's aggregate is published as soon as its tiles' unique counts are known, before this group's own
    // walk, so that the next groups' walks can sum it without waiting for the walk
    if (lane == 0) s_u[wave] = u;
    __syncthreads();
    if (wave == 0) {
        const uint32_t agg = wave_sum(lane < kClusterWaves ? s_u[lane] : 0u);
        if (lane == 0) wave_lookback_publish(tile_status + fd.tile_base / VREN_TILE_GROUP, t0 / VREN_TILE_GROUP, agg);
    }
    // decoupled look-back over the frame's tile groups (row-major order), one walk per workgroup: a round trip of
    // the walk covers 64 groups = 512 tiles
    if (wave == 0) {
        const uint32_t agg = wave_sum(lane < kClusterWaves ? s_u[lane] : 0u);
        const uint32_t prefix = wave_lookback_walk(tile_status + fd.tile_base / VREN_TILE_GROUP, t0 / VREN_TILE_GROUP, agg);
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    if (!live) return;
    uint32_t prefix = s_prefix;
    for (uint32_t w = 0; w < wave; w++) prefix += s_u[w];
    if (lane == 0) tile_cl[g] = make_uint2(prefix, u);
    for (uint32_t i = lane; i < u; i += VREN_WAVE) {
        const uint32_t idx = prefix + i;
        if (idx < fd.key_cap) fd.keys[idx] = uq[i];
    }
    if (lane == 0 && t == tiles - 1) {
        const uint32_t total = prefix + u;
        *fd.cluster_count = total;
        if (total > fd.key_cap) atomic_or_agent(fd.flags, VREN_FLAG_KEY_CAPACITY);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t y = tj * 32 + j * 8 + (lane >> 3);
        if (y >= fd.H) continue;
        uint32_t ref[4];
#pragma unroll
        for (int q = 0; q < 4; q++) ref[q] = (prefix - 0x10000u) + key[4 * j + q]; // key = 0x10000 + index
        uint32_t* row = fd.refs + (uint64_t)y * fd.W;
        if (x0 + 3 < fd.W && ((reinterpret_cast<uintptr_t>(row + x0) & 15) == 0)) {
            *reinterpret_cast<uint4*>(row + x0) = make_uint4(ref[0], ref[1], ref[2], ref[3]);
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (x0 + q < fd.W) row[x0 + q] = ref[q];
        }
    }
}

} // namespace

void chain_work_bind(ChainWork* w, void* mem, uint32_t total_tiles, uint32_t total_offs_tiles, uint32_t table_frames)
{
    w->tickets = static_cast<uint32_t*>(mem);
    w->tile_status = reinterpret_cast<uint64_t*>(static_cast<char*>(mem) + 256);
    w->offs_status = w->tile_status + total_tiles;
    w->tile_cl = reinterpret_cast<uint2*>(w->offs_status + total_offs_tiles);
    w->tile_range = w->tile_cl + total_tiles;
    w->slice_thr = table_frames ? reinterpret_cast<float*>(static_cast<char*>(mem) +
                                                           chain_work_table_offset(total_tiles, total_offs_tiles))
                                : nullptr;
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
    if (!w.slice_thr) return vren_fail(ctx, VREN_EINVAL, "clusters: chain workspace has no slice table");
    slice_table_kernel<<<F, 1024, 0, s>>>(d_frames, w.slice_thr);
    VREN_LAUNCH_CHECK(ctx);
    const uint32_t blocks = (w.total_tiles + kClusterWaves - 1) / kClusterWaves;
    clusters_kernel<<<blocks, kClusterWaves * 64, 0, s>>>(d_frames, F > 1 ? w.tile_frame : nullptr, w.total_tiles,
                                                          w.tile_status, w.tickets + 0, w.slice_thr, w.tile_cl);
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
    st = vren_ensure_scratch(ctx, chain_work_bytes(tiles, 0, 1), s);
    if (st) return st;
    chain_work_bind(&w, ctx->scratch, tiles, 0, 1);
    FrameDev* slot = vren_next_frame_slot(ctx);
    st = vren_upload_frames(ctx, s, slot, &f, 1);
    if (!st) st = launch_chain_init(ctx, s, w);
    if (!st) st = launch_clusters(ctx, s, slot, 1, w);
    return st;
}
