// chain.hpp -- internal launchers of the clustered-shading chain, shared by the per-stage entry points (one
// frame) and the batched chain (many frames in one launch per stage).
#pragma once

#include "clustered.hpp"

int vren_validate_cfg(vren_ctx* ctx, const vren_cluster_cfg* cfg);
int vren_validate_camera(vren_ctx* ctx, const vren_camera* cam);
int vren_upload_frames(vren_ctx* ctx, hipStream_t s, FrameDev* dst, const FrameDev* src, uint32_t n);
int vren_validate_material_set(vren_ctx* ctx, const vren_material_set* m, const char* what);
void vren_frame_set_materials(FrameDev* f, const vren_material_set* m);
FrameDev* vren_next_frame_slot(vren_ctx* ctx);

#define VREN_LEAF_ALIGN 256u
// a frame's 32x32 tiles start at a multiple of this in the flattened tile numbering of a table (the clusters kernel
// chains its look-back over groups of this many tiles, one workgroup each); tiles of the padding do nothing
#define VREN_TILE_GROUP 8u
#define VREN_SCAN_TILE 4096u

// look-back statuses and ticket counters of one chain launch
struct ChainWork {
    uint64_t* tile_status; // cluster compaction: one word per 32x32 tile of every frame
    uint64_t* offs_status; // light-offset scan: one word per 4096 clusters of capacity of every frame
    uint32_t* tickets;     // [0] cluster tiles, [1] offset tiles
    float* slice_thr;      // depth-slice threshold table of every frame (VREN_SLICE_TABLE floats each)
    uint2* tile_cl;        // per 32x32 tile: (first cluster index, cluster count), written by the clusters kernel
    uint2* tile_range;     // per 32x32 tile: [begin, end) of the index-buffer entries of its clusters' light lists
    const uint32_t* tile_frame; // frame of every flattened 32x32 tile (null for a one-frame table)
    uint32_t total_tiles;
    uint32_t total_offs_tiles;
    uint32_t total_leaf_slots;
    uint32_t max_L;
    uint32_t any_materials; // some frame of the table takes the material path (shade picks its kernel by it)
    // launch shape (vren_chain_shape; 0 = the default): shade strips per workgroup, persistent assignment waves
    // per CU, digit width of the light sort
    uint32_t shade_ppt = 0;
    uint32_t assign_wpc = 0;
    uint32_t light_bits = 0;
    uint32_t light_concurrent = 0; // batched runs: the light BVH on a second stream beside the cluster stage
    // optional per-kernel timing: an event is recorded after every kernel of the chain
    hipEvent_t* events = nullptr;
    int* n_events = nullptr;
};

// a failed record surfaces at the checked synchronisation on the last event (vren_batch_run_profiled)
inline void chain_mark(const ChainWork& w, hipStream_t s)
{
    if (w.events) vren_teardown(hipEventRecord(w.events[(*w.n_events)++], s));
}

// kernels of one chain launch, in order (chain_mark follows each)
#define VREN_CHAIN_KERNELS 9

// floats per frame of the depth-slice threshold table (clusters.hip)
#define VREN_SLICE_TABLE 1024u

// workspace: 256 bytes of tickets, the look-back statuses, the per-tile cluster spans and list ranges, then
// (256-aligned) the slice tables of table_frames frames
inline size_t chain_work_table_offset(uint32_t total_tiles, uint32_t total_offs_tiles)
{
    return (256 + ((size_t)total_tiles + total_offs_tiles) * sizeof(uint64_t) + (size_t)total_tiles * 16 + 255) &
           ~(size_t)255;
}
inline size_t chain_work_bytes(uint32_t total_tiles, uint32_t total_offs_tiles, uint32_t table_frames)
{
    return chain_work_table_offset(total_tiles, total_offs_tiles) + (size_t)table_frames * VREN_SLICE_TABLE * 4;
}
void chain_work_bind(ChainWork* w, void* mem, uint32_t total_tiles, uint32_t total_offs_tiles, uint32_t table_frames);

// stage launchers; d_frames is a device table of F FrameDev records
int launch_chain_init(vren_ctx* ctx, hipStream_t s, const ChainWork& w);
int launch_lights(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w);
// which 32-leaf groups the leaf kernel writes: all, those holding a light, or the all-padding ones
enum : uint32_t { kLeavesAll = 0, kLeavesLive = 1, kLeavesPadding = 2 };
int launch_bvh(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, uint32_t total_leaf_slots,
               bool leaves_from_lights, const ChainWork* w, uint32_t span);
int launch_clusters(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w);
int launch_assign(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w);
int launch_shade(vren_ctx* ctx, hipStream_t s, const FrameDev* d_frames, uint32_t F, const ChainWork& w);

// device: frame of a flattened 32x32 tile: one load from the tile -> frame table (uniform across the workgroup)
__device__ __forceinline__ uint32_t frame_of_tile(const uint32_t* tile_frame, uint32_t g)
{
    return tile_frame ? tile_frame[g] : 0u;
}

// device: frame of a flattened index, given a per-frame base field (bases ascending)
template <typename Get>
__device__ __forceinline__ uint32_t find_frame(const FrameDev* frames, uint32_t F, uint32_t g, Get base)
{
    uint32_t lo = 0, hi = F;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (base(frames[mid]) <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// device: the same search done by one whole wave at once (every lane calls it with the same g): lane i reads frame
// c + i's base, and a ballot gives the last frame whose base is <= g. One round of loads instead of log2(F)
// dependent ones.
template <typename Get>
__device__ __forceinline__ uint32_t find_frame_wave(const FrameDev* frames, uint32_t F, uint32_t g, Get base)
{
    const uint32_t lane = lane_id();
    uint32_t f = 0;
    for (uint32_t c = 0; c < F; c += VREN_WAVE) {
        const uint32_t i = c + lane;
        const uint64_t m = ballot(i < F && base(frames[i]) <= g);
        if (m == 0ull) break;
        f = c + 63u - (uint32_t)__builtin_clzll(m);
        if (m != ~0ull) break;
    }
    return f;
}
