// batch.hip -- cluster_and_shade (clustered_shading.cpp:817-1172) over a batch of frames.
//
// The reference records one frame's four stages into a render-graph node and submits it; its buffers are sized
// once at fixed maxima (1920x1080, 2^17 clusters, 2^23 assignments; clustered_shading.cpp:817-922). A batch
// here owns one device arena with every frame's intermediates sized for that frame, a device table of FrameDev
// records, and the look-back workspace. The chain is the same kernels for one frame or for all of them:
//   chain init -> lights -> BVH leaves + first level -> BVH upper levels -> clusters -> light counts ->
//   light offsets -> light lists -> shade
// VREN_RUN_PER_FRAME launches it once per frame (with that frame's one-entry table) and synchronises the stream
// after each frame; VREN_RUN_BATCHED launches it once for all frames; VREN_RUN_GRAPH captures the batched launches
// into a hipGraph on first use and replays it.
#include <cstring>
#include <vector>

#include "chain.hpp"

struct vren_batch {
    vren_ctx* ctx = nullptr;
    uint32_t F = 0;
    std::vector<FrameDev> host;     // batch table (flattened bases)
    FrameDev* d_table = nullptr;    // F records
    FrameDev* d_single = nullptr;   // F one-entry tables (bases 0)
    char* arena = nullptr;
    size_t arena_bytes = 0;
    void* work = nullptr;           // ChainWork memory
    ChainWork w;                    // batch-wide
    std::vector<ChainWork> w_frame; // per-frame views of the same memory
    std::vector<uint32_t> leaf_slots;
    std::vector<uint32_t> live_slots; // per frame: leaf slots of the groups that hold lights (256-aligned)
    uint32_t total_live_slots = 0;
    uint32_t* d_tile_frame = nullptr; // frame of every flattened tile
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture = nullptr;
    uint64_t size_class = 0;        // chain autotuner key (batch_size_class)
    vren_chain_shape shape{};       // the chain's launch shape (zero fields: defaults)
    // the second stream of a concurrent light stage, with its fork and join events (created with the batch)
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

namespace {

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

uint32_t ilog2_floor(uint64_t v)
{
    uint32_t l = 0;
    while (v > 1) {
        v >>= 1;
        l++;
    }
    return l;
}

// the chain autotuner's size class: log2 of the batch's pixels, of its largest light count and of its frame count
uint64_t batch_size_class(uint64_t pixels, uint32_t max_L, uint32_t frames)
{
    return ((uint64_t)ilog2_floor(pixels) << 16) | ((uint64_t)ilog2_floor((uint64_t)max_L + 1) << 8) |
           ilog2_floor(frames);
}

bool valid_shape(const vren_chain_shape& s)
{
    return (s.shade_pixels_per_thread == 0 || s.shade_pixels_per_thread == 1 || s.shade_pixels_per_thread == 2) &&
           (s.assign_waves_per_cu == 0 || s.assign_waves_per_cu == 8 || s.assign_waves_per_cu == 16 ||
            s.assign_waves_per_cu == 24 || s.assign_waves_per_cu == 32) &&
           (s.light_sort_bits == 0 || s.light_sort_bits == 4 || s.light_sort_bits == 8) &&
           (s.light_stage_concurrent == 0 || s.light_stage_concurrent == 1);
}

void apply_shape(ChainWork& w, const vren_chain_shape& s, bool batched)
{
    w.shade_ppt = s.shade_pixels_per_thread;
    w.assign_wpc = s.assign_waves_per_cu;
    w.light_bits = s.light_sort_bits;
    w.light_concurrent = batched ? s.light_stage_concurrent : 0u; // the per-frame path stays one stream
}

// a new shape invalidates the captured graph (its kernels and grids were captured with the old one)
void set_batch_shape(vren_batch* b, const vren_chain_shape& s)
{
    b->shape = s;
    apply_shape(b->w, s, true);
    for (ChainWork& wf : b->w_frame) apply_shape(wf, s, false);
    if (b->exec) vren_teardown(hipGraphExecDestroy(b->exec));
    if (b->graph) vren_teardown(hipGraphDestroy(b->graph));
    b->exec = nullptr;
    b->graph = nullptr;
}

// live_slots: the leaf-kernel grid over the frames' light-holding leaf groups (their padding groups were written
// when the batch was created). With a concurrent light stage (and no per-kernel timing), the lights and their BVH
// run on the batch's second stream between a fork and a join while `s` keys the clusters: the two stages share no
// buffer, and light assignment, the first consumer of both, waits for the join. The join is recorded even when a
// light-stage launch fails, so a capture always ends with both streams joined.
int run_chain(vren_batch* b, hipStream_t s, const FrameDev* table, uint32_t F, const ChainWork& w,
              uint32_t live_slots)
{
    vren_ctx* ctx = b->ctx;
    int st = launch_chain_init(ctx, s, w);
    if (!st && w.light_concurrent && !w.events && b->side) {
        VREN_HIP(ctx, hipEventRecord(b->fork, s));
        VREN_HIP(ctx, hipStreamWaitEvent(b->side, b->fork, 0));
        int sl = launch_lights(ctx, b->side, table, F, w);
        if (!sl && live_slots) sl = launch_bvh(ctx, b->side, table, F, live_slots, true, &w, kLeavesLive);
        VREN_HIP(ctx, hipEventRecord(b->join, b->side));
        st = launch_clusters(ctx, s, table, F, w);
        VREN_HIP(ctx, hipStreamWaitEvent(s, b->join, 0));
        if (!st) st = sl;
    } else {
        if (!st) st = launch_lights(ctx, s, table, F, w);
        if (!st && live_slots) st = launch_bvh(ctx, s, table, F, live_slots, true, &w, kLeavesLive);
        if (!st) st = launch_clusters(ctx, s, table, F, w);
    }
    if (!st) st = launch_assign(ctx, s, table, F, w);
    if (!st) st = launch_shade(ctx, s, table, F, w);
    return st;
}

} // namespace

extern "C" {

VREN_EXPORT int vren_batch_create(vren_ctx* ctx, const vren_frame_desc* frames, uint32_t n_frames,
                                  const vren_cluster_cfg* cfg, vren_batch** out)
{
    if (!ctx || !out) return VREN_EINVAL;
    *out = nullptr;
    int st = vren_validate_cfg(ctx, cfg);
    if (st) return st;
    if (!frames || n_frames == 0 || n_frames > 4096) return vren_fail(ctx, VREN_EINVAL, "batch: 1..4096 frames");
    for (uint32_t i = 0; i < n_frames; i++) {
        const vren_frame_desc& d = frames[i];
        st = vren_validate_camera(ctx, &d.camera);
        if (st) return st;
        if (d.width == 0 || d.height == 0 || d.width > 8192 || d.height > 8192)
            return vren_fail(ctx, VREN_EINVAL, "batch: frame %u screen must be in [1, 8192]^2", i);
        if (!d.depth || !d.normals || (!d.materials && (!d.albedo_roughness || !d.metallic)) || !d.output)
            return vren_fail(ctx, VREN_EINVAL, "batch: frame %u has a null G-buffer or output", i);
        if (d.materials) {
            st = vren_validate_material_set(ctx, d.materials, "batch");
            if (st) return st;
        }
        if (d.light_count > 0 && (!d.light_positions || !d.point_lights))
            return vren_fail(ctx, VREN_EINVAL, "batch: frame %u has null light buffers", i);
        if (d.light_count > (1u << 20)) return vren_fail(ctx, VREN_EINVAL, "batch: frame %u has > 2^20 lights", i);
        if (d.directional_light_count > 0 && !d.directional_lights)
            return vren_fail(ctx, VREN_EINVAL, "batch: frame %u has null directional lights", i);
    }
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "batch: context has no device");
    vren_batch* b = new vren_batch();
    b->ctx = ctx;
    b->F = n_frames;
    b->host.resize(n_frames);
    b->leaf_slots.resize(n_frames);
    b->live_slots.resize(n_frames);
    // sizes and arena layout
    uint32_t tile_base = 0, leaf_base = 0, live_base = 0, offs_base = 0, max_L = 0, any_materials = 0;
    std::vector<size_t> off(n_frames * 14);
    size_t arena = 0;
    for (uint32_t i = 0; i < n_frames; i++) {
        const vren_frame_desc& d = frames[i];
        FrameDev& f = b->host[i];
        memset(&f, 0, sizeof(f));
        vren_frame_setup(&f, d.width, d.height, &d.camera, cfg);
        const uint32_t L = d.light_count;
        const uint32_t tiles = f.Sx * f.Sy;
        uint32_t key_cap = d.key_capacity;
        if (key_cap == 0) key_cap = (uint32_t)std::min<uint64_t>((uint64_t)tiles * 1024, 1u << 20);
        uint32_t index_cap = d.index_capacity;
        if (index_cap == 0) index_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((uint64_t)key_cap * L, 1), 1u << 24);
        f.tile_base = tile_base;
        f.leaf_base = leaf_base;
        f.live_leaf_base = live_base;
        f.offs_base = offs_base;
        f.L = L;
        f.light_pos = d.light_positions;
        f.lights = d.point_lights;
        f.n_dir = d.directional_light_count;
        f.dir_lights = d.directional_lights;
        f.depth = d.depth;
        f.normals = d.normals;
        f.albedo_rough = d.albedo_roughness;
        f.metallic = d.metallic;
        if (d.materials) {
            vren_frame_set_materials(&f, d.materials);
            any_materials = 1;
        }
        f.out = d.output;
        f.key_cap = key_cap;
        f.index_cap = index_cap;
        f.flags = ctx->d_flags;
        f.bvh_padded = L ? vren_calc_bvh_padded_leaf_count(L) : 0;
        f.bvh_root = L ? vren_calc_bvh_root_index(L) : 0;
        f.bvh_levels = L ? vren_calc_bvh_level_count(L) : 0;
        f.bvh_length = L ? vren_calc_bvh_buffer_length(L) : 0;
        const uint32_t slots = L ? (f.bvh_padded + VREN_LEAF_ALIGN - 1) / VREN_LEAF_ALIGN * VREN_LEAF_ALIGN : 0;
        b->leaf_slots[i] = slots;
        const uint32_t live = (std::min(slots, (L + 31u) & ~31u) + VREN_LEAF_ALIGN - 1) / VREN_LEAF_ALIGN * VREN_LEAF_ALIGN;
        b->live_slots[i] = live;
        tile_base += (tiles + VREN_TILE_GROUP - 1) / VREN_TILE_GROUP * VREN_TILE_GROUP;
        leaf_base += slots;
        live_base += live;
        offs_base += (key_cap + VREN_SCAN_TILE - 1) / VREN_SCAN_TILE;
        max_L = std::max(max_L, L);
        // list slots for the first 32 768 clusters of the frame (768 B each); clusters past them are traversed again
        const uint32_t slot_cap = std::min<uint32_t>(key_cap, 1u << 15);
        f.slot_cap = L ? slot_cap : 0u;
        size_t* o = &off[i * 14];
        o[0] = arena; arena += a256((size_t)L * 16);                 // view_pos
        o[1] = arena; arena += a256((size_t)L * 8);                  // codes
        o[2] = arena; arena += a256(L > 4096 ? (size_t)L * 8 : 0);   // codes_tmp
        o[3] = arena; arena += a256((size_t)f.bvh_length * 32);      // bvh
        o[4] = arena; arena += a256((size_t)key_cap * 4);            // keys
        o[5] = arena; arena += 256;                                  // cluster_count
        o[6] = arena; arena += a256((size_t)d.width * d.height * 4); // refs
        o[7] = arena; arena += a256((size_t)key_cap * 4);            // counts
        o[8] = arena; arena += a256((size_t)key_cap * 4);            // offsets
        o[9] = arena; arena += a256((size_t)index_cap * 4);          // indices
        o[10] = arena; arena += 256;                                 // assigned_total
        o[11] = arena; arena += a256((size_t)L * 16);                // leaf_pos
        o[12] = arena; arena += a256((size_t)f.slot_cap * VREN_LIST_SLOT * 4); // list slots
    }
    hipError_t e = hipMalloc(&b->arena, std::max<size_t>(arena, 256));
    if (e == hipSuccess) e = hipMemset(b->arena, 0, std::max<size_t>(arena, 256));
    const size_t wbytes = chain_work_bytes(tile_base, offs_base, n_frames);
    if (e == hipSuccess) e = hipMalloc(&b->work, wbytes);
    if (e == hipSuccess) e = hipMalloc(&b->d_table, sizeof(FrameDev) * n_frames);
    if (e == hipSuccess) e = hipMalloc(&b->d_single, sizeof(FrameDev) * n_frames);
    if (e == hipSuccess) e = hipMalloc(&b->d_tile_frame, sizeof(uint32_t) * std::max<uint32_t>(tile_base, 1));
    if (e == hipSuccess) {
        std::vector<uint32_t> tf(tile_base);
        for (uint32_t i = 0; i < n_frames; i++) {
            const uint32_t end = i + 1 < n_frames ? b->host[i + 1].tile_base : tile_base;
            for (uint32_t t = b->host[i].tile_base; t < end; t++) tf[t] = i; // the frame's tiles and its padding
        }
        e = hipMemcpy(b->d_tile_frame, tf.data(), sizeof(uint32_t) * tile_base, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        vren_batch_destroy(b);
        return vren_hip_fail(ctx, e, "batch allocation");
    }
    b->arena_bytes = arena;
    for (uint32_t i = 0; i < n_frames; i++) {
        FrameDev& f = b->host[i];
        size_t* o = &off[i * 14];
        f.view_pos = reinterpret_cast<float*>(b->arena + o[0]);
        f.codes = reinterpret_cast<uint32_t*>(b->arena + o[1]);
        f.codes_tmp = f.L > 4096 ? reinterpret_cast<uint32_t*>(b->arena + o[2]) : nullptr;
        f.bvh = reinterpret_cast<vren_bvh_node*>(b->arena + o[3]);
        f.keys = reinterpret_cast<uint32_t*>(b->arena + o[4]);
        f.cluster_count = reinterpret_cast<uint32_t*>(b->arena + o[5]);
        f.refs = reinterpret_cast<uint32_t*>(b->arena + o[6]);
        f.counts = reinterpret_cast<uint32_t*>(b->arena + o[7]);
        f.offsets = reinterpret_cast<uint32_t*>(b->arena + o[8]);
        f.indices = reinterpret_cast<uint32_t*>(b->arena + o[9]);
        f.assigned_total = reinterpret_cast<uint32_t*>(b->arena + o[10]);
        f.leaf_pos = f.L ? reinterpret_cast<float*>(b->arena + o[11]) : nullptr;
        f.list_slots = f.slot_cap ? reinterpret_cast<uint32_t*>(b->arena + o[12]) : nullptr;
    }
    chain_work_bind(&b->w, b->work, tile_base, offs_base, n_frames);
    b->w.tile_frame = b->d_tile_frame;
    b->w.total_leaf_slots = leaf_base;
    b->total_live_slots = live_base;
    b->w.max_L = max_L;
    b->w.any_materials = any_materials;
    {
        uint64_t pixels = 0;
        for (uint32_t i = 0; i < n_frames; i++) pixels += (uint64_t)frames[i].width * frames[i].height;
        b->size_class = batch_size_class(pixels, max_L, n_frames);
    }
    // per-frame tables: same buffers, bases 0, the workspace viewed from its start
    std::vector<FrameDev> single(b->host);
    b->w_frame.resize(n_frames);
    for (uint32_t i = 0; i < n_frames; i++) {
        single[i].tile_base = 0;
        single[i].leaf_base = 0;
        single[i].live_leaf_base = 0;
        single[i].offs_base = 0;
        ChainWork& wf = b->w_frame[i];
        chain_work_bind(&wf, b->work, single[i].Sx * single[i].Sy, (single[i].key_cap + VREN_SCAN_TILE - 1) / VREN_SCAN_TILE,
                        1);
        wf.total_leaf_slots = b->leaf_slots[i];
        wf.max_L = single[i].L;
        wf.any_materials = single[i].texcoords != nullptr;
    }
    {
        // the shape the context tuned for this size class, if any
        auto it = ctx->chain_tuned.find(b->size_class);
        set_batch_shape(b, it != ctx->chain_tuned.end() ? it->second : vren_chain_shape{});
    }
    hipStream_t s = nullptr;
    // the second stream and events of a concurrent light stage (default priority, like every stream here)
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
        st = vren_fail(ctx, VREN_EHIP, "batch: second stream or its events could not be created");
    if (!st) st = vren_upload_frames(ctx, s, b->d_table, b->host.data(), n_frames);
    if (!st) st = vren_upload_frames(ctx, s, b->d_single, single.data(), n_frames);
    // the BVH leaf groups past each frame's last light, written once: every run of the chain skips them
    if (!st && leaf_base) st = launch_bvh(ctx, s, b->d_table, n_frames, leaf_base, true, nullptr, kLeavesPadding);
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = vren_fail(ctx, VREN_EHIP, "batch: upload failed");
    if (st) {
        vren_batch_destroy(b);
        return st;
    }
    *out = b;
    return VREN_OK;
}

VREN_EXPORT int vren_batch_run(vren_batch* b, vren_stream_t stream, int mode)
{
    if (!b) return VREN_EINVAL;
    vren_ctx* ctx = b->ctx;
    hipStream_t s = (hipStream_t)stream;
    if (mode == VREN_RUN_PER_FRAME) {
        for (uint32_t i = 0; i < b->F; i++) {
            int st = run_chain(b, s, b->d_single + i, 1, b->w_frame[i], b->live_slots[i]);
            if (st) return st;
            VREN_HIP(ctx, hipStreamSynchronize(s));
        }
        return VREN_OK;
    }
    if (mode == VREN_RUN_BATCHED) return run_chain(b, s, b->d_table, b->F, b->w, b->total_live_slots);
    if (mode == VREN_RUN_GRAPH) {
        if (!b->exec) {
            // the capture stream lives as long as the batch (created once, destroyed by vren_batch_destroy); a
            // failed capture or instantiation leaves no graph behind, so the next call starts over
            if (!b->capture) VREN_HIP(ctx, hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
            VREN_HIP(ctx, hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
            int st = run_chain(b, b->capture, b->d_table, b->F, b->w, b->total_live_slots);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->capture, &g);
            if (st || e != hipSuccess) {
                if (g) vren_teardown(hipGraphDestroy(g));
                return st ? st : vren_hip_fail(ctx, e, "hipStreamEndCapture");
            }
            e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                b->exec = nullptr;
                vren_teardown(hipGraphDestroy(g));
                return vren_hip_fail(ctx, e, "hipGraphInstantiate");
            }
            b->graph = g;
        }
        VREN_HIP(ctx, hipGraphLaunch(b->exec, s));
        return VREN_OK;
    }
    return vren_fail(ctx, VREN_EINVAL, "batch_run: unknown mode %d", mode);
}

VREN_EXPORT int vren_batch_run_profiled(vren_batch* b, vren_stream_t stream, float* kernel_ms)
{
    if (!b || !kernel_ms) return VREN_EINVAL;
    vren_ctx* ctx = b->ctx;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t ev[VREN_CHAIN_KERNELS + 1];
    for (int i = 0; i <= VREN_CHAIN_KERNELS; i++) VREN_HIP(ctx, hipEventCreate(&ev[i]));
    int n = 0;
    ChainWork w = b->w;
    w.events = ev + 1;
    w.n_events = &n;
    VREN_HIP(ctx, hipEventRecord(ev[0], s));
    int st = run_chain(b, s, b->d_table, b->F, w, b->total_live_slots);
    if (!st) VREN_HIP(ctx, hipEventSynchronize(ev[n]));
    for (int i = 0; i < VREN_CHAIN_KERNELS; i++) {
        kernel_ms[i] = 0.0f;
        if (!st && i < n && hipEventElapsedTime(&kernel_ms[i], ev[i], ev[i + 1]) != hipSuccess)
            st = vren_fail(ctx, VREN_EHIP, "batch_run_profiled: no time between events %d and %d", i, i + 1);
    }
    for (int i = 0; i <= VREN_CHAIN_KERNELS; i++) vren_teardown(hipEventDestroy(ev[i]));
    return st;
}

VREN_EXPORT int vren_batch_frame_outputs(vren_batch* b, uint32_t frame, vren_frame_outputs* out)
{
    if (!b || !out || frame >= b->F) return VREN_EINVAL;
    const FrameDev& f = b->host[frame];
    out->view_positions = f.view_pos;
    out->light_codes = f.codes;
    out->bvh = f.bvh;
    out->bvh_length = f.bvh_length;
    out->cluster_keys = f.keys;
    out->cluster_count = f.cluster_count;
    out->cluster_refs = f.refs;
    out->key_capacity = f.key_cap;
    out->light_counts = f.counts;
    out->light_offsets = f.offsets;
    out->light_indices = f.indices;
    out->index_capacity = f.index_cap;
    out->assigned_total = f.assigned_total;
    return VREN_OK;
}

VREN_EXPORT int vren_batch_get_shape(vren_batch* b, vren_chain_shape* shape)
{
    if (!b || !shape) return VREN_EINVAL;
    *shape = b->shape;
    return VREN_OK;
}

VREN_EXPORT int vren_batch_set_shape(vren_batch* b, const vren_chain_shape* shape)
{
    if (!b || !shape) return VREN_EINVAL;
    if (!valid_shape(*shape))
        return vren_fail(b->ctx, VREN_EINVAL, "batch_set_shape: shade 1|2, assign waves 8|16|24|32, light bits 4|8");
    set_batch_shape(b, *shape);
    return VREN_OK;
}

VREN_EXPORT int vren_batch_autotune(vren_batch* b, vren_stream_t stream, uint32_t reps, vren_chain_shape* best,
                                    float* best_ms)
{
    if (!b) return VREN_EINVAL;
    vren_ctx* ctx = b->ctx;
    if (reps == 0 || reps > 100) return vren_fail(ctx, VREN_EINVAL, "batch_autotune: reps must be in [1, 100]");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) vren_teardown(hipEventDestroy(e0));
        return vren_fail(ctx, VREN_EHIP, "batch_autotune: event creation failed");
    }
    const vren_chain_shape before = b->shape;
    vren_chain_shape top{};
    float top_ms = -1.0f;
    int st = VREN_OK;
    const uint32_t ppts[2] = {1, 2}, wpcs[4] = {8, 16, 24, 32}, bits[2] = {4, 8};
    // every shape of the grid: {1, 2} pixels per thread x {8, 16, 24, 32} waves per CU x {4, 8} bits x light stage
    // in line or concurrent, as index i = ((a * 4 + c) * 2 + d) * 2 + e
    for (uint32_t i = 0; i < 2 * 4 * 2 * 2 && !st; i++) {
        const uint32_t a = i / 16, c = (i / 4) % 4, d = (i / 2) % 2, e = i % 2;
        const vren_chain_shape cand{ppts[a], wpcs[c], bits[d], e};
        set_batch_shape(b, cand);
        st = run_chain(b, s, b->d_table, b->F, b->w, b->total_live_slots); // warm-up
        if (!st && hipEventRecord(e0, s) != hipSuccess) st = vren_fail(ctx, VREN_EHIP, "batch_autotune: record");
        for (uint32_t r = 0; r < reps && !st; r++) st = run_chain(b, s, b->d_table, b->F, b->w, b->total_live_slots);
        float ms = 0.0f;
        if (!st && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            st = vren_fail(ctx, VREN_EHIP, "batch_autotune: timing a shape failed");
        ms /= (float)reps;
        if (!st && (top_ms < 0.0f || ms < top_ms)) {
            top = cand;
            top_ms = ms;
        }
    }
    vren_teardown(hipEventDestroy(e0));
    vren_teardown(hipEventDestroy(e1));
    if (st) {
        set_batch_shape(b, before);
        return st;
    }
    set_batch_shape(b, top);
    ctx->chain_tuned[b->size_class] = top;
    if (best) *best = top;
    if (best_ms) *best_ms = top_ms;
    return VREN_OK;
}

VREN_EXPORT void vren_batch_destroy(vren_batch* b)
{
    if (!b) return;
    vren_teardown(hipDeviceSynchronize());
    if (b->exec) vren_teardown(hipGraphExecDestroy(b->exec));
    if (b->graph) vren_teardown(hipGraphDestroy(b->graph));
    if (b->capture) vren_teardown(hipStreamDestroy(b->capture));
    if (b->side) vren_teardown(hipStreamDestroy(b->side));
    if (b->fork) vren_teardown(hipEventDestroy(b->fork));
    if (b->join) vren_teardown(hipEventDestroy(b->join));
    if (b->arena) vren_teardown(hipFree(b->arena));
    if (b->work) vren_teardown(hipFree(b->work));
    if (b->d_table) vren_teardown(hipFree(b->d_table));
    if (b->d_single) vren_teardown(hipFree(b->d_single));
    if (b->d_tile_frame) vren_teardown(hipFree(b->d_tile_frame));
    delete b;
}

} // extern "C"
