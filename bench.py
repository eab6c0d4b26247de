#!/usr/bin/env python3
"""Benchmark of the clustered-shading compute path on one MI355X.

Workload (the metric's configuration): batches of 64 independent 1920x1080 frames, each with its own procedural
scene (64 boxes + ground plane, depth/normals ray-cast) and its own 4096 point lights (positions U[-50,50]^3,
radius U[1,8], colour U[0,1]^3, intensity logU[0.1,100]), per-pixel materials; camera at (0,0,-50), 90 deg fov,
near 0.1, far 100. All inputs are synthetic and generated on the device before timing; they stay resident in HBM.

One step = the whole clustered-shading chain (light BVH, unique clusters, light assignment, shade) over one
batch of 64 frames:
  - batched: the chain launched once for all 64 frames, replayed from a captured hipGraph  -> `value`
  - plain:   each frame's chain launched separately and followed by hipStreamSynchronize (the reference's
             record -> vkQueueSubmit -> fence wait per frame)                                 -> `plain_frames_per_s`
Both produce identical bytes (tests/test_clustered_gpu.py). Also measured: the 2^26-pair onesweep radix sort
(`sort`), per-kernel durations of the batched chain with HIP events, and the CPU restatement on the host cores
(`cpu_baseline`).
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "clustered-shade frames/s (1080p, 4096 lights) batched vs plain; sort Gkeys/s %HBM"
HBM_PEAK_GBS = 8000.0
KERNELS = ["chain_init", "lights", "bvh_leaf", "bvh_upper", "clusters", "assign_count", "assign_offsets",
           "assign_write", "shade"]
W, H, L_LIGHTS = 1920, 1080, 4096
SEED_FRAMES = 42 + 1000 * 4
# algorithmic bytes per unit of the chain's kernels (per pixel / per light of one frame)
BYTES_PER_PIXEL = {"clusters": 16.0, "shade": 52.0}
BYTES_PER_LIGHT = {"lights": 40.0}
SORT_BYTES_PER_PAIR = 68.0
# FP32 operations of one (pixel, assigned light) pair of the PBR shade (distance, window, GGX NDF, Smith G,
# Schlick F, diffuse/specular combine, accumulate), against the MI355X vector FP32 peak
SHADE_OPS_PER_PAIR = 102.0
FP32_PEAK_TFLOPS = 157.3


def profiled_traffic(label="traffic-bytes-per-launch"):
    """HBM bytes per launch of each kernel (label "traffic-bytes-per-launch") or per call of each row's entry point
    (label "row-traffic-bytes-per-call"), from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/hbm_traffic_pmc.md; counters scaled by a 1 GiB copy-kernel calibration)."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_pmc.md")) as f:
            text = f.read()
    except OSError:
        return {}
    m = re.search(label + r": (.*)", text)
    if not m:
        return {}
    out = {}
    for item in m.group(1).split(";"):
        if "=" in item:
            k, v = item.strip().rsplit("=", 1)
            out[k] = float(v)
    return out


def make_camera(vren_amd):
    import numpy as np
    c = vren_amd.Camera()
    c.position[0], c.position[1], c.position[2] = 0.0, 0.0, -50.0
    c.fov_y = float(np.float32(np.deg2rad(90.0)))
    c.aspect_ratio = float(np.float32(W) / np.float32(H))
    c.near_plane, c.far_plane, c.yaw, c.pitch = 0.1, 100.0, 0.0, 0.0
    return c


def build_batch(ctx, vren_amd, n_frames):
    """The bench's frames and a Batch over them (reference key encoder); `keep` holds each frame's tensors and,
    last, the list of frame descriptors (for another Batch over the same frames)."""
    import torch
    cam = make_camera(vren_amd)
    keep, descs = [], []
    for f in range(n_frames):
        seed = SEED_FRAMES + f
        d = torch.empty(W * H, dtype=torch.float32, device="cuda")
        n = torch.empty(W * H * 4, dtype=torch.int16, device="cuda")
        a = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
        m = torch.empty(W * H, dtype=torch.float32, device="cuda")
        p = torch.empty(L_LIGHTS * 4, dtype=torch.float32, device="cuda")
        l = torch.empty(L_LIGHTS * 4, dtype=torch.float32, device="cuda")
        o = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
        ctx.synth_frame(seed, W, H, cam, 64, d, n)
        ctx.synth_materials(seed, W * H, a, m)
        ctx.synth_lights(seed + 7, L_LIGHTS, p, l)
        fd = vren_amd.FrameDesc()
        fd.width, fd.height, fd.camera = W, H, cam
        fd.depth, fd.normals, fd.albedo_roughness, fd.metallic = d.data_ptr(), n.data_ptr(), a.data_ptr(), m.data_ptr()
        fd.light_positions, fd.point_lights, fd.light_count = p.data_ptr(), l.data_ptr(), L_LIGHTS
        fd.directional_lights, fd.directional_light_count, fd.output = None, 0, o.data_ptr()
        fd.key_capacity, fd.index_capacity = 0, 0
        descs.append(fd)
        keep.append((d, n, a, m, p, l, o))
    torch.cuda.synchronize()
    keep.append(descs)
    return vren_amd.Batch(ctx, descs, vren_amd.default_cfg()), keep


def _from_device(ptr, nbytes, dtype):
    """Copy nbytes from a raw device pointer to the host (hipMemcpy through the HIP runtime torch loaded)."""
    import numpy as np
    import torch
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.zeros(max(nbytes, 1), np.uint8)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes),
                       ctypes.c_int(2))
    if rc != 0:
        raise RuntimeError("hipMemcpy device->host failed (%d)" % rc)
    return out[:nbytes].view(dtype)


def shade_pairs(ctx, vren_amd, batch, keep):
    """(pixel, assigned light) pairs the shade kernel evaluates over the batch: for every pixel in front of the
    background threshold (view depth < far - 0.001, the kernel's test), the length of its cluster's light list.
    Read back from the batch's outputs after a run."""
    import numpy as np
    cam = make_camera(vren_amd)
    _, proj = ctx.camera_matrices(cam)
    inv = np.linalg.inv(proj.astype(np.float64).reshape(4, 4).T)  # stored column-major
    bg = float(np.float32(cam.far_plane) - np.float32(0.001))
    total = 0
    for f, k in enumerate(keep[:batch.n]):
        o = batch.outputs(f)
        d = k[0].cpu().numpy().astype(np.float64)
        z = (inv[2, 2] * d + inv[2, 3]) / (inv[3, 2] * d + inv[3, 3])
        refs = _from_device(o.cluster_refs, 4 * W * H, np.uint32)
        counts = _from_device(o.light_counts, 4 * o.key_capacity, np.uint32)
        ok = (z < bg) & (refs < o.key_capacity)
        total += int(counts[refs[ok]].sum(dtype=np.int64))
    return total


def time_events(fn, steps):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1), wall * 1e3


def bench_sort(ctx, log2n, steps, warmup, tune):
    import torch
    n = 1 << log2n
    keys0 = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.synth_keys(42, 54, n, keys0)  # the sort's key generator: pcg32(initstate 42, initseq 54)
    vals0 = torch.arange(n, dtype=torch.int32, device="cuda")
    keys, vals = torch.empty_like(keys0), torch.empty_like(vals0)
    scratch = torch.empty(ctx.radix_sort_scratch_bytes(n, True), dtype=torch.uint8, device="cuda")
    cfg_id, tuned_ms = (None, None)
    if tune:
        cfg_id, tuned_ms = ctx.autotune_radix_sort(n, True, 32)
    s = torch.cuda.current_stream()
    times = []
    for r in range(warmup + steps):
        keys.copy_(keys0)
        vals.copy_(vals0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        ctx.radix_sort(keys, vals, n, scratch)
        e1.record(s)
        torch.cuda.synchronize()
        if r >= warmup:
            times.append(e0.elapsed_time(e1))
    # correctness spot check on the device: sorted
    ku = keys.to(torch.int64) & 0xFFFFFFFF
    assert bool((ku[1:] >= ku[:-1]).all())
    ms = sorted(times)[len(times) // 2]
    achieved = SORT_BYTES_PER_PAIR * n / (ms * 1e-3) / 1e9
    configs = ctx.radix_configs()
    cfg = configs[cfg_id] if cfg_id is not None else (8, 23, 16, 0)  # the library's untuned default
    bits, kpt, waves, algo = cfg
    passes = (32 + bits - 1) // bits
    traffic = profiled_traffic()
    sort_traffic = None
    if log2n == 26:
        if algo == 0:  # histogram of every pass + onesweep passes
            parts = [traffic.get("rs_hist_kernel<%d, false>" % bits)] + \
                    [traffic.get("rs_onesweep_kernel<%d, %d, %d, true, false, false, false, false>"
                                 % (bits, kpt, waves))] * passes
        else:          # per pass: tile counts, scan, scatter
            parts = [traffic.get("rs_upsweep_kernel<%d, %d, %d>" % (bits, kpt, waves)), traffic.get("rs_tile_scan"),
                     traffic.get("rs_onesweep_kernel<%d, %d, %d, true, true, false, false, false>"
                                 % (bits, kpt, waves))] * passes
        if all(x is not None for x in parts):
            sort_traffic = sum(parts)
    return {
        "n": n, "pairs": True, "ms": round(ms, 4), "gpairs_per_s": round(n / (ms * 1e-3) / 1e9, 3),
        "config": list(cfg),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": sort_traffic,
                     "algorithmic_bytes_per_pair": SORT_BYTES_PER_PAIR},
    }


def _median_ms(fn, reps, warmup=2, before=None):
    """Median duration of fn() over reps runs, HIP events on the current stream around each run; before() (not
    timed) runs ahead of each."""
    import torch
    s = torch.cuda.current_stream()
    times = []
    for r in range(warmup + reps):
        if before:
            before()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fn()
        e1.record(s)
        torch.cuda.synchronize()
        if r >= warmup:
            times.append(e0.elapsed_time(e1))
    return sorted(times)[len(times) // 2]


def _hbm(ms, nbytes, **extra):
    gbs = nbytes / (ms * 1e-3) / 1e9
    d = {"ms": round(ms, 4), "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": nbytes}
    d.update(extra)
    return d


def bench_primitives(ctx, vren_amd, reps, log2n=26, bucket_log2=24):
    """The path's other rows at roofline sizes: uint32 add-reduce (the up-sweep tree and the scalar total) and
    exclusive scan of 2^26 keys, bucket sort of 2^24 pairs into 65536 buckets, the point-light animation of 2^24
    lights, and the shade forward and adjoint on one 1920x1080 frame with 4096 lights (bench_shade_frame)."""
    import numpy as np
    import torch
    n = 1 << log2n
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.synth_keys(42, 54, n, keys)
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    one = torch.empty(1, dtype=torch.int32, device="cuda")
    res = {}
    ms = _median_ms(lambda: ctx.reduce_u32(keys, n, out, vren_amd.REDUCE_ADD), reps)
    res["reduce_tree_u32"] = _hbm(ms, 8 * n, n=n, algorithmic_bytes_per_key=8)
    ms = _median_ms(lambda: ctx.reduce_total_u32(keys, n, one, vren_amd.REDUCE_ADD), reps)
    res["reduce_total_u32"] = _hbm(ms, 4 * n, n=n, algorithmic_bytes_per_key=4)
    ms = _median_ms(lambda: ctx.exclusive_scan_u32(keys, out, n), reps)
    res["exclusive_scan_u32"] = _hbm(ms, 8 * n, n=n, algorithmic_bytes_per_key=8)
    del out, one
    nb = 1 << bucket_log2
    pairs_in = torch.stack([keys[:nb], torch.arange(nb, dtype=torch.int32, device="cuda")], dim=1).contiguous()
    pairs_out = torch.empty_like(pairs_in)
    offs = torch.empty(65536, dtype=torch.int32, device="cuda")
    ms = _median_ms(lambda: ctx.bucket_sort(pairs_in, nb, pairs_out, offs, 65536, 0), reps)
    # counting-sort unit: count read 8 B + scatter read 8 B + write 8 B per pair (pairs in and out alone: 16 B)
    res["bucket_sort_65536"] = _hbm(ms, 24 * nb, n=nb, algorithmic_bytes_per_pair=24)
    del pairs_in, pairs_out, offs
    # blelloch_scan::downsweep of the up-sweep tree (radix_sort.cpp:281-289's use): one read and one write per key
    tree = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.reduce_u32(keys, n, tree, vren_amd.REDUCE_ADD)
    work = torch.empty_like(tree)
    ms = _median_ms(lambda: ctx.blelloch_downsweep_u32(work, n, 1, True), reps, before=lambda: work.copy_(tree))
    res["blelloch_downsweep_u32"] = _hbm(ms, 8 * n, n=n, algorithmic_bytes_per_key=8)
    del keys, tree, work
    # point-light animation of 2^24 lights in place, one 60 Hz frame at speed 10 per call (position and direction
    # read and written: 64 B/light)
    nl = 1 << 24
    gen = torch.Generator(device="cuda").manual_seed(7)
    lpos = torch.rand(nl, 4, device="cuda", generator=gen) * 120.0 - 60.0
    ldir = torch.randn(nl, 4, device="cuda", generator=gen)
    ldir[:, 3] = 0.0
    ldir = (ldir / ldir.norm(dim=1, keepdim=True)).contiguous()
    ms = _median_ms(lambda: ctx.bounce_point_lights(lpos, ldir, nl, (-50.0,) * 3, (50.0,) * 3, 10.0, 1.0 / 60.0), reps)
    res["bounce_point_lights_2p24"] = _hbm(ms, 64 * nl, n=nl, algorithmic_bytes_per_light=64)
    del lpos, ldir
    res.update(bench_shade_frame(ctx, vren_amd, reps))
    res.update(bench_pyramid(ctx, vren_amd, reps))
    row_traffic = profiled_traffic("row-traffic-bytes-per-call")
    for name, r in res.items():
        if "bound" in r:
            r["traffic"] = row_traffic.get(name)
    return res


def bench_pyramid(ctx, vren_amd, reps, frames=64):
    """Hi-Z max pyramid of 64 stacked 1920x1080 depth buffers in one call: the depth read once (4 B/px), level 0
    and every smaller level written once (4 B per texel of the pyramid)."""
    import torch
    cam = make_camera(vren_amd)
    d = torch.empty(frames * W * H, dtype=torch.float32, device="cuda")
    n = torch.empty(W * H * 4, dtype=torch.int16, device="cuda")
    for f in range(frames):
        ctx.synth_frame(SEED_FRAMES + f, W, H, cam, 64, d[f * W * H:(f + 1) * W * H], n)
    plen = int(ctx.lib.vren_depth_pyramid_length(W, H))
    out = torch.empty(frames * plen, dtype=torch.float32, device="cuda")
    ms = _median_ms(lambda: ctx.depth_pyramid(d, W, H, frames, out), reps)
    nbytes = frames * (4 * W * H + 4 * plen)
    return {"depth_pyramid_64x1080p": _hbm(ms, nbytes, images=frames, texels_per_pyramid=plen,
                                           algorithmic_bytes_per_pixel=round((4 * W * H + 4 * plen) / (W * H), 3))}


def bench_material_path(ctx, vren_amd, keep, args, plain_steps):
    """The bench's 64 frames with the reference's material path in shade (texcoord + material-index G-buffer, 8
    materials over 6 RGBA8 textures of 1x1 to 256x256 in every filter and address mode): batched and plain
    frames/s."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    rng = np.random.default_rng(11)
    tex_dev, tab = [], (vren_amd.Texture * 6)()
    for k, (w, h) in enumerate([(256, 256), (64, 128), (7, 5), (1, 1), (33, 17), (128, 64)]):
        t = torch.from_numpy(rng.integers(0, 256, (h, w, 4), dtype=np.uint8)).cuda()
        tex_dev.append(t)
        tab[k].texels, tab[k].width, tab[k].height = t.data_ptr(), w, h
        tab[k].filter, tab[k].address_u, tab[k].address_v = k % 2, k % 3, (k + 1) % 3
    mats = (vren_amd.Material * 8)()
    for m in range(8):
        mats[m].base_color_texture_idx, mats[m].metallic_roughness_texture_idx = (2 * m) % 6, (2 * m + 1) % 6
        mats[m].metallic_factor, mats[m].roughness_factor = 0.1 * m, 0.3 + 0.08 * m
        for c in range(4):
            mats[m].base_color_factor[c] = 0.5 + 0.05 * c
    d_tab = torch.from_numpy(np.frombuffer(bytes(tab), np.uint8).copy()).cuda()
    d_mats = torch.from_numpy(np.frombuffer(bytes(mats), np.uint8).copy()).cuda()
    gy, gx = torch.meshgrid(torch.arange(H, device="cuda"), torch.arange(W, device="cuda"), indexing="ij")
    tc = torch.stack([gx.reshape(-1) / W * 3.0 - 1.0, gy.reshape(-1) / H * 2.5 - 0.7], dim=1).float().contiguous()
    mi = ((gx // 97 + gy // 61) % 8).reshape(-1).to(torch.int16).contiguous()
    sets, descs = [], []
    for fd0 in keep[-1]:
        ms = vren_amd.MaterialSet()
        ms.texcoords, ms.material_indices = tc.data_ptr(), mi.data_ptr()
        ms.materials, ms.material_count, ms.textures, ms.texture_count = d_mats.data_ptr(), 8, d_tab.data_ptr(), 6
        sets.append(ms)
        fd = vren_amd.FrameDesc.from_buffer_copy(fd0)
        fd.materials = ctypes.pointer(ms)
        descs.append(fd)
    b = vren_amd.Batch(ctx, descs, vren_amd.default_cfg())
    for _ in range(max(1, args.warmup)):
        b.run(b.GRAPH)
    torch.cuda.synchronize()
    ms_b, _ = time_events(lambda: b.run(b.GRAPH), args.steps)
    for _ in range(max(1, args.warmup)):
        b.run(b.PER_FRAME)
    _, wall_p = time_events(lambda: b.run(b.PER_FRAME), plain_steps)
    shade_ms = 0.0
    for _ in range(3):
        shade_ms += b.run_profiled()[KERNELS.index("shade")] / 3
    ctx.sync_status()
    b.close()
    fps = args.frames * args.steps / (ms_b * 1e-3)
    plain = args.frames * plain_steps / (wall_p * 1e-3)
    px = args.frames * W * H
    print("[bench] material path: %.3f ms/step batched, %.1f frames/s plain" % (ms_b / args.steps, plain),
          file=sys.stderr, flush=True)
    # shade on the material path reads depth 4, normal 8, texcoord 8, material index 2, reference 4 and writes 16 B
    # per pixel (the material and texture tables are a few hundred KB and stay in the caches)
    return {"frames_per_s": round(fps, 2), "ms_per_step": round(ms_b / args.steps, 4),
            "plain_frames_per_s": round(plain, 2), "batched_over_plain": round(fps / plain, 3),
            "shade": _hbm(shade_ms, 42 * px, algorithmic_bytes_per_pixel=42)}


def bench_heterogeneous(ctx, vren_amd, args):
    """The batched configuration: 64 frames, resolutions drawn from {640x360, 1280x720, 1920x1080, 2560x1440,
    3840x2160}, light counts log-uniform in [256, 65536] (the same draw as the 64-frame heterogeneous parity test),
    batched (one graph replay) and plain (per frame launches + stream sync) frames/s."""
    import numpy as np
    import torch
    rng = np.random.default_rng(4444)
    res = [(640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]
    keep, descs = [], []
    total_px = 0
    for f in range(64):
        w, h = res[rng.integers(0, len(res))]
        L = int(round(float(np.exp(rng.uniform(np.log(256.0), np.log(65536.0))))))
        cam = vren_amd.Camera()
        cam.position[2] = -50.0
        cam.fov_y = float(np.float32(np.deg2rad(90.0)))
        cam.aspect_ratio = float(np.float32(w) / np.float32(h))
        cam.near_plane, cam.far_plane = 0.1, 100.0
        t = dict(d=torch.empty(w * h, dtype=torch.float32, device="cuda"),
                 n=torch.empty(w * h * 4, dtype=torch.int16, device="cuda"),
                 a=torch.empty(w * h * 4, dtype=torch.float32, device="cuda"),
                 m=torch.empty(w * h, dtype=torch.float32, device="cuda"),
                 p=torch.empty(L * 4, dtype=torch.float32, device="cuda"),
                 l=torch.empty(L * 4, dtype=torch.float32, device="cuda"),
                 o=torch.empty(w * h * 4, dtype=torch.float32, device="cuda"))
        ctx.synth_frame(7000 + f, w, h, cam, 64, t["d"], t["n"])
        ctx.synth_materials(7000 + f, w * h, t["a"], t["m"])
        ctx.synth_lights(7007 + f, L, t["p"], t["l"])
        fd = vren_amd.FrameDesc()
        fd.width, fd.height, fd.camera = w, h, cam
        fd.depth, fd.normals = t["d"].data_ptr(), t["n"].data_ptr()
        fd.albedo_roughness, fd.metallic = t["a"].data_ptr(), t["m"].data_ptr()
        fd.light_positions, fd.point_lights, fd.light_count = t["p"].data_ptr(), t["l"].data_ptr(), L
        fd.output = t["o"].data_ptr()
        descs.append(fd)
        keep.append(t)
        total_px += w * h
    torch.cuda.synchronize()
    b = vren_amd.Batch(ctx, descs, vren_amd.default_cfg())
    shape = None
    if not args.no_tune:
        shape, _ = b.autotune(reps=2)  # this batch's size class has its own shape
    for _ in range(max(1, args.warmup)):
        b.run(b.GRAPH)
    torch.cuda.synchronize()
    steps = max(3, args.steps // 2)
    ms_b, _ = time_events(lambda: b.run(b.GRAPH), steps)
    acc = [0.0] * len(KERNELS)
    for _ in range(3):
        acc = [a + m for a, m in zip(acc, b.run_profiled())]
    kernel_ms = {k: round(a / 3, 5) for k, a in zip(KERNELS, acc)}
    b.run(b.PER_FRAME)
    # as many plain runs as batched ones (the plain path's time varies more from run to run)
    _, wall_p = time_events(lambda: b.run(b.PER_FRAME), steps)
    ctx.sync_status()
    b.close()
    fps = 64 * steps / (ms_b * 1e-3)
    plain = 64 * steps / (wall_p * 1e-3)
    print("[bench] 64 heterogeneous frames: %.3f ms/step batched, %.1f frames/s plain" % (ms_b / steps, plain),
          file=sys.stderr, flush=True)
    return {"frames": 64, "pixels": total_px, "frames_per_s": round(fps, 2), "ms_per_step": round(ms_b / steps, 4),
            "plain_frames_per_s": round(plain, 2), "batched_over_plain": round(fps / plain, 3),
            "megapixels_per_s": round(total_px * steps / (ms_b * 1e-3) / 1e6, 1), "kernel_ms": kernel_ms,
            "chain_shape": list(shape) if shape else None}


def bench_shade_frame(ctx, vren_amd, reps):
    """Shade forward and adjoint on one 1920x1080 frame with 4096 lights (smooth windowed attenuation), light
    lists from the chain's stages."""
    import numpy as np
    import torch
    res = {}
    cam = make_camera(vren_amd)
    cf_lists = vren_amd.default_cfg()
    cf = vren_amd.default_cfg()
    cf.attenuation = 1
    npx = W * H
    key_cap, index_cap = 1 << 20, 1 << 22
    d = torch.empty(npx, dtype=torch.float32, device="cuda")
    nrm = torch.empty(npx * 4, dtype=torch.int16, device="cuda")
    ar = torch.empty(npx * 4, dtype=torch.float32, device="cuda")
    me = torch.empty(npx, dtype=torch.float32, device="cuda")
    lp = torch.empty(L_LIGHTS * 4, dtype=torch.float32, device="cuda")
    lc = torch.empty(L_LIGHTS * 4, dtype=torch.float32, device="cuda")
    ctx.synth_frame(SEED_FRAMES, W, H, cam, 64, d, nrm)
    ctx.synth_materials(SEED_FRAMES, npx, ar, me)
    ctx.synth_lights(SEED_FRAMES, L_LIGHTS, lp, lc)
    u8 = lambda nbytes: torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    vp, codes = u8(16 * L_LIGHTS), u8(8 * L_LIGHTS)
    bvh = u8(32 * ctx.lib.vren_calc_bvh_buffer_length(L_LIGHTS))
    ctx.construct_point_light_bvh(lp, lc, L_LIGHTS, cam, cf_lists, vp, codes, bvh)
    ckeys, cnt, refs = u8(4 * key_cap), u8(4), u8(4 * npx)
    ctx.find_unique_cluster_list(W, H, cam, cf_lists, d, nrm, ckeys, key_cap, cnt, refs)
    counts, offsets, idx, tot = u8(4 * key_cap), u8(4 * key_cap), u8(4 * index_cap), u8(4)
    ctx.assign_lights(W, H, cam, cf_lists, ckeys, cnt, key_cap, bvh, L_LIGHTS, codes, vp, counts, offsets, idx,
                      index_cap, tot)
    img = torch.empty(npx * 4, dtype=torch.float32, device="cuda")
    grad = torch.rand(npx * 4, dtype=torch.float32, device="cuda")
    g_pos = torch.zeros(L_LIGHTS * 4, dtype=torch.float32, device="cuda")
    g_lig = torch.zeros(L_LIGHTS * 4, dtype=torch.float32, device="cuda")
    g_ar = torch.empty(npx * 4, dtype=torch.float32, device="cuda")

    def clear():
        g_pos.zero_()
        g_lig.zero_()
        g_ar.zero_()

    fwd = lambda: ctx.shade(W, H, cam, cf, d, nrm, ar, me, refs, key_cap, counts, offsets, idx, index_cap, lp, lc,
                            L_LIGHTS, img)
    bwd = lambda: ctx.shade_backward(W, H, cam, cf, d, nrm, ar, me, refs, key_cap, counts, offsets, idx, index_cap,
                                     lp, lc, L_LIGHTS, grad, g_pos, g_lig, g_ar)
    # Karras LBVH over the frame's 4096 sorted (Morton code, light index) pairs: 4095 internal nodes
    lb_nodes, lb_parents = u8(32 * (2 * L_LIGHTS - 1)), u8(4 * (2 * L_LIGHTS - 1))
    ms_lb = _median_ms(lambda: ctx.build_lbvh(codes, L_LIGHTS, vp, lc, lb_nodes, lb_parents), reps)
    res["lbvh_4096_lights"] = {"ms": round(ms_lb, 4), "lights": L_LIGHTS, "nodes": 2 * L_LIGHTS - 1,
                               "bound": "latency (bottom-up refit along the tree depth)"}
    ms_f = _median_ms(fwd, reps)
    ms_b = _median_ms(bwd, reps, before=clear)
    torch.cuda.synchronize()
    # pairs the two kernels evaluate: every shaded pixel's cluster list
    _, proj = ctx.camera_matrices(cam)
    inv = np.linalg.inv(proj.astype(np.float64).reshape(4, 4).T)
    zd = d.cpu().numpy().astype(np.float64)
    z = (inv[2, 2] * zd + inv[2, 3]) / (inv[3, 2] * zd + inv[3, 3])
    r = refs.cpu().numpy().view(np.uint32)
    c = counts.cpu().numpy().view(np.uint32)
    ok = (z < float(np.float32(cam.far_plane) - np.float32(0.001))) & (r < key_cap)
    pairs = int(c[r[ok]].sum(dtype=np.int64))
    res["shade_forward_1080p"] = _hbm(ms_f, 52 * npx, pixels=npx, algorithmic_bytes_per_pixel=52, pairs=pairs,
                                      gpairs_per_s=round(pairs / (ms_f * 1e-3) / 1e9, 3))
    # adjoint: G-buffer, reference and dL/d(rgba) in (52 B), dL/d(albedo, roughness) accumulated (16 B read, 16 B
    # written) per pixel, plus 28 B of light gradient per light
    res["shade_adjoint_1080p"] = _hbm(ms_b, 84 * npx + 28 * L_LIGHTS, pixels=npx, algorithmic_bytes_per_pixel=84,
                                      pairs=pairs, gpairs_per_s=round(pairs / (ms_b * 1e-3) / 1e9, 3),
                                      lights=L_LIGHTS)
    return res


def _cpu_threads():
    """std::thread::hardware_concurrency() of the host, capped at this GPU's share of it (16 cores)."""
    return max(1, min(16, os.cpu_count() or 1))


def _chain_frames(co, ol, seeds, threads):
    """Wall seconds of the oracle chain (BVH, clusters, assignment, the reference's fp32 shade) over one 1080p frame
    per seed, `threads` frames at a time, each frame on one host thread; inputs generated before the timing."""
    import concurrent.futures
    import numpy as np
    cam = ol.camera(W, H)
    cf = ol.cfg()

    def synth(seed):
        depth, normals = co.synth_frame(seed, W, H, cam)
        ar, me = co.synth_materials(seed, W * H)
        pos, lig = co.synth_lights(seed + 7, L_LIGHTS)
        return depth, normals, ar, me, pos, lig

    def chain(inp):
        depth, normals, ar, me, pos, lig = inp
        vp, codes, bvh = co.light_bvh(pos, lig, cam, cf)
        keys, U, refs = co.unique_clusters(W, H, cam, cf, depth, normals)
        counts, offsets, idx, tot = co.assign(W, H, cam, cf, keys, U, bvh, codes, vp)
        cc = np.zeros(1 << 20, np.uint32)
        oo = np.zeros(1 << 20, np.uint32)
        cc[:counts.size] = counts
        oo[:offsets.size] = offsets
        co.shade(W, H, cam, cf, depth, normals, ar, me, refs, 1 << 20, cc, oo, idx, pos, lig, f32=True)

    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        inputs = list(ex.map(synth, seeds))
        t0 = time.perf_counter()
        list(ex.map(chain, inputs))
        return time.perf_counter() - t0


def cpu_baseline(sort_log2=(22, 26), scan_log2=26):
    """The CPU restatement (oracle/) on the host: 1 thread and N = hardware_concurrency() threads (capped at the
    GPU's 16-core share; std::thread, no TBB). The chain runs one 1080p frame per thread (N frames at once for the
    N-thread figure); std::stable_sort of configuration-1 (key, index) pairs, std::exclusive_scan and std::accumulate
    of configuration-1 keys are split over the threads (chunks, then merge levels / a prefix over chunk totals).
    Input generation is outside every timing; each figure is the best of 3 runs (the chain: one run)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    o = ol.Oracle()
    co = ol.ClusteredOracle(o)
    nt = _cpu_threads()
    t1 = _chain_frames(co, ol, [SEED_FRAMES], 1)
    tn = _chain_frames(co, ol, [SEED_FRAMES + f for f in range(nt)], nt)
    res = {
        "value": round(nt / tn, 3), "unit": "frames/s", "cores": nt, "kind": "port",
        "sample": "%d frames of 1920x1080 / 4096 lights through the oracle chain (BVH, clusters, assignment, reference "
                  "fp32 shade), one frame per thread, %d threads at once; input generation excluded" % (nt, nt),
        "single_thread": {"value": round(1.0 / t1, 3), "unit": "frames/s", "cores": 1,
                          "sample": "1 frame of the same chain on one thread"},
        "host_cpu": _cpu_model(),
    }
    sorts = {}
    for lg in sort_log2:
        n = 1 << lg
        keys = o.pcg32(42, 54, n)
        vals = np.arange(n, dtype=np.uint32)
        entry = {}
        for th in (1, nt):
            best = min(o.cpu_stable_sort_pairs(keys, vals, th)[0] for _ in range(1 if lg >= 26 and th == 1 else 3))
            entry["threads_%d" % th] = round(n / best / 1e9, 4)
        sorts["2^%d" % lg] = entry
    res["sort"] = {"unit": "Gpairs/s", "cores": nt, "sample": "std::stable_sort of configuration-1 (key, index) "
                   "pairs, per thread count", "results": sorts}
    n = 1 << scan_log2
    x = o.pcg32(42, 54, n)
    scan, red = {}, {}
    for th in (1, nt):
        scan["threads_%d" % th] = round(n / min(o.cpu_exclusive_scan(x, th)[0] for _ in range(3)) / 1e9, 4)
        red["threads_%d" % th] = round(n / min(o.cpu_reduce_add(x, th)[0] for _ in range(3)) / 1e9, 4)
    res["exclusive_scan"] = {"unit": "Gkeys/s", "cores": nt, "sample": "std::exclusive_scan of 2^%d uint32 "
                             "configuration-1 keys" % scan_log2, "results": scan}
    res["reduce_add"] = {"unit": "Gkeys/s", "cores": nt, "sample": "std::accumulate of 2^%d uint32 configuration-1 "
                         "keys" % scan_log2, "results": red}
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip() + " (%d logical cpus)" % os.cpu_count()
    except OSError:
        pass
    return "unknown"


def bench_key_variant(ctx, vren_amd, descs, args, plain_steps):
    """Batched and plain frames/s of the bench's frames with the 32x32 tile x 32 log-depth slices x 6 normal-cone
    bins cluster-key encoder."""
    import torch
    cfg32 = vren_amd.default_cfg(key_mode=vren_amd.KEYS_SLICED, depth_slices=32)
    batch32 = vren_amd.Batch(ctx, descs, cfg32)
    for _ in range(max(1, args.warmup)):
        batch32.run(batch32.GRAPH)
    torch.cuda.synchronize()
    ms32, _ = time_events(lambda: batch32.run(batch32.GRAPH), args.steps)
    for _ in range(max(1, args.warmup)):
        batch32.run(batch32.PER_FRAME)
    _, wall32 = time_events(lambda: batch32.run(batch32.PER_FRAME), plain_steps)
    ctx.sync_status()
    batch32.close()
    fps32 = args.frames * args.steps / (ms32 * 1e-3)
    plain32 = args.frames * plain_steps / (wall32 * 1e-3)
    print("[bench] 32-slice keys: %.3f ms/step batched, %.1f frames/s plain" % (ms32 / args.steps, plain32),
          file=sys.stderr, flush=True)
    return {"key_encoder": "32x32 tile x 32 log-depth slices x 6 normal-cone bins", "frames_per_s": round(fps32, 2),
            "ms_per_step": round(ms32 / args.steps, 4), "plain_frames_per_s": round(plain32, 2),
            "batched_over_plain": round(fps32 / plain32, 3)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--sort-log2", type=int, default=26)
    ap.add_argument("--no-sort", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the radix-sort and chain-shape autotuners")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rows", action="store_true", help="skip reduce, scan, bucket sort and the shade adjoint")
    ap.add_argument("--no-key-variant", action="store_true",
                    help="skip the second batch with the 32-slice x 6-bin cluster-key encoder")
    ap.add_argument("--no-heterogeneous", action="store_true", help="skip the 64-frame heterogeneous batch")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy this many MiB with one plain device copy kernel (scales HBM counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: the clustered-shading batches shard by frame with no exchange between "
              "GPUs, so there is no multi-GPU run to make here (asked for --gpus %d, WORLD_SIZE %d)"
              % (args.gpus, world), file=sys.stderr)
        sys.exit(2)
    import torch
    import vren_amd
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    ctx = vren_amd.Context(0)
    if args.copy_calibration_mib:
        nb = args.copy_calibration_mib << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] copy calibration: %d bytes read, %d bytes written" % (nb, nb), file=sys.stderr, flush=True)
        del src, dst
    print("[bench] building %d frames of %dx%d with %d lights" % (args.frames, W, H, L_LIGHTS), file=sys.stderr,
          flush=True)
    batch, keep = build_batch(ctx, vren_amd, args.frames)
    chain_tune = None
    if not args.no_tune:
        # the chain autotuner: every launch shape timed on this batch, the fastest kept (setup, not timed)
        shape, tms = batch.autotune(reps=3)
        chain_tune = {"shade_pixels_per_thread": shape[0], "assign_waves_per_cu": shape[1],
                      "light_sort_bits": shape[2], "light_stage_concurrent": shape[3],
                      "tuned_ms_per_batched_run": round(tms, 4)}
        print("[bench] chain shape %s (%.3f ms per batched run)" % (shape, tms), file=sys.stderr, flush=True)

    # batched: one graph replay per step
    for _ in range(max(1, args.warmup)):
        batch.run(batch.GRAPH)
    torch.cuda.synchronize()
    ms_b, wall_b = time_events(lambda: batch.run(batch.GRAPH), args.steps)
    ms_per_step = ms_b / args.steps
    value = args.frames * args.steps / (ms_b * 1e-3)
    print("[bench] batched: %.3f ms/step (wall %.3f)" % (ms_per_step, wall_b / args.steps), file=sys.stderr,
          flush=True)

    # plain: per frame launches + stream sync
    for _ in range(max(1, args.warmup)):
        batch.run(batch.PER_FRAME)
    plain_steps = max(2, args.steps // 2)
    ms_p, wall_p = time_events(lambda: batch.run(batch.PER_FRAME), plain_steps)
    plain_fps = args.frames * plain_steps / (wall_p * 1e-3)
    print("[bench] plain: %.3f ms/step (wall %.3f)" % (ms_p / plain_steps, wall_p / plain_steps), file=sys.stderr,
          flush=True)
    ctx.sync_status()

    # per-kernel durations of the batched chain (HIP events between the kernels, same stream)
    acc = [0.0] * len(KERNELS)
    prof_steps = max(3, args.steps // 2)
    for _ in range(prof_steps):
        ms = batch.run_profiled()
        for i, v in enumerate(ms):
            acc[i] += v
    kernel_ms = {k: round(a / prof_steps, 5) for k, a in zip(KERNELS, acc)}
    dom = max(kernel_ms, key=lambda k: kernel_ms[k])
    pixels = args.frames * W * H
    if dom in BYTES_PER_PIXEL:
        alg_bytes = BYTES_PER_PIXEL[dom] * pixels
    elif dom in BYTES_PER_LIGHT:
        alg_bytes = BYTES_PER_LIGHT[dom] * args.frames * L_LIGHTS
    else:
        alg_bytes = None
    achieved = alg_bytes / (kernel_ms[dom] * 1e-3) / 1e9 if alg_bytes else None
    traffic = profiled_traffic()
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                "traffic": traffic.get(dom + "_kernel") if args.frames == 64 else None, "launch_ms": kernel_ms[dom],
                "algorithmic_bytes_per_launch": alg_bytes}
    # the shade kernel's arithmetic view: FP32 operations of the pairs it evaluates per launch
    pairs = shade_pairs(ctx, vren_amd, batch, keep)
    tflops = SHADE_OPS_PER_PAIR * pairs / (kernel_ms["shade"] * 1e-3) / 1e12
    shade_valu = {"kernel": "shade", "bound": "valu", "achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS,
                  "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "pairs_per_launch": pairs,
                  "pairs_per_pixel": round(pairs / pixels, 3), "ops_per_pair": SHADE_OPS_PER_PAIR,
                  "launch_ms": kernel_ms["shade"]}

    # the same frames with the 32x32 tile x 32 log-depth slice x 6 normal-cone bin cluster-key encoder
    # instead of the reference's (1024 log-depth slices x 64 normal codes)
    sliced = None
    if not args.no_key_variant:
        sliced = bench_key_variant(ctx, vren_amd, keep[-1], args, plain_steps)

    material = None
    if not args.no_key_variant:
        material = bench_material_path(ctx, vren_amd, keep, args, plain_steps)
    hetero = None
    if not args.no_heterogeneous:
        print("[bench] 64 heterogeneous frames", file=sys.stderr, flush=True)
        hetero = bench_heterogeneous(ctx, vren_amd, args)

    sort = None
    if not args.no_sort:
        print("[bench] sort 2^%d pairs" % args.sort_log2, file=sys.stderr, flush=True)
        sort = bench_sort(ctx, args.sort_log2, args.steps, args.warmup, not args.no_tune)

    rows = None
    if not args.no_rows:
        print("[bench] reduce, scan, bucket sort, shade forward and adjoint", file=sys.stderr, flush=True)
        rows = bench_primitives(ctx, vren_amd, max(5, args.steps))

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline()

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (device-generated procedural scenes and lights)",
        "config": {"workload": "clustered shading, batches of %d frames of 1920x1080 with 4096 point lights each "
                               "(light BVH + unique clusters + BVH light assignment + PBR shade), hipGraph replay"
                               % args.frames,
                   "frames_per_step": args.frames, "resolution": "1920x1080", "lights_per_frame": L_LIGHTS,
                   "key_encoder": "reference (32x32 tile x 1024 log-depth slices x 64 normal codes)",
                   "parallelism": "1 GPU"},
        "plain_frames_per_s": round(plain_fps, 2),
        "batched_over_plain": round(value / plain_fps, 3),
        "config_key_encoder": sliced,
        "config_material_path": material,
        "heterogeneous_64_frames": hetero,
        "chain_shape": chain_tune,
        "kernel_ms": kernel_ms,
        "roofline": roofline,
        "shade_valu": shade_valu,
        "cpu_baseline": cpu,
        "sort": sort,
        "rows": rows,
    }
    batch.close()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
