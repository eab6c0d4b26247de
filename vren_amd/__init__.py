"""vren_amd -- MI355X (gfx950) implementation of vren's clustered-shading compute path.

The product is the C ABI in include/vren_hip/capi.h, implemented by libvren_amd.so (hand-written HIP kernels),
with the drop-in C++ operator classes in include/vren/. This package loads that library with ctypes and wraps
each entry point so that device buffers can be passed as torch tensors (torch provides device memory and
streams only). Every call goes to the library; there is no other implementation to fall back to.
"""
import ctypes

from . import _lib
from ._lib import (ADDRESS_CLAMP_TO_EDGE, ADDRESS_MIRRORED_REPEAT, ADDRESS_REPEAT, ATTEN_REFERENCE, ATTEN_SMOOTH,
                   BVH_INVALID_NODE, BVH_LEAF_NODE, CLUSTER_AABB_REFERENCE, CLUSTER_AABB_TIGHT, FILTER_LINEAR,
                   FILTER_NEAREST, KEYS_REFERENCE, KEYS_SLICED, MORTON_GIVEN_AABB, MORTON_LIGHTS_AABB, REDUCE_ADD,
                   REDUCE_MAX, REDUCE_MIN, Camera, ChainShape, ClusterCfg, FrameDesc, FrameOutputs, Material,
                   MaterialSet, Texture, VrenError)

__all__ = ["Context", "Camera", "ClusterCfg", "FrameDesc", "FrameOutputs", "VrenError", "lib", "default_cfg",
           "REDUCE_ADD", "REDUCE_MIN", "REDUCE_MAX", "KEYS_REFERENCE", "KEYS_SLICED", "CLUSTER_AABB_REFERENCE",
           "CLUSTER_AABB_TIGHT", "MORTON_LIGHTS_AABB", "MORTON_GIVEN_AABB", "ATTEN_REFERENCE", "ATTEN_SMOOTH",
           "BVH_LEAF_NODE", "BVH_INVALID_NODE", "Material", "MaterialSet", "Texture", "FILTER_NEAREST",
           "FILTER_LINEAR", "ADDRESS_REPEAT", "ADDRESS_MIRRORED_REPEAT", "ADDRESS_CLAMP_TO_EDGE"]


def lib():
    return _lib.load()


def default_cfg(**overrides):
    cfg = ClusterCfg()
    lib().vren_cluster_cfg_default(ctypes.byref(cfg))
    for k, v in overrides.items():
        if k in ("morton_aabb_min", "morton_aabb_max"):
            for i in range(3):
                getattr(cfg, k)[i] = v[i]
        else:
            setattr(cfg, k, v)
    return cfg


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """Owns a vren_ctx on one HIP device. Methods enqueue on torch's current stream of that device."""

    def __init__(self, device=0):
        self.lib = lib()
        h = ctypes.c_void_p()
        st = self.lib.vren_create(device, ctypes.byref(h))
        if st != 0:
            raise VrenError(st, "vren_create(%d) failed" % device)
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.vren_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def check(self, status):
        if status != 0:
            raise VrenError(status, self.lib.vren_last_error(self.h).decode())
        return status

    def sync_status(self):
        return self.check(self.lib.vren_sync_status(self.h, self.stream()))

    # ---- primitives -----------------------------------------------------------------------------------------
    def reduce_u32(self, inp, n, out_tree, op=REDUCE_ADD, blocks=1):
        return self.check(self.lib.vren_reduce_u32(self.h, self.stream(), _ptr(inp), n, _ptr(out_tree), op, blocks))

    def reduce_vec4(self, inp, n, out_tree, op=REDUCE_ADD, blocks=1):
        return self.check(self.lib.vren_reduce_vec4(self.h, self.stream(), _ptr(inp), n, _ptr(out_tree), op, blocks))

    def reduce_total_u32(self, inp, n, out, op=REDUCE_ADD):
        return self.check(self.lib.vren_reduce_total_u32(self.h, self.stream(), _ptr(inp), n, _ptr(out), op))

    def exclusive_scan_u32(self, inp, out, n, blocks=1):
        return self.check(self.lib.vren_exclusive_scan_u32(self.h, self.stream(), _ptr(inp), _ptr(out), n, blocks))

    def blelloch_downsweep_u32(self, data, length, blocks=1, clear_last=True):
        return self.check(self.lib.vren_blelloch_downsweep_u32(self.h, self.stream(), _ptr(data), length, blocks,
                                                               1 if clear_last else 0))

    def radix_sort_scratch_bytes(self, n, with_values):
        return self.lib.vren_radix_sort_scratch_bytes(n, 1 if with_values else 0)

    def radix_sort(self, keys, vals, n, scratch, begin_bit=0, end_bit=32):
        return self.check(self.lib.vren_radix_sort(self.h, self.stream(), _ptr(keys), _ptr(vals), n, begin_bit,
                                                   end_bit, _ptr(scratch), scratch.numel() * scratch.element_size()))

    def radix_force_config(self, cfg):
        return self.check(self.lib.vren_radix_sort_force_config(self.h, cfg))

    def autotune_radix_sort(self, n, with_values, key_bits=32):
        best = ctypes.c_int(-1)
        ms = ctypes.c_float(0.0)
        self.check(self.lib.vren_autotune_radix_sort(self.h, self.stream(), n, 1 if with_values else 0, key_bits,
                                                     ctypes.byref(best), ctypes.byref(ms)))
        return best.value, ms.value

    @staticmethod
    def radix_configs():
        L = lib()
        out = []
        for i in range(L.vren_radix_sort_config_count()):
            b, k, w, a = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
            L.vren_radix_sort_config_info(i, ctypes.byref(b), ctypes.byref(k), ctypes.byref(w), ctypes.byref(a))
            out.append((b.value, k.value, w.value, a.value))
        return out

    def bucket_sort(self, in_pairs, n, out_pairs, offsets, n_buckets=65536, key_shift=0):
        return self.check(self.lib.vren_bucket_sort(self.h, self.stream(), _ptr(in_pairs), n, _ptr(out_pairs),
                                                    _ptr(offsets), n_buckets, key_shift))

    def bounce_point_lights(self, positions, directions, n, aabb_min, aabb_max, speed, dt):
        """Moves n lights (float32 [n, 4] position and direction tensors, in place) by speed * dt inside the AABB."""
        lo, hi = (ctypes.c_float * 3)(*aabb_min), (ctypes.c_float * 3)(*aabb_max)
        return self.check(self.lib.vren_bounce_point_lights(self.h, self.stream(), _ptr(positions), _ptr(directions),
                                                            n, lo, hi, speed, dt))

    def build_bvh(self, nodes, padded_leaf_count):
        return self.check(self.lib.vren_build_bvh(self.h, self.stream(), _ptr(nodes), padded_leaf_count))

    def build_lbvh(self, codes, n, view_positions, point_lights, nodes, parents):
        """Karras LBVH over n sorted (code, light index) pairs: nodes[2n-1] (32 B each), parents[2n-1]."""
        return self.check(self.lib.vren_build_lbvh(self.h, self.stream(), _ptr(codes), n, _ptr(view_positions),
                                                   _ptr(point_lights), _ptr(nodes), _ptr(parents)))

    # ---- clustered shading stages (one frame) --------------------------------------------------------------
    def camera_matrices(self, cam):
        import numpy as np
        view = (ctypes.c_float * 16)()
        proj = (ctypes.c_float * 16)()
        self.lib.vren_camera_matrices(ctypes.byref(cam), view, proj)
        return np.array(view, np.float32), np.array(proj, np.float32)

    def construct_point_light_bvh(self, light_pos, lights, L, cam, cfg, view_pos, codes, bvh):
        return self.check(self.lib.vren_construct_point_light_bvh(
            self.h, self.stream(), _ptr(light_pos), _ptr(lights), L, ctypes.byref(cam), ctypes.byref(cfg),
            _ptr(view_pos), _ptr(codes), _ptr(bvh)))

    def find_unique_cluster_list(self, W, H, cam, cfg, depth, normals, keys, key_cap, count, refs):
        return self.check(self.lib.vren_find_unique_cluster_list(
            self.h, self.stream(), W, H, ctypes.byref(cam), ctypes.byref(cfg), _ptr(depth), _ptr(normals),
            _ptr(keys), key_cap, _ptr(count), _ptr(refs)))

    def assign_lights(self, W, H, cam, cfg, keys, count, key_cap, bvh, L, codes, view_pos, counts, offsets, indices,
                      index_cap, total):
        return self.check(self.lib.vren_assign_lights(
            self.h, self.stream(), W, H, ctypes.byref(cam), ctypes.byref(cfg), _ptr(keys), _ptr(count), key_cap,
            _ptr(bvh), L, _ptr(codes), _ptr(view_pos), _ptr(counts), _ptr(offsets), _ptr(indices), index_cap,
            _ptr(total)))

    def shade(self, W, H, cam, cfg, depth, normals, albedo_rough, metallic, refs, key_cap, counts, offsets, indices,
              index_cap, light_pos, lights, L, out, dir_lights=None, n_dir=0):
        return self.check(self.lib.vren_shade(
            self.h, self.stream(), W, H, ctypes.byref(cam), ctypes.byref(cfg), _ptr(depth), _ptr(normals),
            _ptr(albedo_rough), _ptr(metallic), _ptr(refs), key_cap, _ptr(counts), _ptr(offsets), _ptr(indices),
            index_cap, _ptr(light_pos), _ptr(lights), L, _ptr(dir_lights), n_dir, _ptr(out)))

    def shade_backward(self, W, H, cam, cfg, depth, normals, albedo_rough, metallic, refs, key_cap, counts, offsets,
                       indices, index_cap, light_pos, lights, L, grad_out, g_pos, g_lights, g_albedo_rough):
        return self.check(self.lib.vren_shade_backward(
            self.h, self.stream(), W, H, ctypes.byref(cam), ctypes.byref(cfg), _ptr(depth), _ptr(normals),
            _ptr(albedo_rough), _ptr(metallic), _ptr(refs), key_cap, _ptr(counts), _ptr(offsets), _ptr(indices),
            index_cap, _ptr(light_pos), _ptr(lights), L, _ptr(grad_out), _ptr(g_pos), _ptr(g_lights),
            _ptr(g_albedo_rough)))

    def shade_materials(self, W, H, cam, cfg, depth, normals, material_set, refs, key_cap, counts, offsets, indices,
                        index_cap, light_pos, lights, L, out, dir_lights=None, n_dir=0):
        """`material_set` is a MaterialSet whose pointers are device addresses (texcoords, material indices, the
        material table and the texture descriptor table)."""
        return self.check(self.lib.vren_shade_materials(
            self.h, self.stream(), W, H, ctypes.byref(cam), ctypes.byref(cfg), _ptr(depth), _ptr(normals),
            ctypes.byref(material_set), _ptr(refs), key_cap, _ptr(counts), _ptr(offsets), _ptr(indices), index_cap,
            _ptr(light_pos), _ptr(lights), L, _ptr(dir_lights), n_dir, _ptr(out)))

    # ---- depth pyramid and debug views ---------------------------------------------------------------------
    def depth_pyramid(self, depth, W, H, images, pyramid):
        return self.check(self.lib.vren_depth_pyramid(self.h, self.stream(), _ptr(depth), W, H, images,
                                                      _ptr(pyramid)))

    def show_clusters(self, W, H, mode, refs, keys, key_cap, counts, offsets, indices, index_cap, out):
        return self.check(self.lib.vren_show_clusters(self.h, self.stream(), W, H, mode, _ptr(refs), _ptr(keys),
                                                      key_cap, _ptr(counts), _ptr(offsets), _ptr(indices), index_cap,
                                                      _ptr(out)))

    def visualize_bvh(self, bvh, level_count, vertices):
        return self.check(self.lib.vren_visualize_bvh(self.h, self.stream(), _ptr(bvh), level_count, _ptr(vertices)))

    # ---- synthetic inputs ----------------------------------------------------------------------------------
    def synth_keys(self, initstate, initseq, n, out):
        return self.check(self.lib.vren_synth_keys(self.h, self.stream(), initstate, initseq, n, _ptr(out)))

    def synth_lights(self, initstate, L, light_pos, lights):
        return self.check(self.lib.vren_synth_lights(self.h, self.stream(), initstate, L, _ptr(light_pos),
                                                     _ptr(lights)))

    def synth_frame(self, initstate, W, H, cam, n_boxes, depth, normals):
        return self.check(self.lib.vren_synth_frame(self.h, self.stream(), initstate, W, H, ctypes.byref(cam),
                                                    n_boxes, _ptr(depth), _ptr(normals)))

    def synth_materials(self, initstate, n, albedo_rough, metallic):
        return self.check(self.lib.vren_synth_materials(self.h, self.stream(), initstate, n, _ptr(albedo_rough),
                                                        _ptr(metallic)))


class Batch:
    """cluster_and_shade over a list of frames (vren_batch_*). `frames` is a list of FrameDesc."""

    PER_FRAME, BATCHED, GRAPH = 0, 1, 2

    def __init__(self, ctx, frames, cfg):
        """A frame on the material path points its `materials` field at a MaterialSet, which must stay alive until
        the batch has been created."""
        self.ctx = ctx
        arr = (FrameDesc * len(frames))(*frames)
        self._frames = arr
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.vren_batch_create(ctx.h, arr, len(frames), ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.n = len(frames)

    def run(self, mode):
        return self.ctx.check(self.ctx.lib.vren_batch_run(self.h, self.ctx.stream(), mode))

    def run_profiled(self):
        ms = (ctypes.c_float * 9)()
        self.ctx.check(self.ctx.lib.vren_batch_run_profiled(self.h, self.ctx.stream(), ms))
        return list(ms)

    def outputs(self, f):
        o = FrameOutputs()
        self.ctx.check(self.ctx.lib.vren_batch_frame_outputs(self.h, f, ctypes.byref(o)))
        return o

    def shape(self):
        """(shade pixels per thread, assignment waves per CU, light-sort digit bits, concurrent light stage);
        0 = the default."""
        s = ChainShape()
        self.ctx.check(self.ctx.lib.vren_batch_get_shape(self.h, ctypes.byref(s)))
        return (s.shade_pixels_per_thread, s.assign_waves_per_cu, s.light_sort_bits, s.light_stage_concurrent)

    def set_shape(self, shade_ppt, assign_wpc, light_bits, light_concurrent=0):
        s = ChainShape(shade_ppt, assign_wpc, light_bits, light_concurrent)
        return self.ctx.check(self.ctx.lib.vren_batch_set_shape(self.h, ctypes.byref(s)))

    def autotune(self, reps=3):
        """Times every chain shape on this batch, keeps the fastest; returns (shape tuple, ms per batched run)."""
        s, ms = ChainShape(), ctypes.c_float()
        self.ctx.check(self.ctx.lib.vren_batch_autotune(self.h, self.ctx.stream(), reps, ctypes.byref(s),
                                                        ctypes.byref(ms)))
        return (s.shade_pixels_per_thread, s.assign_waves_per_cu, s.light_sort_bits,
                s.light_stage_concurrent), ms.value

    def close(self):
        if self.h:
            self.ctx.lib.vren_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
