"""ctypes binding of libvren_amd.so (the C ABI declared in include/vren_hip/capi.h).

The library is loaded from this package directory only; if it has not been built, importing raises instead of
falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvren_amd.so")

c_int, c_uint32, c_uint64, c_size_t, c_float, c_void_p, c_char_p = (
    ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_float, ctypes.c_void_p, ctypes.c_char_p)
P = ctypes.POINTER

VREN_OK, VREN_EINVAL, VREN_ECAPACITY, VREN_EHIP, VREN_ENOMEM = 0, -1, -2, -3, -4
REDUCE_ADD, REDUCE_MIN, REDUCE_MAX = 0, 1, 2
KEYS_REFERENCE, KEYS_SLICED = 0, 1
CLUSTER_AABB_REFERENCE, CLUSTER_AABB_TIGHT = 0, 1
MORTON_LIGHTS_AABB, MORTON_GIVEN_AABB = 0, 1
ATTEN_REFERENCE, ATTEN_SMOOTH = 0, 1
BVH_LEAF_NODE, BVH_INVALID_NODE = 0xFFFFFFFF, 0xFFFFFFFE


class Camera(ctypes.Structure):
    _fields_ = [("position", c_float * 3), ("fov_y", c_float), ("aspect_ratio", c_float), ("near_plane", c_float),
                ("far_plane", c_float), ("yaw", c_float), ("pitch", c_float)]


class ClusterCfg(ctypes.Structure):
    _fields_ = [("key_mode", c_uint32), ("depth_slices", c_uint32), ("cluster_aabb_mode", c_uint32),
                ("morton_bits", c_uint32), ("morton_aabb_mode", c_uint32), ("morton_aabb_min", c_float * 3),
                ("morton_aabb_max", c_float * 3), ("attenuation", c_uint32)]


FILTER_NEAREST, FILTER_LINEAR = 0, 1
ADDRESS_REPEAT, ADDRESS_MIRRORED_REPEAT, ADDRESS_CLAMP_TO_EDGE = 0, 1, 2


class Material(ctypes.Structure):
    _fields_ = [("base_color_texture_idx", c_uint32), ("metallic_roughness_texture_idx", c_uint32),
                ("metallic_factor", c_float), ("roughness_factor", c_float), ("base_color_factor", c_float * 4)]


class Texture(ctypes.Structure):
    _fields_ = [("texels", c_void_p), ("width", c_uint32), ("height", c_uint32), ("filter", c_uint32),
                ("address_u", c_uint32), ("address_v", c_uint32), ("pad", c_uint32)]


class MaterialSet(ctypes.Structure):
    _fields_ = [("texcoords", c_void_p), ("material_indices", c_void_p), ("materials", c_void_p),
                ("material_count", c_uint32), ("texture_count", c_uint32), ("textures", c_void_p)]


class FrameDesc(ctypes.Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("camera", Camera),
                ("depth", c_void_p), ("normals", c_void_p), ("albedo_roughness", c_void_p), ("metallic", c_void_p),
                ("light_positions", c_void_p), ("point_lights", c_void_p), ("light_count", c_uint32),
                ("directional_lights", c_void_p), ("directional_light_count", c_uint32), ("output", c_void_p),
                ("key_capacity", c_uint32), ("index_capacity", c_uint32), ("materials", P(MaterialSet))]


class ChainShape(ctypes.Structure):
    _fields_ = [("shade_pixels_per_thread", c_uint32), ("assign_waves_per_cu", c_uint32),
                ("light_sort_bits", c_uint32), ("light_stage_concurrent", c_uint32)]


class FrameOutputs(ctypes.Structure):
    _fields_ = [("view_positions", c_void_p), ("light_codes", c_void_p), ("bvh", c_void_p),
                ("bvh_length", c_uint32), ("cluster_keys", c_void_p), ("cluster_count", c_void_p),
                ("cluster_refs", c_void_p), ("key_capacity", c_uint32), ("light_counts", c_void_p),
                ("light_offsets", c_void_p), ("light_indices", c_void_p), ("index_capacity", c_uint32),
                ("assigned_total", c_void_p)]


# name -> (restype, argtypes); the declarations of include/vren_hip/capi.h
PROTOTYPES = {
    "vren_create": (c_int, [c_int, P(c_void_p)]),
    "vren_destroy": (None, [c_void_p]),
    "vren_status_string": (c_char_p, [c_int]),
    "vren_last_error": (c_char_p, [c_void_p]),
    "vren_sync_status": (c_int, [c_void_p, c_void_p]),
    "vren_cluster_cfg_default": (None, [P(ClusterCfg)]),
    "vren_reduce_u32": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_int, c_uint32]),
    "vren_reduce_vec4": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_int, c_uint32]),
    "vren_reduce_total_u32": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_void_p, c_int]),
    "vren_exclusive_scan_u32": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, c_uint32]),
    "vren_blelloch_downsweep_u32": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_int]),
    "vren_exclusive_scan_u32_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, c_uint32]),
    "vren_radix_sort_scratch_bytes": (c_size_t, [c_uint64, c_int]),
    "vren_radix_sort": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint32, c_void_p,
                                c_size_t]),
    "vren_radix_sort_config_count": (c_int, []),
    "vren_radix_sort_config_info": (c_int, [c_int, P(c_uint32), P(c_uint32), P(c_uint32), P(c_uint32)]),
    "vren_radix_sort_force_config": (c_int, [c_void_p, c_int]),
    "vren_autotune_radix_sort": (c_int, [c_void_p, c_void_p, c_uint64, c_int, c_uint32, P(c_int), P(c_float)]),
    "vren_autotune_save": (c_int, [c_void_p, c_char_p]),
    "vren_autotune_load": (c_int, [c_void_p, c_char_p]),
    "vren_bucket_sort": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_uint32, c_uint32]),
    "vren_bounce_point_lights": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, P(c_float), P(c_float),
                                         c_float, c_float]),
    "vren_calc_bvh_padded_leaf_count": (c_uint32, [c_uint32]),
    "vren_calc_bvh_buffer_length": (c_uint32, [c_uint32]),
    "vren_calc_bvh_buffer_size": (c_size_t, [c_uint32]),
    "vren_calc_bvh_root_index": (c_uint32, [c_uint32]),
    "vren_calc_bvh_level_count": (c_uint32, [c_uint32]),
    "vren_build_bvh": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
    "vren_build_lbvh": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "vren_camera_matrices": (None, [P(Camera), P(c_float), P(c_float)]),
    "vren_construct_point_light_bvh": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, P(Camera),
                                               P(ClusterCfg), c_void_p, c_void_p, c_void_p]),
    "vren_find_unique_cluster_list": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, P(Camera), P(ClusterCfg),
                                              c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p]),
    "vren_assign_lights": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, P(Camera), P(ClusterCfg), c_void_p,
                                   c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_uint32, c_void_p]),
    "vren_shade": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, P(Camera), P(ClusterCfg), c_void_p, c_void_p,
                           c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p,
                           c_void_p, c_uint32, c_void_p, c_uint32, c_void_p]),
    "vren_shade_materials": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, P(Camera), P(ClusterCfg), c_void_p,
                                     c_void_p, P(MaterialSet), c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
                                     c_uint32, c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, c_void_p]),
    "vren_depth_pyramid_level_count": (c_uint32, [c_uint32, c_uint32]),
    "vren_depth_pyramid_level_offset": (c_uint64, [c_uint32, c_uint32, c_uint32]),
    "vren_depth_pyramid_length": (c_uint64, [c_uint32, c_uint32]),
    "vren_depth_pyramid": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_void_p]),
    "vren_show_clusters": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, ctypes.c_int32, c_void_p, c_void_p,
                                   c_uint32, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p]),
    "vren_visualize_bvh": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p]),
    "vren_shade_backward": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, P(Camera), P(ClusterCfg), c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
                                    c_uint32, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
                                    c_void_p]),
    "vren_batch_create": (c_int, [c_void_p, P(FrameDesc), c_uint32, P(ClusterCfg), P(c_void_p)]),
    "vren_batch_run": (c_int, [c_void_p, c_void_p, c_int]),
    "vren_batch_run_profiled": (c_int, [c_void_p, c_void_p, P(c_float)]),
    "vren_batch_frame_outputs": (c_int, [c_void_p, c_uint32, P(FrameOutputs)]),
    "vren_batch_destroy": (None, [c_void_p]),
    "vren_batch_get_shape": (c_int, [c_void_p, P(ChainShape)]),
    "vren_batch_set_shape": (c_int, [c_void_p, P(ChainShape)]),
    "vren_batch_autotune": (c_int, [c_void_p, c_void_p, c_uint32, P(ChainShape), P(c_float)]),
    "vren_synth_keys": (c_int, [c_void_p, c_void_p, c_uint64, c_uint64, c_uint64, c_void_p]),
    "vren_synth_lights": (c_int, [c_void_p, c_void_p, c_uint64, c_uint32, c_void_p, c_void_p]),
    "vren_synth_frame": (c_int, [c_void_p, c_void_p, c_uint64, c_uint32, c_uint32, P(Camera), c_uint32, c_void_p,
                                 c_void_p]),
    "vren_synth_materials": (c_int, [c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_void_p]),
}


class VrenError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s (%d): %s" % (_status_name(status), status, message))
        self.status = status


def _status_name(status):
    return {VREN_EINVAL: "VREN_EINVAL", VREN_ECAPACITY: "VREN_ECAPACITY", VREN_EHIP: "VREN_EHIP",
            VREN_ENOMEM: "VREN_ENOMEM"}.get(status, "VREN_STATUS")


_lib = None


def load():
    """Load libvren_amd.so and attach prototypes for every symbol it exports (missing ones are skipped)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("vren_amd: %s is not built (run __graft_entry__.build() or vren_amd/build.py); there is "
                          "no fallback path" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
