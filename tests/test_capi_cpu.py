"""The C ABI without a GPU: the library loads, exports every entry point the headers declare, its host helpers agree
with the oracle, and every entry point validates its arguments (a host-only context, vren_create(-1), returns
VREN_EINVAL for bad arguments and for valid ones, since there is no device to run on)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EINVAL = -1


@pytest.fixture(scope="module")
def lib():
    import vren_amd
    return vren_amd.lib()


@pytest.fixture(scope="module")
def hctx(lib):
    h = ctypes.c_void_p()
    assert lib.vren_create(-1, ctypes.byref(h)) == 0
    yield h
    lib.vren_destroy(h)


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for dirpath, _, files in os.walk(inc):
        for f in files:
            if f.endswith(".h"):
                text = open(os.path.join(dirpath, f)).read()
                names.update(re.findall(r"VREN_EXPORT\s+[\w\s\*]+?\b(vren_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(lib):
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_the_abi():
    from vren_amd import _lib
    missing = [n for n in declared_symbols() if n not in _lib.PROTOTYPES]
    assert not missing, missing


def test_status_strings(lib):
    assert lib.vren_status_string(0) == b"ok"
    assert lib.vren_status_string(-1) == b"invalid argument"
    assert lib.vren_status_string(-2) == b"output capacity exceeded"


def test_calc_bvh_known_answers_library(lib, oracle):
    # build_bvh.utils (vren_test/vren_test/primitives/build_bvh.cpp:223-254), through the product's host helpers
    for n in (0, 1, 17, 32, 129, 582, 1024, 4096, 65536, 2193819):
        assert lib.vren_calc_bvh_padded_leaf_count(n) == oracle.calc("padded_leaf_count", n)
        assert lib.vren_calc_bvh_buffer_length(n) == oracle.calc("buffer_length", n)
        assert lib.vren_calc_bvh_root_index(n) == oracle.calc("root_index", n)
        assert lib.vren_calc_bvh_level_count(n) == oracle.calc("level_count", n)
        assert lib.vren_calc_bvh_buffer_size(n) == 32 * oracle.calc("buffer_length", n)
    from test_oracle_primitives import golden
    kat = golden("build_bvh_utils.json")
    for fn in ("padded_leaf_count", "buffer_length", "root_index", "level_count"):
        for n, v in kat[fn]:
            assert getattr(lib, "vren_calc_bvh_" + fn)(n) == v, (fn, n)


CAMERAS = [dict(W=1920, H=1080), dict(W=640, H=360, fov_deg=60.0, pos=(3.0, -2.0, 7.5)),
           dict(W=1000, H=700, fov_deg=75.0, yaw=0.3, pitch=-0.2, near=0.5, far=250.0)]


@pytest.mark.parametrize("spec", CAMERAS)
def test_camera_matrices_match_oracle(lib, oracle, spec):
    co = ol.ClusteredOracle(oracle)
    s = dict(spec)
    W, H = s.pop("W"), s.pop("H")
    cam = ol.camera(W, H, **s)
    v = (ctypes.c_float * 16)()
    p = (ctypes.c_float * 16)()
    lib.vren_camera_matrices(ctypes.byref(cam), v, p)
    ov, op = co.camera_matrices(cam)
    assert np.array_equal(np.array(v, np.float32), ov)
    assert np.array_equal(np.array(p, np.float32), op)
    if spec.get("yaw", 0.0) == 0.0 and spec.get("pitch", 0.0) == 0.0:
        # closed form of camera.cpp:32-50 for an axis-aligned camera: view = translate(-position)
        pos = spec.get("pos", (0.0, 0.0, -50.0))
        want = np.eye(4, dtype=np.float32)
        want[3, :3] = -np.array(pos, np.float32)
        assert np.array_equal(ov.reshape(4, 4), want)
        t = np.float32(np.tan(np.float64(np.float32(cam.fov_y) / np.float32(2))))
        assert op[5] == np.float32(1) / t and op[11] == 1.0


def test_cfg_default(lib):
    import vren_amd
    cfg = vren_amd.default_cfg()
    assert (cfg.key_mode, cfg.cluster_aabb_mode, cfg.morton_bits, cfg.morton_aabb_mode, cfg.attenuation) == \
        (0, 0, 5, 0, 0)


def test_primitive_argument_validation(lib, hctx):
    P = ctypes.c_void_p
    buf = P(4096)  # never dereferenced: every call below fails validation or has no device
    # reduce: n == 0, blocks == 0, blocks > 65535, bad op, then valid arguments without a device
    assert lib.vren_reduce_u32(hctx, None, buf, 0, buf, 0, 1) == EINVAL
    assert lib.vren_reduce_u32(hctx, None, buf, 10, buf, 0, 0) == EINVAL
    assert lib.vren_reduce_u32(hctx, None, buf, 10, buf, 0, 70000) == EINVAL
    assert lib.vren_reduce_u32(hctx, None, buf, 10, buf, 7, 1) == EINVAL
    assert lib.vren_reduce_u32(hctx, None, buf, 10, buf, 0, 1) == EINVAL
    assert b"no device" in lib.vren_last_error(hctx)
    assert lib.vren_reduce_u32(hctx, None, buf, 10, buf, 0, 2) == EINVAL  # in place, 2 blocks, n not a power of 2
    assert lib.vren_reduce_vec4(hctx, None, None, 10, buf, 0, 1) == EINVAL
    assert lib.vren_reduce_total_u32(hctx, None, buf, 0, buf, 0) == EINVAL
    assert lib.vren_exclusive_scan_u32(hctx, None, buf, buf, 0, 1) == EINVAL
    # allocation-free scan: beyond its fixed workspace of 2^16 tiles of 16384 keys, then no device
    assert lib.vren_exclusive_scan_u32_noalloc(hctx, None, buf, buf, 16, 65537) == EINVAL
    assert b"fixed workspace" in lib.vren_last_error(hctx)
    assert lib.vren_exclusive_scan_u32_noalloc(hctx, None, buf, buf, 16, 65536) == EINVAL
    assert b"no device" in lib.vren_last_error(hctx)
    # radix sort: n == 0, n >= 2^30, bad bit range, too little scratch (radix_sort.cpp:158-161 throws on bad length)
    sb = lib.vren_radix_sort_scratch_bytes(1000, 1)
    assert lib.vren_radix_sort(hctx, None, buf, buf, 0, 0, 32, buf, sb) == EINVAL
    assert lib.vren_radix_sort(hctx, None, buf, buf, 1 << 30, 0, 32, buf, sb) == EINVAL
    assert lib.vren_radix_sort(hctx, None, buf, buf, 1000, 8, 4, buf, sb) == EINVAL
    assert lib.vren_radix_sort(hctx, None, buf, buf, 1000, 0, 33, buf, sb) == EINVAL
    assert lib.vren_radix_sort(hctx, None, buf, buf, 1000, 0, 32, buf, sb - 1) == EINVAL
    assert b"scratch" in lib.vren_last_error(hctx)
    assert lib.vren_radix_sort(hctx, None, buf, buf, 1000, 0, 32, buf, sb) == EINVAL
    assert b"no device" in lib.vren_last_error(hctx)
    # bucket sort: bucket count must be a power of 2 in [1, 65536]
    for nb in (0, 3, 65537, 1 << 17):
        assert lib.vren_bucket_sort(hctx, None, buf, 10, P(8192), buf, nb, 0) == EINVAL
    assert lib.vren_bucket_sort(hctx, None, buf, 10, buf, buf, 64, 0) == EINVAL  # in place
    # the radix passes load and store 8-byte pairs
    assert lib.vren_bucket_sort(hctx, None, P(4100), 10, P(8192), buf, 64, 0) == EINVAL
    assert b"8-byte aligned" in lib.vren_last_error(hctx)
    assert lib.vren_bucket_sort(hctx, None, buf, 10, P(8196), buf, 64, 0) == EINVAL
    assert b"8-byte aligned" in lib.vren_last_error(hctx)
    assert lib.vren_bucket_sort(hctx, None, buf, 10, P(8192), buf, 64, 0) == EINVAL
    assert b"no device" in lib.vren_last_error(hctx)
    # bounce_point_lights: non-null AABB and (for n > 0) buffers, 16-byte aligned vec4 buffers
    box = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    assert lib.vren_bounce_point_lights(hctx, None, buf, buf, 10, None, box, 1.0, 1.0) == EINVAL
    assert lib.vren_bounce_point_lights(hctx, None, None, buf, 10, box, box, 1.0, 1.0) == EINVAL
    assert lib.vren_bounce_point_lights(hctx, None, P(4104), buf, 10, box, box, 1.0, 1.0) == EINVAL
    assert b"16-byte aligned" in lib.vren_last_error(hctx)
    assert lib.vren_bounce_point_lights(hctx, None, buf, buf, 10, box, box, 1.0, 1.0) == EINVAL
    assert b"no device" in lib.vren_last_error(hctx)
    # build_lbvh: at least one leaf, at most 2^20, non-null buffers
    assert lib.vren_build_lbvh(hctx, None, buf, 0, buf, buf, buf, buf) == EINVAL
    assert lib.vren_build_lbvh(hctx, None, buf, (1 << 20) + 1, buf, buf, buf, buf) == EINVAL
    assert lib.vren_build_lbvh(hctx, None, None, 10, buf, buf, buf, buf) == EINVAL
    assert lib.vren_build_lbvh(hctx, None, buf, 10, buf, buf, buf, buf) == EINVAL
    assert b"no device" in lib.vren_last_error(hctx)
    # build_bvh: leaf count must be a power of 32 and >= 32 (build_bvh.cpp:45)
    for n in (0, 1, 31, 33, 64, 1000):
        assert lib.vren_build_bvh(hctx, None, buf, n) == EINVAL
    assert lib.vren_build_bvh(hctx, None, buf, 1024) == EINVAL
    assert b"no device" in lib.vren_last_error(hctx)


def test_radix_sort_configs_listed(lib):
    import vren_amd
    cfgs = vren_amd.Context.radix_configs()
    assert len(cfgs) == lib.vren_radix_sort_config_count() >= 20
    assert {c[0] for c in cfgs} == {4, 6, 8}
    assert {c[1] for c in cfgs} == {8, 15, 23}
    assert {c[2] for c in cfgs} == {4, 8, 16}


def test_clustered_argument_validation(lib, hctx):
    import vren_amd
    P = ctypes.c_void_p
    b = P(4096)
    good = ol.camera(640, 360)
    cfg = vren_amd.default_cfg()
    bad_cams = [ol.camera(640, 360, near=0.0), ol.camera(640, 360, near=10.0, far=5.0), ol.camera(640, 360, fov_deg=0)]
    bad_cfgs = [vren_amd.default_cfg(morton_bits=0), vren_amd.default_cfg(morton_bits=11),
                vren_amd.default_cfg(key_mode=1, depth_slices=0), vren_amd.default_cfg(key_mode=5),
                vren_amd.default_cfg(attenuation=9), vren_amd.default_cfg(cluster_aabb_mode=3)]
    R = ctypes.byref

    def fuc(W, H, cam, cf, key_cap=100, keys=b):
        return lib.vren_find_unique_cluster_list(hctx, None, W, H, R(cam), R(cf), b, b, keys, key_cap, b, b)

    for cam in bad_cams:
        assert fuc(640, 360, cam, cfg) == EINVAL
        assert lib.vren_construct_point_light_bvh(hctx, None, b, b, 10, R(cam), R(cfg), b, b, b) == EINVAL
    for cf in bad_cfgs:
        assert fuc(640, 360, good, cf) == EINVAL
    assert fuc(0, 360, good, cfg) == EINVAL
    assert fuc(9000, 360, good, cfg) == EINVAL
    assert fuc(640, 360, good, cfg, key_cap=0) == EINVAL
    assert fuc(640, 360, good, cfg, keys=None) == EINVAL
    assert fuc(640, 360, good, cfg) == EINVAL and b"no device" in lib.vren_last_error(hctx)
    # more than 2^20 lights (config.hpp:18)
    assert lib.vren_construct_point_light_bvh(hctx, None, b, b, (1 << 20) + 1, R(good), R(cfg), b, b, b) == EINVAL
    assert lib.vren_construct_point_light_bvh(hctx, None, b, b, 0, R(good), R(cfg), b, b, b) == 0  # no lights: no-op
    assert lib.vren_assign_lights(hctx, None, 640, 360, R(good), R(cfg), b, b, 0, b, 10, b, b, b, b, b, 10, b) == EINVAL
    assert lib.vren_assign_lights(hctx, None, 640, 360, R(good), R(cfg), b, b, 10, b, 10, b, b, b, b, b, 0, b) == EINVAL
    assert lib.vren_shade(hctx, None, 640, 360, R(good), R(cfg), b, b, b, b, None, 10, b, b, b, 10, b, b, 10, None, 0,
                          b) == EINVAL
    assert lib.vren_shade(hctx, None, 640, 360, R(good), R(cfg), b, b, b, b, b, 10, b, b, b, 10, b, b, 10, None, 2,
                          b) == EINVAL  # directional lights without a buffer
    assert lib.vren_shade_backward(hctx, None, 640, 360, R(good), R(cfg), b, b, b, b, b, 10, b, b, b, 10, b, b, 10,
                                   None, b, b, b) == EINVAL
    # batch: no frames / a frame without a depth buffer
    out = ctypes.c_void_p()
    assert lib.vren_batch_create(hctx, None, 0, R(cfg), R(out)) == EINVAL
    fd = vren_amd.FrameDesc()
    fd.width, fd.height, fd.camera = 640, 360, good
    fd.normals = fd.albedo_roughness = fd.metallic = fd.output = 4096
    arr = (vren_amd.FrameDesc * 1)(fd)
    assert lib.vren_batch_create(hctx, arr, 1, R(cfg), R(out)) == EINVAL
    assert b"null G-buffer" in lib.vren_last_error(hctx)


def test_synth_host_validation(lib, hctx):
    cam = ol.camera(64, 64)
    assert lib.vren_synth_frame(hctx, None, 1, 64, 64, ctypes.byref(cam), 300, ctypes.c_void_p(16),
                                ctypes.c_void_p(16)) == EINVAL  # more than 256 boxes
    assert lib.vren_synth_keys(hctx, None, 42, 54, 0, None) == 0
