"""numpy-facing ctypes binding of the CPU oracle (oracle/_build/liboracle.so). Test infrastructure only."""
import ctypes
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

c_u32, c_u64, c_int, c_f, c_vp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(c_vp)


def _ensure_built():
    if not os.path.exists(LIB):
        subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "build.py")], check=True)


class Oracle:
    def __init__(self):
        _ensure_built()
        self.lib = ctypes.CDLL(LIB)
        L = self.lib
        for name, res in [("o_logf", c_f), ("o_expf", c_f)]:
            getattr(L, name).restype = res
            getattr(L, name).argtypes = [c_f]
        L.o_powi.restype = c_f
        L.o_powi.argtypes = [c_f, c_u32]
        L.o_reduce_total_u32.restype = c_u32
        for n in ["o_calc_bvh_padded_leaf_count", "o_calc_bvh_buffer_length", "o_calc_bvh_root_index",
                  "o_calc_bvh_level_count"]:
            getattr(L, n).restype = c_u32
            getattr(L, n).argtypes = [c_u32]
        L.o_find_unique_clusters.restype = c_u32
        L.o_assign_lights.restype = c_u64
        L.o_assign_lights_brute.restype = c_u64

    # rng ---------------------------------------------------------------------------------------------------
    def pcg32(self, initstate, initseq, n, start=0, jump=False):
        out = np.zeros(n, np.uint32)
        fn = self.lib.o_pcg32_fill_at if jump else self.lib.o_pcg32_fill
        fn(c_u64(initstate), c_u64(initseq), c_u64(start), _p(out), c_u64(n))
        return out

    def logf(self, x):
        return self.lib.o_logf(float(x))

    def expf(self, x):
        return self.lib.o_expf(float(x))

    def powi(self, a, k):
        return self.lib.o_powi(float(a), c_u32(k))

    # primitives --------------------------------------------------------------------------------------------
    def reduce_tree_u32(self, x, op, blocks=1):
        x = np.ascontiguousarray(x, np.uint32)
        n = x.size // blocks
        p = 1 << max(0, (n - 1).bit_length())
        out = np.zeros(p * blocks, np.uint32)
        self.lib.o_reduce_tree_u32(_p(x), c_u32(n), _p(out), c_int(op), c_u32(blocks))
        return out

    def reduce_tree_vec4(self, x, op, blocks=1):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        n = x.shape[0] // blocks
        p = 1 << max(0, (n - 1).bit_length())
        out = np.zeros((p * blocks, 4), np.float32)
        self.lib.o_reduce_tree_vec4(_p(x), c_u32(n), _p(out), c_int(op), c_u32(blocks))
        return out

    def reduce_total_u32(self, x, op):
        x = np.ascontiguousarray(x, np.uint32)
        return self.lib.o_reduce_total_u32(_p(x), c_u64(x.size), c_int(op))

    def exclusive_scan_u32(self, x, blocks=1):
        x = np.ascontiguousarray(x, np.uint32)
        out = np.zeros_like(x)
        self.lib.o_exclusive_scan_u32(_p(x), _p(out), c_u64(x.size // blocks), c_u32(blocks))
        return out

    def stable_sort_kv(self, keys, vals=None, begin_bit=0, end_bit=32):
        k = np.array(keys, np.uint32, copy=True)
        v = None if vals is None else np.array(vals, np.uint32, copy=True)
        self.lib.o_stable_sort_kv(_p(k), _p(v), c_u64(k.size), c_u32(begin_bit), c_u32(end_bit))
        return k, v

    def bucket_sort(self, pairs, n_buckets=65536, shift=0):
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        out = np.zeros_like(pairs)
        offs = np.zeros(n_buckets, np.uint32)
        self.lib.o_bucket_sort(_p(pairs), c_u32(pairs.shape[0]), _p(out), _p(offs), c_u32(n_buckets), c_u32(shift))
        return out, offs

    def calc(self, name, n):
        return getattr(self.lib, "o_calc_bvh_" + name)(c_u32(n))

    def build_bvh(self, nodes, padded):
        nodes = np.array(nodes, copy=True)
        self.lib.o_build_bvh(_p(nodes), c_u32(padded))
        return nodes
