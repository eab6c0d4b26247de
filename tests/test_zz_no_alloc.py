"""The allocation-free exclusive scan (vren_exclusive_scan_u32_noalloc) through the C ABI: bit-exact against the
allocating entry point (vren_exclusive_scan_u32) and the oracle, on two calls in a row (the second smaller, so
leftover workspace words from the first would show), with sizes beyond the fixed workspace rejected before any
launch. Runs alone in a fresh process and after the other GPU tests."""
import ctypes

import numpy as np
import pytest

from gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

EINVAL = -1
TILE = 16384


def _scan_both(vctx, oracle, keys, n, blocks):
    import torch
    lib, h = vctx.lib, vctx.h
    s = vctx.stream()
    din = to_dev(keys)
    a = empty_dev(4 * keys.size, fill=0xAB)
    b = empty_dev(4 * keys.size, fill=0xCD)
    print("noalloc scan n=%d blocks=%d" % (n, blocks), flush=True)
    assert lib.vren_exclusive_scan_u32_noalloc(h, s, ctypes.c_void_p(din.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                                               n, blocks) == 0, lib.vren_last_error(h)
    assert lib.vren_exclusive_scan_u32(h, s, ctypes.c_void_p(din.data_ptr()), ctypes.c_void_p(b.data_ptr()), n,
                                       blocks) == 0, lib.vren_last_error(h)
    torch.cuda.synchronize()
    ga, gb = from_dev(a, np.uint32, (keys.size,)), from_dev(b, np.uint32, (keys.size,))
    assert np.array_equal(ga, gb)
    assert np.array_equal(ga, oracle.exclusive_scan_u32(keys, blocks))
    return din


def test_noalloc_scan_matches_allocating_scan(vctx, oracle):
    # first call: 3 segments of 5,000,001 keys (306 tiles each, a partial last tile) = 918 workspace tiles
    n1, b1 = 5_000_001, 3
    k1 = oracle.pcg32(42, 54, n1 * b1)
    _scan_both(vctx, oracle, k1, n1, b1)
    # second call, smaller: the first call's status words beyond this call's tiles are stale, and must not matter
    n2, b2 = 70_001, 1
    k2 = oracle.pcg32(43, 54, n2 * b2)
    _scan_both(vctx, oracle, k2, n2, b2)
    # at the workspace limit: 65536 segments of 16 keys, one tile each = 2^16 tiles
    n3, b3 = 16, 65536
    k3 = oracle.pcg32(44, 54, n3 * b3)
    _scan_both(vctx, oracle, k3, n3, b3)


def test_noalloc_scan_rejects_sizes_beyond_the_workspace(vctx):
    lib, h = vctx.lib, vctx.h
    buf = ctypes.c_void_p(4096)   # never dereferenced: each call fails its size check before any launch
    s = vctx.stream()
    assert lib.vren_exclusive_scan_u32_noalloc(h, s, buf, buf, 16, 65537) == EINVAL
    assert b"fixed workspace" in lib.vren_last_error(h)
    assert lib.vren_exclusive_scan_u32_noalloc(h, s, buf, buf, TILE * 65536 + 1, 1) == EINVAL
    assert lib.vren_exclusive_scan_u32_noalloc(h, s, buf, buf, TILE * 2 + 1, 1 << 15) == EINVAL
    assert lib.vren_exclusive_scan_u32_noalloc(h, s, buf, buf, 0, 1) == EINVAL
    assert lib.vren_exclusive_scan_u32_noalloc(None, s, buf, buf, 16, 1) == EINVAL
