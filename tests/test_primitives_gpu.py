"""GPU parity of the primitives through the C ABI, against the CPU oracle (bit-exact)."""
import numpy as np
import pytest

from gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

N_C1 = 1 << 20


@pytest.fixture(scope="module")
def pcg_keys(oracle):
    # 2^20 uint32 keys from pcg32(initstate 42, initseq 54)
    return oracle.pcg32(42, 54, N_C1)


@pytest.mark.parametrize("n", [1, 10, 1000, 10000, 100000, N_C1])
@pytest.mark.parametrize("op", [0, 1, 2])
def test_reduce_u32_tree(vctx, oracle, pcg_keys, n, op):
    x = pcg_keys[:n] if op != 0 or n != 10000 else np.ones(n, np.uint32)
    p = 1 << max(0, (n - 1).bit_length())
    out = empty_dev(4 * p, fill=0xAB)
    vctx.reduce_u32(to_dev(x), n, out, op)
    got = from_dev(out, np.uint32, (p,))
    assert np.array_equal(got, oracle.reduce_tree_u32(x, op))


def test_reduce_u32_multi_block_and_in_place(vctx, oracle, pcg_keys):
    x = pcg_keys[:3 * 1000]
    out = empty_dev(4 * 3 * 1024)
    vctx.reduce_u32(to_dev(x), 1000, out, 0, blocks=3)
    assert np.array_equal(from_dev(out, np.uint32, (3 * 1024,)), oracle.reduce_tree_u32(x, 0, blocks=3))
    y = pcg_keys[:4096]
    d = to_dev(y)
    vctx.reduce_u32(d, 4096, d, 0)
    assert np.array_equal(from_dev(d, np.uint32, (4096,)), oracle.reduce_tree_u32(y, 0))
    # segments longer than one 16384-element tile, at input offsets that are not 16-byte aligned (5001 * 4 bytes)
    for n, op in ((20000, 0), (5001, 2), (40001, 1)):
        x = pcg_keys[:3 * n]
        p = 1 << (n - 1).bit_length()
        out = empty_dev(4 * 3 * p, fill=0xCD)
        vctx.reduce_u32(to_dev(x), n, out, op, blocks=3)
        assert np.array_equal(from_dev(out, np.uint32, (3 * p,)), oracle.reduce_tree_u32(x, op, blocks=3)), n


def test_reduce_u32_two_passes_in_sequence(vctx, oracle, pcg_keys):
    # one segment of 16 384 < p <= 2^26 elements (two up-sweep passes, the second over the first's compact tile
    # totals): calls in a row with the scalar total in between (both use the context's scratch), a partial last
    # tile, in place, and 2^24 elements
    for n, op in ((1 << 20, 0), (100000, 2), ((1 << 16) + 3, 1), (1 << 20, 1)):
        x = pcg_keys[:n]
        p = 1 << (n - 1).bit_length()
        out = empty_dev(4 * p, fill=0xEF)
        vctx.reduce_u32(to_dev(x), n, out, op)
        assert np.array_equal(from_dev(out, np.uint32, (p,)), oracle.reduce_tree_u32(x, op)), (n, op)
        tot = empty_dev(4)
        vctx.reduce_total_u32(to_dev(x), n, tot, op)
        assert from_dev(tot, np.uint32)[0] == oracle.reduce_total_u32(x, op)
    d = to_dev(pcg_keys)
    vctx.reduce_u32(d, N_C1, d, 0)
    assert np.array_equal(from_dev(d, np.uint32, (N_C1,)), oracle.reduce_tree_u32(pcg_keys, 0))
    big = oracle.pcg32(7, 54, 1 << 24)
    out = empty_dev(4 << 24)
    vctx.reduce_u32(to_dev(big), 1 << 24, out, 0)
    assert np.array_equal(from_dev(out, np.uint32, (1 << 24,)), oracle.reduce_tree_u32(big, 0))


@pytest.mark.parametrize("op", [0, 1, 2])
def test_reduce_vec4_tree_bit_exact(vctx, oracle, op):
    # random (non-integer) floats: bit-exact only if the combination tree is the reference's
    rng = np.random.default_rng(11)
    for n in (1, 100, 10000):
        x = rng.uniform(-100, 100, (n, 4)).astype(np.float32)
        p = 1 << max(0, (n - 1).bit_length())
        out = empty_dev(16 * p)
        vctx.reduce_vec4(to_dev(x), n, out, op)
        got = from_dev(out, np.float32, (p, 4))
        assert np.array_equal(got.view(np.uint32), oracle.reduce_tree_vec4(x, op).view(np.uint32)), n


@pytest.mark.parametrize("op", [0, 1, 2])
def test_reduce_total(vctx, oracle, pcg_keys, op):
    for n in (1, 5, 4099, N_C1):
        out = empty_dev(4)
        vctx.reduce_total_u32(to_dev(pcg_keys[:n]), n, out, op)
        assert from_dev(out, np.uint32)[0] == oracle.reduce_total_u32(pcg_keys[:n], op)


@pytest.mark.parametrize("n", [1, 7, 4096, 4097, 100003, N_C1])
def test_exclusive_scan(vctx, oracle, pcg_keys, n):
    x = pcg_keys[:n]
    out = empty_dev(4 * n)
    vctx.exclusive_scan_u32(to_dev(x), out, n)
    assert np.array_equal(from_dev(out, np.uint32, (n,)), oracle.exclusive_scan_u32(x))


def test_exclusive_scan_segments_in_place_and_ones(vctx, oracle, pcg_keys):
    x = pcg_keys[:3 * 5000]
    d = to_dev(x)
    vctx.exclusive_scan_u32(d, d, 5000, blocks=3)
    assert np.array_equal(from_dev(d, np.uint32, (15000,)), oracle.exclusive_scan_u32(x, blocks=3))
    # several 16384-key tiles per segment, segments at offsets that are not 16-byte aligned (40001 * 4 bytes)
    x = pcg_keys[:3 * 40001]
    o = empty_dev(4 * 3 * 40001)
    vctx.exclusive_scan_u32(to_dev(x), o, 40001, blocks=3)
    assert np.array_equal(from_dev(o, np.uint32, (3 * 40001,)), oracle.exclusive_scan_u32(x, blocks=3))
    for e in range(0, 20):  # blelloch_scan.main: all ones, N = 2^0 .. 2^19
        n = 1 << e
        o = empty_dev(4 * n)
        vctx.exclusive_scan_u32(to_dev(np.ones(n, np.uint32)), o, n)
        assert np.array_equal(from_dev(o, np.uint32, (n,)), np.arange(n, dtype=np.uint32))


@pytest.mark.parametrize("blocks", [1, 16])
@pytest.mark.parametrize("length", [16, 1024, 1 << 12, 1 << 17, 1 << 20])
def test_reduce_then_downsweep_is_exclusive_scan(vctx, oracle, pcg_keys, length, blocks):
    """The reference's radix-sort use (radix_sort.cpp:230-289): reduce (add, in place) over `blocks` segments, then
    downsweep(clear_last = true): every segment holds std::exclusive_scan of its keys."""
    n = length * blocks
    x = np.resize(pcg_keys, n)
    d = to_dev(x)
    vctx.reduce_u32(d, length, d, 0, blocks)
    vctx.blelloch_downsweep_u32(d, length, blocks, True)
    assert np.array_equal(from_dev(d, np.uint32, (n,)), oracle.exclusive_scan_u32(x, blocks=blocks))


@pytest.mark.parametrize("length", [1, 2, 4, 8, 16, 32, 512, 1024, 2048, 4096, 8192, 1 << 16, 1 << 22])
def test_downsweep_arbitrary_input_vs_oracle(vctx, oracle, pcg_keys, length):
    """On arbitrary input (not an up-sweep tree), both clear_last settings, 1 and 3 segments (the second and third
    not 16-byte aligned when length < 4): bit for bit against the oracle's level-by-level restatement."""
    for blocks in (1, 3):
        n = length * blocks
        x = np.resize(pcg_keys[::-1], n)
        for clear in (True, False):
            d = to_dev(x)
            vctx.blelloch_downsweep_u32(d, length, blocks, clear)
            want = oracle.blelloch_downsweep_u32(x, length, blocks, clear)
            assert np.array_equal(from_dev(d, np.uint32, (n,)), want), (blocks, clear)


def _sort_gpu(vctx, keys, vals, begin=0, end=32):
    n = keys.size
    dk = to_dev(keys)
    dv = to_dev(vals) if vals is not None else None
    scratch = empty_dev(vctx.radix_sort_scratch_bytes(n, vals is not None))
    vctx.radix_sort(dk, dv, n, scratch, begin, end)
    k = from_dev(dk, np.uint32, (n,))
    v = from_dev(dv, np.uint32, (n,)) if dv is not None else None
    return k, v


def test_radix_sort_pcg_keys_kv(vctx, oracle, pcg_keys):
    vals = np.arange(N_C1, dtype=np.uint32)
    k, v = _sort_gpu(vctx, pcg_keys, vals)
    ok, ov = oracle.stable_sort_kv(pcg_keys, vals)
    assert np.array_equal(k, ok) and np.array_equal(v, ov)


def test_radix_sort_reference_case_and_edges(vctx, oracle, pcg_keys):
    # radix_sort.main: N = 1024, reversed iota, vs std::sort
    k, _ = _sort_gpu(vctx, np.arange(1024, dtype=np.uint32)[::-1].copy(), None)
    assert np.array_equal(k, np.arange(1024, dtype=np.uint32))
    rng = np.random.default_rng(5)
    for n in (1, 2, 33, 5000, N_C1 + 7):
        keys = rng.integers(0, 50, n, dtype=np.uint32) if n == 5000 else rng.integers(0, 2 ** 32, n,
                                                                                        dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        k, v = _sort_gpu(vctx, keys, vals)
        ok, ov = oracle.stable_sort_kv(keys, vals)
        assert np.array_equal(k, ok) and np.array_equal(v, ov), n
    keys = pcg_keys[:70000]
    for b, e in ((0, 15), (4, 12), (16, 32), (0, 30)):
        k, v = _sort_gpu(vctx, keys, np.arange(keys.size, dtype=np.uint32), b, e)
        ok, ov = oracle.stable_sort_kv(keys, np.arange(keys.size, dtype=np.uint32), b, e)
        assert np.array_equal(v, ov), (b, e)


def test_radix_sort_every_config_identical(vctx, oracle, pcg_keys):
    keys = pcg_keys[:100003]
    vals = np.arange(keys.size, dtype=np.uint32)
    ok, ov = oracle.stable_sort_kv(keys, vals)
    configs = vctx.radix_configs()
    try:
        for i, cfg in enumerate(configs):
            vctx.radix_force_config(i)
            k, v = _sort_gpu(vctx, keys, vals)
            assert np.array_equal(k, ok) and np.array_equal(v, ov), cfg
            k, _ = _sort_gpu(vctx, keys, None)
            assert np.array_equal(k, ok), cfg
    finally:
        vctx.radix_force_config(-1)


def test_radix_sort_headline_size_properties(vctx, oracle):
    # 2^26 pairs (the headline size): checked through size-independent properties on the device
    import torch
    n = 1 << 26
    keys = oracle.pcg32(42, 54, n)
    dk = to_dev(keys).view(torch.int32)
    orig = dk.clone()
    dv = torch.arange(n, dtype=torch.int32, device="cuda")
    scratch = empty_dev(vctx.radix_sort_scratch_bytes(n, True))
    vctx.radix_sort(dk, dv, n, scratch)
    torch.cuda.synchronize()
    ku = dk.to(torch.int64) & 0xFFFFFFFF
    assert bool((ku[1:] >= ku[:-1]).all())                                   # sorted
    seen = torch.zeros(n, dtype=torch.int32, device="cuda")
    seen.index_add_(0, dv.long(), torch.ones(n, dtype=torch.int32, device="cuda"))
    assert bool((seen == 1).all())                                           # values are a permutation
    assert bool((orig[dv.long()] == dk).all())                               # each key travelled with its value
    eq = ku[1:] == ku[:-1]
    assert bool((dv[1:][eq] > dv[:-1][eq]).all())                            # stable: ties keep input order
    del seen, orig


@pytest.mark.parametrize("n_buckets,shift,keys_kind", [(65536, 0, "u16"), (64, 24, "u30"), (1, 0, "u16"),
                                                        (65536, 24, "u32"), (4096, 7, "u32"), (65536, 0, "few")])
def test_bucket_sort(vctx, oracle, n_buckets, shift, keys_kind):
    # bucket_sort.main (vren_test/vren_test/primitives/bucket_sort.cpp:182-192): N = 1, then x(4..6) while < 10^7,
    # key = rand() % 65536, value = i -- here seeded PCG32 draws and a fixed growth sequence. "u32": full 32-bit
    # keys with random values (the bucket field of (65536, 24) runs past bit 31 and holds 8 bits)
    n = 1
    sizes = []
    while n < 3_000_000:
        sizes.append(n)
        n *= 5
    for n in sizes:
        raw = oracle.pcg32(4242, 54, n)
        if keys_kind == "u16":
            keys, vals = raw % 65536, np.arange(n, dtype=np.uint32)
        elif keys_kind == "few":
            # three keys far apart: long runs of one bucket and long gaps of empty buckets between them, inside a
            # tile's digit run and across tiles
            keys, vals = (raw % 3) * 20011 + 7, np.arange(n, dtype=np.uint32)
        elif keys_kind == "u30":
            keys, vals = raw >> 2, np.arange(n, dtype=np.uint32)  # 30-bit codes for the 64 coarse buckets
        else:
            keys, vals = raw, oracle.pcg32(77, 3, n)
        pairs = np.stack([keys.astype(np.uint32), vals.astype(np.uint32)], axis=1)
        out = empty_dev(8 * n)
        offs = empty_dev(4 * n_buckets)
        vctx.bucket_sort(to_dev(pairs), n, out, offs, n_buckets, shift)
        opairs, ooffs = oracle.bucket_sort(pairs, n_buckets, shift)
        assert np.array_equal(from_dev(out, np.uint32, (n, 2)), opairs), n
        assert np.array_equal(from_dev(offs, np.uint32, (n_buckets,)), ooffs), n
    # no pairs: every bucket starts at 0
    offs = to_dev(np.full(n_buckets, 0xFFFFFFFF, np.uint32))
    vctx.bucket_sort(empty_dev(16), 0, empty_dev(16), offs, n_buckets, shift)
    assert not from_dev(offs, np.uint32, (n_buckets,)).any()


@pytest.mark.parametrize("n", [0, 1, 10, 100, 1000, 10000])
def test_build_bvh_standalone(vctx, oracle, n):
    # build_bvh.main (vren_test/vren_test/primitives/build_bvh.cpp:213-221): random AABBs, pos and extent U[0,100]
    from test_oracle_primitives import BVH_DTYPE, random_leaves
    rng = np.random.default_rng(100 + n)
    padded = oracle.calc("padded_leaf_count", n)
    length = oracle.calc("buffer_length", n)
    nodes = random_leaves(n, padded, length, rng)
    d = to_dev(nodes)
    vctx.build_bvh(d, padded)
    got = from_dev(d, BVH_DTYPE, (length,))
    want = oracle.build_bvh(nodes, padded)
    assert np.array_equal(got.view(np.uint32).reshape(-1, 8)[:, :7], want.view(np.uint32).reshape(-1, 8)[:, :7])
