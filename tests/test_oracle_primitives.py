"""Pins of the CPU oracle's primitives, independent of the oracle's own code.

Each check uses something the oracle does not: the published PCG32 demo vector, numpy's own sort/scan/min/max,
the closed form of the Blelloch up-sweep tree, the build_bvh.utils known answers of the reference's test
(vren_test/vren_test/primitives/build_bvh.cpp:223-254) and its point-query property test (:175-210).
"""
import math

import numpy as np
import pytest

BVH_DTYPE = np.dtype([("min", "<f4", 3), ("next", "<u4"), ("max", "<f4", 3), ("pad", "<u4")])


def test_pcg32_published_vector(oracle):
    # pcg32-demo, initstate 42, initseq 54 (O'Neill's reference implementation output), from tests/golden
    g = golden("pcg32_demo.json")
    assert list(oracle.pcg32(g["initstate"], g["initseq"], len(g["outputs"]))) == g["outputs"]


def test_pcg32_jump_ahead_matches_sequential(oracle):
    seq = oracle.pcg32(1042, 55, 5000)
    for start in (0, 1, 17, 4095):
        jmp = oracle.pcg32(1042, 55, 64, start=start, jump=True)
        assert np.array_equal(jmp, seq[start:start + 64])


def _ulp_err(got, ref):
    ref32 = np.float32(ref)
    return abs(float(got) - ref) / float(np.spacing(np.abs(ref32)))


def test_vmath_logf_error_bound(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-3, 1e3, 4000), 1.0 + rng.uniform(-1e-3, 1e-3, 2000),
                         [1e-40, 1e-45, 3.4e38, 2.0, 0.5, 1.0000001]]).astype(np.float32)
    worst = 0.0
    for x in xs:
        r = math.log(float(x))
        g = oracle.logf(x)
        if r == 0.0:
            assert g == 0.0
            continue
        worst = max(worst, _ulp_err(g, r))
    assert worst <= 2.0, worst
    assert oracle.logf(1.0) == 0.0
    assert math.isinf(oracle.logf(0.0)) and oracle.logf(0.0) < 0
    assert math.isnan(oracle.logf(-1.0))


def test_vmath_logf_nondecreasing_over_all_nonnegative_floats(oracle):
    """vren_logf(b) <= vren_logf(b + 1) for every bit pattern b of [+0, +inf): the GPU's depth-slice threshold
    table equals the slice formula only if the log never decreases. Exhaustive, 8 threads (the scan releases the
    GIL); a planted glitch must be reported at exactly its bit pattern."""
    from concurrent.futures import ThreadPoolExecutor
    assert oracle.logf_first_decrease(0x3F000000, 0x3F010000) == 0xFFFFFFFF
    assert oracle.logf_first_decrease(0x3F000000, 0x3F010000, planted=0x3F00ABCD) == 0x3F00ABCD
    end = 0x7F800000
    chunks = [(i * end // 8, (i + 1) * end // 8) for i in range(8)]
    with ThreadPoolExecutor(8) as ex:
        found = list(ex.map(lambda c: oracle.logf_first_decrease(c[0], c[1]), chunks))
    assert found == [0xFFFFFFFF] * 8, [hex(f) for f in found]


def test_reciprocal_of_f16_divisors_is_correctly_rounded(oracle):
    """The cluster-key kernel divides by the dominant normal component, an f16 value, as v_rcp_f32 (within 1 ulp)
    plus one FMA Newton step. For every finite positive f16 divisor and every reciprocal within 1 ulp, the step
    lands on the correctly rounded quotient; without the step (planted) the neighbours do not. Exact rational
    arithmetic agrees on a sample of divisors."""
    from fractions import Fraction
    assert oracle.rcp_f16_newton_mismatches() == 0
    assert oracle.rcp_f16_newton_mismatches(planted=1) > 0
    for h in (1, 0x3C01, 0x3555, 0x5A5A, 0x7BFF):
        a = float(np.array(h, np.uint16).view(np.float16))
        r = np.float32(1.0) / np.float32(a)
        assert abs(Fraction(float(r)) - 1 / Fraction(a)) <= abs(Fraction(float(np.nextafter(r, np.float32(0)))) - 1 / Fraction(a))


def test_third_index_comparisons_match_formula_on_every_float(oracle):
    """The normal code's cell index f2u_sat(floor((u + 1) / 2 * 3)) as three threshold comparisons
    (vren_third_index), on every float bit pattern below 1.5 and every NaN; 8 threads. A threshold moved by one
    float must be caught."""
    from concurrent.futures import ThreadPoolExecutor
    assert oracle.third_index_mismatches(0xBEAAAA00, 0xBEAAAB00, planted=1) > 0
    chunks = [(i << 29, (i + 1) << 29) for i in range(8)]  # every 32-bit pattern
    with ThreadPoolExecutor(8) as ex:
        bad = list(ex.map(lambda c: oracle.third_index_mismatches(c[0], c[1]), chunks))
    assert bad == [0] * 8, bad


def test_vmath_expf_and_powi(oracle):
    rng = np.random.default_rng(2)
    worst = 0.0
    for x in rng.uniform(-30, 30, 4000).astype(np.float32):
        worst = max(worst, _ulp_err(oracle.expf(x), math.exp(float(x))))
    assert worst <= 2.0, worst
    # a^k by square-and-multiply: exact for powers of two, within k ulp-ish otherwise
    assert oracle.powi(2.0, 10) == 1024.0
    assert oracle.powi(1.5, 0) == 1.0
    a = np.float32(1.0 + 2.0 / 34.0)
    for k in (1, 7, 33, 121, 400):
        rel = abs(oracle.powi(a, k) - float(a) ** k) / float(a) ** k
        assert rel < 1e-6 * max(1, math.log2(k)), (k, rel)


def _tree_closed_form(x, op, ident):
    """tree[i] = op over the aligned block of 2^ctz(i+1) elements ending at i (input padded with ident)."""
    n = x.size
    p = 1 << max(0, (n - 1).bit_length())
    xp = np.full(p, ident, dtype=x.dtype)
    xp[:n] = x
    out = np.empty(p, dtype=x.dtype)
    for i in range(p):
        k = ((i + 1) & -(i + 1)).bit_length() - 1
        blk = xp[i + 1 - (1 << k): i + 1]
        out[i] = op(blk)
    return out


@pytest.mark.parametrize("n", [1, 2, 10, 1000, 10000])
def test_reduce_tree_u32_closed_form(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    for op, fn, ident in [(0, lambda b: np.uint32(np.sum(b, dtype=np.uint64) & 0xFFFFFFFF), 0),
                          (1, np.min, 0xFFFFFFFF), (2, np.max, 0)]:
        got = oracle.reduce_tree_u32(x, op)
        want = _tree_closed_form(x, fn, ident) if n <= 1000 else None
        if want is not None:
            assert np.array_equal(got, want), op
        assert got[-1] == fn(x), op
        assert got[-1] == oracle.reduce_total_u32(x, op)


def test_reduce_tree_u32_all_ones_reference_case(oracle):
    # reduce.type_uint: N = 10000 ones, padded to 16384 (vren_test/.../reduce.cpp:256-282)
    x = np.ones(10000, np.uint32)
    t = oracle.reduce_tree_u32(x, 0)
    assert t.size == 16384 and t[-1] == 10000
    i = np.arange(16384)
    block = (i + 1) & -(i + 1)
    lo = i + 1 - block
    want = np.clip(i + 1, 0, 10000) - np.clip(lo, 0, 10000)
    assert np.array_equal(t, want.astype(np.uint32))


def test_reduce_tree_multi_block(oracle):
    rng = np.random.default_rng(5)
    x = rng.integers(0, 1000, 3 * 100, dtype=np.uint32)
    t = oracle.reduce_tree_u32(x, 0, blocks=3)
    add = lambda b: np.uint32(np.sum(b, dtype=np.uint64) & 0xFFFFFFFF)
    for b in range(3):
        assert np.array_equal(t[b * 128:(b + 1) * 128], _tree_closed_form(x[b * 100:(b + 1) * 100], add, 0))


def test_reduce_tree_vec4_integer_valued(oracle):
    # integer-valued floats make the sum order-independent (as in reduce.type_vec4, reduce.cpp:174-196)
    rng = np.random.default_rng(7)
    x = rng.integers(0, 100, (1000, 4)).astype(np.float32)
    for op, fn in [(0, np.sum), (1, np.min), (2, np.max)]:
        t = oracle.reduce_tree_vec4(x, op)
        assert t.shape == (1024, 4)
        assert np.array_equal(t[-1], fn(x, axis=0).astype(np.float32)), op
    # identities of the padding (VRen.cmake:70-75)
    one = oracle.reduce_tree_vec4(np.array([[1, 2, 3, 4]], np.float32), 1)
    assert np.array_equal(one[0], [1, 2, 3, 4])
    t = oracle.reduce_tree_vec4(np.zeros((5, 4), np.float32), 1)
    assert t[6, 0] == np.float32(1e35) and t[5, 0] == 0.0 and t[7, 0] == 0.0
    t = oracle.reduce_tree_vec4(np.zeros((5, 4), np.float32), 2)
    assert t[6, 0] == np.float32(-1e35)


def test_exclusive_scan_vs_numpy(oracle):
    rng = np.random.default_rng(3)
    for n in (1, 7, 4096, 100003):
        x = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        want = np.concatenate([[0], np.cumsum(x.astype(np.uint64))[:-1]]).astype(np.uint64) & 0xFFFFFFFF
        assert np.array_equal(oracle.exclusive_scan_u32(x), want.astype(np.uint32))
    ones = np.ones(1 << 12, np.uint32)
    assert np.array_equal(oracle.exclusive_scan_u32(ones), np.arange(1 << 12, dtype=np.uint32))


def _numpy_upsweep(x):
    """Up-sweep tree of x (a power-of-2 length), built by pairwise sums level by level in numpy."""
    t = x.astype(np.uint64).copy()
    n = t.size
    step = 1
    while step < n:
        t[2 * step - 1::2 * step] += t[step - 1::2 * step]
        step *= 2
    return (t & 0xFFFFFFFF).astype(np.uint32)


def _exclusive(x):
    c = np.concatenate([np.zeros(1, np.uint64), np.cumsum(x.astype(np.uint64))[:-1]])
    return (c & np.uint64(0xFFFFFFFF)).astype(np.uint32)


@pytest.mark.parametrize("length", [1, 2, 8, 16, 512, 1024, 2048, 1 << 14])
@pytest.mark.parametrize("blocks", [1, 3])
def test_blelloch_downsweep_of_upsweep_trees(oracle, length, blocks):
    """downsweep(clear_last = true) of an up-sweep tree is the exclusive scan of the leaves (the use in
    radix_sort.cpp:230-289 and blelloch_scan.cpp:141-166); with clear_last = false the root's value becomes the
    prefix of the whole block, except in blocks shorter than 1024, where the reference's 1024-wide LDS pass
    overwrites the root with its zero padding's prefix."""
    rng = np.random.default_rng(length + blocks)
    x = rng.integers(0, 2 ** 32, (blocks, length), dtype=np.uint64).astype(np.uint32)
    trees = np.concatenate([_numpy_upsweep(r) for r in x])
    want = np.concatenate([_exclusive(r) for r in x])
    assert np.array_equal(oracle.blelloch_downsweep_u32(trees, length, blocks, True), want)
    kept = oracle.blelloch_downsweep_u32(trees, length, blocks, False)
    for b in range(blocks):
        total = np.uint32(np.sum(x[b], dtype=np.uint64) & 0xFFFFFFFF)
        root = total if length >= 1024 else np.uint32(0)
        assert np.array_equal(kept[b * length:(b + 1) * length], want[b * length:(b + 1) * length] + root), b


@pytest.mark.parametrize("length", [1, 4, 32, 1024, 8192])
def test_blelloch_downsweep_closed_form_on_arbitrary_input(oracle, length):
    """On any input (not only an up-sweep tree) each output is P plus, for every set bit l of its index i, the input
    at the last element of i's left sibling at that level, ((i >> (l + 1)) << (l + 1)) + 2^l - 1 (no level above l
    writes that element); P = 0 when clear_last or length < 1024, else the last input element."""
    rng = np.random.default_rng(length)
    d = rng.integers(0, 2 ** 32, 2 * length, dtype=np.uint64).astype(np.uint32)
    for clear in (True, False):
        got = oracle.blelloch_downsweep_u32(d, length, 2, clear)
        for b in range(2):
            seg = d[b * length:(b + 1) * length].astype(np.uint64)
            P = 0 if (clear or length < 1024) else int(seg[-1])
            i = np.arange(length)
            want = np.full(length, P, np.uint64)
            lvl = 0
            while (1 << lvl) < length:
                sib = ((i >> (lvl + 1)) << (lvl + 1)) + (1 << lvl) - 1
                want += np.where((i >> lvl) & 1, seg[sib], 0)
                lvl += 1
            assert np.array_equal(got[b * length:(b + 1) * length], (want & 0xFFFFFFFF).astype(np.uint32)), (clear, b)


def test_stable_sort_vs_numpy(oracle):
    rng = np.random.default_rng(4)
    k = rng.integers(0, 64, 5000, dtype=np.uint32)  # many ties
    v = np.arange(5000, dtype=np.uint32)
    ks, vs = oracle.stable_sort_kv(k, v)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(ks, k[order]) and np.array_equal(vs, v[order])
    # bit range [4, 12)
    k2 = rng.integers(0, 2 ** 32, 3000, dtype=np.uint64).astype(np.uint32)
    ks, vs = oracle.stable_sort_kv(k2, np.arange(3000, dtype=np.uint32), 4, 12)
    order = np.argsort((k2 >> 4) & 0xFF, kind="stable")
    assert np.array_equal(vs, order.astype(np.uint32))


def test_bucket_sort_vs_numpy(oracle):
    rng = np.random.default_rng(6)
    pairs = np.stack([rng.integers(0, 2 ** 32, 7000, dtype=np.uint64).astype(np.uint32),
                      np.arange(7000, dtype=np.uint32)], axis=1)
    for nb, shift in [(65536, 0), (64, 24)]:
        out, offs = oracle.bucket_sort(pairs, nb, shift)
        b = (pairs[:, 0] >> shift) & (nb - 1)
        order = np.argsort(b, kind="stable")
        assert np.array_equal(out, pairs[order])
        cnt = np.bincount(b, minlength=nb)
        assert np.array_equal(offs, np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint32))


def golden(name):
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)) as f:
        return json.load(f)


def test_calc_bvh_known_answers(oracle):
    # build_bvh.utils (vren_test/vren_test/primitives/build_bvh.cpp:223-254), from tests/golden
    kat = golden("build_bvh_utils.json")
    for fn in ("padded_leaf_count", "buffer_length", "root_index", "level_count"):
        assert len(kat[fn]) >= 5
        for n, v in kat[fn]:
            assert oracle.calc(fn, n) == v, (fn, n)


def random_leaves(n, padded, length, rng):
    nodes = np.zeros(length, BVH_DTYPE)
    pos = rng.uniform(0, 100, (n, 3)).astype(np.float32)
    ext = rng.uniform(0, 100, (n, 3)).astype(np.float32)
    nodes["min"][:n] = pos
    nodes["max"][:n] = pos + ext
    nodes["next"][:n] = 0xFFFFFFFF
    nodes["min"][n:padded] = np.inf
    nodes["max"][n:padded] = -np.inf
    nodes["next"][n:padded] = 0xFFFFFFFE
    return nodes


@pytest.mark.parametrize("n", [1, 10, 100, 1000, 10000])
def test_build_bvh_levels_and_point_queries(oracle, n):
    rng = np.random.default_rng(n)
    padded = oracle.calc("padded_leaf_count", n)
    nodes = random_leaves(n, padded, oracle.calc("buffer_length", n), rng)
    built = oracle.build_bvh(nodes, padded)
    # every parent = min/max over its valid children; next = first child or invalid
    start, size = 0, padded
    while size > 1:
        ch = built[start:start + size].reshape(-1, 32)
        par = built[start + size:start + size + size // 32]
        valid = ch["next"] != 0xFFFFFFFE
        anyv = valid.any(axis=1)
        mn = np.where(valid[..., None], ch["min"], np.inf).min(axis=1)
        mx = np.where(valid[..., None], ch["max"], -np.inf).max(axis=1)
        assert np.array_equal(par["min"][anyv], mn[anyv]) and np.array_equal(par["max"][anyv], mx[anyv])
        assert np.all(par["next"][~anyv] == 0xFFFFFFFE)
        assert np.array_equal(par["next"][anyv], (start + 32 * np.arange(size // 32))[anyv].astype(np.uint32))
        start, size = start + size, size // 32
    # the reference's property test: BVH-accelerated point query == linear scan over all nodes
    root = oracle.calc("root_index", n)
    for _ in range(8):
        p = rng.uniform(0, 100, 3).astype(np.float32)
        inside = np.all((built["min"] <= p) & (p <= built["max"]), axis=1)
        want = sorted(np.nonzero(inside)[0].tolist())
        got = []
        stack = [root]
        while stack:
            i = stack.pop()
            nd = built[i]
            if not (np.all(nd["min"] <= p) and np.all(p <= nd["max"])):
                continue
            got.append(i)
            if nd["next"] not in (0xFFFFFFFF, 0xFFFFFFFE):
                stack.extend(range(int(nd["next"]), int(nd["next"]) + 32))
        assert sorted(got) == want


def test_bounce_point_lights_closed_forms(oracle):
    # bounce_point_lights.comp:33-73 pinned by the motion it describes, not by its own formula
    lo, hi = (-50.0, -50.0, -50.0), (50.0, 50.0, 50.0)
    # no wall within reach: one segment of length speed * dt = 2 along d, direction kept, w = 1 and 0
    p0 = np.array([[1, 2, 3, 7]], np.float32)
    d0 = np.array([[0.5, -0.25, 0.125, 9]], np.float32)
    p, d = oracle.bounce_point_lights(p0, d0, lo, hi, 4.0, 0.5)
    assert np.array_equal(p[0], [2.0, 1.5, 3.25, 1.0]) and np.array_equal(d[0], [0.5, -0.25, 0.125, 0.0])
    # a position outside the box is clamped onto it before moving (x to -50, y to 50)
    p, d = oracle.bounce_point_lights([[-100, 60, 5, 1]], [[0.5, -0.5, 0, 0]], lo, hi, 1.0, 1.0)
    assert np.array_equal(p[0, :3], [-49.5, 49.5, 5.0])
    # billiard in [0, 10]^3: the unfolded path folds back with period 20 per axis; each reflection stops 1e-5
    # short of its wall, hence the tolerance
    box = ((0.0, 0.0, 0.0), (10.0, 10.0, 10.0))
    p, d = oracle.bounce_point_lights([[5, 5, 5, 1]], [[1, 0, 0, 0]], *box, 4.0, 5.0)  # x: 5 + 20 -> 5, moving +
    assert abs(p[0, 0] - 5.0) < 1e-3 and p[0, 1] == 5.0 and p[0, 2] == 5.0 and d[0, 0] == 1.0

    def fold(u):
        m = u % 20.0
        return (m, 1.0) if m <= 10.0 else (20.0 - m, -1.0)

    for dx, dy, length in ((0.6, 0.8, 37.0), (-0.8, 0.6, 51.0), (0.28, -0.96, 23.5)):
        p, d = oracle.bounce_point_lights([[5, 5, 5, 1]], [[dx, dy, 0, 0]], *box, length, 1.0)
        (wx, sx), (wy, sy) = fold(5.0 + dx * length), fold(5.0 + dy * length)
        assert abs(p[0, 0] - wx) < 1e-3 and abs(p[0, 1] - wy) < 1e-3, (dx, dy, p[0])
        assert np.sign(d[0, 0]) == sx * np.sign(dx) and np.sign(d[0, 1]) == sy * np.sign(dy)
    # the shader's quirk: a light clamped onto a wall and moving outward has both wall times of that axis <= 0
    # (replaced by 1e35), so it is never reflected and leaves the box
    p, d = oracle.bounce_point_lights([[0, -60, 0, 1]], [[0, -1, 0, 0]], lo, hi, 2.0, 1.0)
    assert np.array_equal(p[0, :3], [0.0, -52.0, 0.0]) and d[0, 1] == -1.0
    # invariants on random lights that start inside: direction components keep their magnitudes, and over a
    # frame-sized travel (30 units) positions stay in the box. (Over hundreds of units a few lights do leave it in
    # the shader's arithmetic too: a segment stops 1e-5 short in time, which can round onto the wall at |p| ~ 50.)
    rng = np.random.default_rng(5)
    pos = np.concatenate([rng.uniform(-49, 49, (2000, 3)), np.ones((2000, 1))], 1).astype(np.float32)
    dirs = rng.normal(size=(2000, 3))
    dirs = np.concatenate([dirs / np.linalg.norm(dirs, axis=1, keepdims=True), np.zeros((2000, 1))], 1).astype(np.float32)
    p, d = oracle.bounce_point_lights(pos, dirs, lo, hi, 300.0, 1.0)
    assert np.array_equal(np.abs(d[:, :3]), np.abs(dirs[:, :3]))
    p, d = oracle.bounce_point_lights(pos, dirs, lo, hi, 30.0, 1.0)
    assert (p[:, :3] >= -50.001).all() and (p[:, :3] <= 50.001).all()
