"""Pins of the oracle's material resolve, depth pyramid, debug views and CPU-baseline helpers (CPU only). The
expected values are worked out here from the reference's shaders and the Vulkan sampling rules, with numpy, never by
calling the oracle function under test:

* material resolve (shade.comp:89-96): hand-made textures with known texels, NEAREST and LINEAR, REPEAT,
  MIRRORED_REPEAT and CLAMP_TO_EDGE at coordinates whose texels and weights are worked out by hand;
* depth pyramid (depth_buffer_reduce.comp): a numpy 2x2 max over the in-bounds texels of each level;
* show_clusters (show_clusters.comp:60-121) and visualize_bvh (show_bvh.comp) in closed form;
* the multi-threaded CPU baselines give the serial results.
"""
import numpy as np
import pytest

import oracle_lib as ol
from vren_amd._lib import Material

F32 = np.float32


@pytest.fixture(scope="module")
def co(oracle):
    return ol.ClusteredOracle(oracle)


def _materials(*rows):
    arr = (Material * len(rows))()
    for i, (bt, mt, mf, rf, bcf) in enumerate(rows):
        arr[i].base_color_texture_idx, arr[i].metallic_roughness_texture_idx = bt, mt
        arr[i].metallic_factor, arr[i].roughness_factor = mf, rf
        for k in range(4):
            arr[i].base_color_factor[k] = bcf[k]
    return arr


# a 4x2 RGBA8 texture whose red channel is the texel's index 10 * (y * 4 + x) + 5 and green 200 - red
TEX = np.zeros((2, 4, 4), np.uint8)
for _y in range(2):
    for _x in range(4):
        TEX[_y, _x] = [10 * (_y * 4 + _x) + 5, 200 - (10 * (_y * 4 + _x) + 5), 7 * _x + 3, 255]


def _red(x, y):
    return (10 * (y * 4 + x) + 5) / 255.0


NEAREST_CASES = [  # (s, t, address mode) -> texel (x, y): u = s * 4, v = t * 2, i = floor(u)
    (0.0, 0.0, 0, (0, 0)),
    (0.3, 0.7, 0, (1, 1)),        # u = 1.2, v = 1.4
    (0.999, 0.49, 0, (3, 0)),
    (-0.1, 0.0, 0, (3, 0)),       # REPEAT: floor(-0.4) = -1 -> 3
    (1.3, -0.6, 0, (1, 0)),       # u = 5.2 -> 5 mod 4 = 1; v = -1.2 -> -2 mod 2 = 0
    (1.3, 0.0, 1, (2, 0)),        # MIRRORED: i = 5: t = 5 mod 8 = 5, 5 - 4 = 1 -> 3 - 1 = 2
    (-0.1, 0.0, 1, (0, 0)),       # MIRRORED: i = -1: t = 7, 7 - 4 = 3 -> 3 - 3 = 0
    (1.05, 1.25, 1, (3, 1)),      # u = 4.2 -> i 4: t 4, 0 -> 3; v = 2.5 -> j 2: t 2, 0 -> 1
    (5.0, -3.0, 2, (3, 0)),       # CLAMP
    (-2.0, 0.9, 2, (0, 1)),
]


@pytest.mark.parametrize("s,t,mode,texel", NEAREST_CASES)
def test_material_nearest_texel_selection(co, s, t, mode, texel):
    mats = _materials((0, 0, 1.0, 1.0, (1.0, 1.0, 1.0, 1.0)))
    ar, me = co.resolve_materials(np.array([[s, t]], F32), np.array([0], np.uint16), mats, [(TEX, 0, mode, mode)])
    x, y = texel
    want = TEX[y, x, :3].astype(np.float64) / 255.0
    assert np.array_equal(ar[0, :3], want), (ar[0], want)
    assert ar[0, 3] == TEX[y, x, 1] / 255.0 and me[0] == TEX[y, x, 2] / 255.0


def test_material_linear_weights(co):
    """LINEAR: i0 = floor(u - 0.5), alpha = frac(u - 0.5). s = 0.375 -> u - 0.5 = 1.0: texel 1 exactly; s = 0.5 ->
    1.5: halfway between texels 1 and 2; t = 0.5 -> v - 0.5 = 0.5: halfway between rows 0 and 1; s = 0.0625 ->
    -0.25: i0 = -1 (REPEAT: 3), alpha = 0.75 towards texel 0."""
    mats = _materials((0, 0, 1.0, 1.0, (1.0, 1.0, 1.0, 1.0)))
    tex = [(TEX, 1, 0, 0)]
    tc = np.array([[0.375, 0.25], [0.5, 0.25], [0.375, 0.5], [0.0625, 0.25]], F32)
    ar, _ = co.resolve_materials(tc, np.zeros(4, np.uint16), mats, tex)
    assert ar[0, 0] == _red(1, 0)
    assert abs(ar[1, 0] - 0.5 * (_red(1, 0) + _red(2, 0))) < 1e-15
    assert abs(ar[2, 0] - 0.5 * (_red(1, 0) + _red(1, 1))) < 1e-15
    assert abs(ar[3, 0] - (0.25 * _red(3, 0) + 0.75 * _red(0, 0))) < 1e-15


def test_material_factors_indices_and_defaults(co):
    """albedo = base texel rgb * base_color_factor rgb, roughness = mr texel g * roughness_factor, metallic =
    mr texel b * metallic_factor (shade.comp:94-96); out-of-range material or texture indices read zeros."""
    white = np.full((1, 1, 4), 255, np.uint8)
    mr = np.array([[[0, 128, 64, 255]]], np.uint8)
    mats = _materials((0, 1, 0.5, 0.25, (0.2, 0.4, 0.8, 1.0)),   # base white, mr texture 1
                      (5, 1, 1.0, 1.0, (1.0, 1.0, 1.0, 1.0)),    # base texture out of range
                      (0, 9, 1.0, 1.0, (1.0, 1.0, 1.0, 1.0)))    # mr texture out of range
    tex = [(white, 0, 0, 0), (mr, 0, 0, 0)]
    tc = np.zeros((4, 2), F32)
    ar, me = co.resolve_materials(tc, np.array([0, 1, 2, 3], np.uint16), mats, tex)
    assert np.allclose(ar[0], [float(F32(0.2)), float(F32(0.4)), float(F32(0.8)), 128 / 255 * 0.25], rtol=1e-15)
    assert me[0] == 64 / 255 * 0.5
    assert np.all(ar[1, :3] == 0) and ar[1, 3] == 128 / 255 and me[1] == 64 / 255
    assert np.all(ar[2, :3] == 1.0) and ar[2, 3] == 0 and me[2] == 0
    assert np.all(ar[3] == 0) and me[3] == 0   # material index 3 past the table


def test_material_shade_equals_resolved_shade(co):
    """The material-path FP64 shade equals the per-pixel-material FP64 shade given the resolved values (rounded to
    float32 for the latter's inputs: agreement to ~1e-7)."""
    W, H, L = 96, 64, 200
    cam = ol.camera(W, H)
    cf = ol.cfg()
    pos, lig = co.synth_lights(71, L)
    depth, normals = co.synth_frame(72, W, H, cam)
    vp, codes, bvh = co.light_bvh(pos, lig, cam, cf)
    keys, U, refs = co.unique_clusters(W, H, cam, cf, depth, normals)
    counts, offsets, idx, tot = co.assign(W, H, cam, cf, keys, U, bvh, codes, vp)
    rng = np.random.default_rng(5)
    tex = [(rng.integers(0, 256, (16, 8, 4), dtype=np.uint8), 1, 0, 1),
           (rng.integers(0, 256, (5, 7, 4), dtype=np.uint8), 0, 2, 0)]
    mats = _materials((0, 1, 0.9, 0.8, (1.0, 0.5, 0.7, 1.0)), (1, 0, 0.1, 1.0, (0.3, 0.9, 0.2, 1.0)))
    tc = rng.uniform(-1.5, 2.5, (W * H, 2)).astype(F32)
    mi = rng.integers(0, 3, W * H).astype(np.uint16)   # 2 = past the table
    out, mask = co.shade_materials(W, H, cam, cf, depth, normals, tc, mi, mats, tex, refs, U, counts, offsets, idx,
                                   pos, lig)
    ar, me = co.resolve_materials(tc, mi, mats, tex)
    ref, mask2 = co.shade(W, H, cam, cf, depth, normals, ar.astype(F32), me.astype(F32), refs, U, counts, offsets,
                          idx, pos, lig)
    assert np.array_equal(mask, mask2) and mask.sum() > 1000
    # a zero roughness (the zero material of an index past the table) with N.V = 0 is 0/0 in Smith G, a NaN in
    # both (pbr.glsl:8-13)
    a, b = out[mask, :3], ref[mask, :3]
    assert np.array_equal(np.isnan(a), np.isnan(b))
    fin = ~np.isnan(b)
    rel = np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1e-3)
    assert fin.sum() > 1000 and rel.max() < 2e-6, rel.max()


# ---- depth pyramid (depth_buffer_pyramid.cpp:18, depth_buffer_reduce.comp) ----------------------------------------

def _numpy_pyramid(d, W, H):
    levels = [d.reshape(H, W)]
    n = int(np.floor(np.log2(max(W, H)))) + 1
    for l in range(1, n):
        w, h = max(W >> l, 1), max(H >> l, 1)
        p = levels[-1]
        m = np.zeros((h, w), np.float32)
        for i in (0, 1):
            for j in (0, 1):
                src = np.zeros((h, w), np.float32)
                ys, xs = 2 * np.arange(h)[:, None] + j, 2 * np.arange(w)[None, :] + i
                ok = (ys < p.shape[0]) & (xs < p.shape[1])
                src[ok] = p[np.minimum(ys, p.shape[0] - 1), np.minimum(xs, p.shape[1] - 1)][ok]
                v = np.where(ok & ~np.isnan(src), src, -np.inf)
                m = np.where(v > m, v, m)   # GLSL max(m, v) = m < v ? v : m; NaN never wins
        levels.append(m.astype(np.float32))
    return np.concatenate([x.reshape(-1) for x in levels]), n


@pytest.mark.parametrize("W,H", [(1, 1), (2, 1), (5, 100), (64, 64), (333, 200), (1920, 1080), (4096, 3)])
def test_depth_pyramid_vs_numpy(co, W, H):
    rng = np.random.default_rng(W * 7 + H)
    d = rng.uniform(0.0, 1.0, W * H).astype(F32)
    d[rng.integers(0, W * H, max(1, W * H // 50))] = np.nan
    d[rng.integers(0, W * H, max(1, W * H // 50))] = -0.5
    got, n = co.depth_pyramid(d, W, H)
    want, nw = _numpy_pyramid(d, W, H)
    assert n == nw
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_depth_pyramid_level_count_known_answers(oracle):
    # floor(log2(max(W, H))) + 1 (depth_buffer_pyramid.cpp:18)
    for (W, H, n) in [(1, 1, 1), (2, 1, 2), (3, 3, 2), (4, 1, 3), (1920, 1080, 11), (2048, 1, 12), (65535, 2, 16)]:
        assert oracle.lib.o_depth_pyramid_level_count(W, H) == n


# ---- show_clusters (show_clusters.comp:60-121) and visualize_bvh (show_bvh.comp) ---------------------------------

def _triple32(x):
    """Wellons' triple32 hash as written in show_clusters.comp:48-58, in Python integers."""
    m = 0xFFFFFFFF
    x ^= x >> 17
    x = (x * 0xED5AD4BB) & m
    x ^= x >> 11
    x = (x * 0xAC4C1B51) & m
    x ^= x >> 15
    x = (x * 0x31848BAB) & m
    x ^= x >> 14
    return x


def test_triple32_values(oracle):
    for x in [0, 1, 2, 0xFFFFFFFF, 123456789, 0x80000000]:
        assert oracle.lib.o_triple32(x) == _triple32(x)
    assert _triple32(0) == 0


def test_show_clusters_modes(co):
    W, H = 70, 40   # 3 x 2 tiles
    rng = np.random.default_rng(9)
    U = 50
    i, j = rng.integers(0, 3, U), rng.integers(0, 2, U)
    k, n = rng.integers(0, 1024, U), rng.integers(0, 64, U)
    keys = (i | (j << 8) | (k << 16) | (n << 26)).astype(np.uint32)
    counts = rng.integers(0, 100, U).astype(np.uint32)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32)
    idx = rng.integers(0, 4096, int(counts.sum())).astype(np.uint32)
    refs = rng.integers(0, U + 3, W * H).astype(np.uint32)   # a few references past the table
    h = lambda v: np.float32(np.float32(v) / np.float32(4294967296.0))
    for mode in range(7):
        got = co.show_clusters(W, H, mode, refs, keys, counts, offsets, idx)
        assert np.all(got[:, 3] == 1.0)
        for p in range(0, W * H, 7):
            c = refs[p]
            if c >= U or mode == 6:
                assert got[p, :3].tolist() == [1.0, 0.0, 0.0]
                continue
            kk = int(keys[c])
            if mode == 0:
                want = [h(_triple32(kk))] * 3
            elif mode == 1:
                want = [F32(i[c]) / F32(3), F32(j[c]) / F32(2), F32(k[c]) / F32(1024)]
            elif mode == 2:
                want = [F32(n[c] // 9) / F32(6), F32((n[c] // 3) % 3) / F32(3), F32(n[c] % 3) / F32(3)]
            elif mode == 3:
                want = [min(F32(counts[c]) / F32(64), F32(1))] * 3
            elif mode == 4:
                want = [h(_triple32(int(offsets[c])))] * 3
            else:
                x = 0
                for q in range(int(counts[c])):
                    x ^= int(idx[offsets[c] + q])
                want = [h(_triple32(x))] * 3
            assert np.array_equal(got[p, :3], np.array(want, F32)), (mode, p)


def test_visualize_bvh_edges(co):
    """level_count 1: 32 leaves then the root. Every valid node's 24 vertices are the 12 edges of its AABB (each
    edge joins two corners that differ in one coordinate, every corner is on 3 edges), leaves coloured 0x00ff00,
    the root 0x0000ff; invalid nodes are 24 zero vertices of colour 0."""
    rng = np.random.default_rng(3)
    bvh = np.zeros(33, ol.BVH_DTYPE)
    bvh["min"] = rng.uniform(-5, 0, (33, 3)).astype(F32)
    bvh["max"] = rng.uniform(0, 5, (33, 3)).astype(F32)
    bvh["next"] = 0xFFFFFFFF
    bvh["next"][[3, 17]] = 0xFFFFFFFE
    bvh["next"][32] = 0
    v = co.visualize_bvh(bvh, 1)
    for node in range(33):
        vv = v[24 * node:24 * node + 24]
        if node in (3, 17):
            assert np.all(vv == 0)
            continue
        assert np.all(vv[:, 3] == (0x0000FF if node == 32 else 0x00FF00))
        xyz = vv[:, :3].view(np.float32)
        mn, mx = bvh["min"][node], bvh["max"][node]
        corners = {}
        for e in range(12):
            a, b = xyz[2 * e], xyz[2 * e + 1]
            for pnt in (a, b):
                sel = tuple(int(pnt[c] == mx[c]) for c in range(3))
                assert np.all((pnt == mn) | (pnt == mx))
                corners[sel] = corners.get(sel, 0) + 1
            assert int(np.sum(a != b)) == 1
        assert len(corners) == 8 and all(c == 3 for c in corners.values())


# ---- CPU baseline helpers ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("threads", [1, 3, 8])
def test_cpu_baseline_helpers_give_serial_results(oracle, threads):
    keys = oracle.pcg32(42, 54, 100003) & np.uint32(0xFFF)   # many ties: stability matters
    vals = np.arange(keys.size, dtype=np.uint32)
    _, sk, sv = oracle.cpu_stable_sort_pairs(keys, vals, threads)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(sk, keys[order]) and np.array_equal(sv, vals[order])
    x = oracle.pcg32(43, 54, 100003)
    _, sc = oracle.cpu_exclusive_scan(x, threads)
    assert np.array_equal(sc, oracle.exclusive_scan_u32(x))
    _, r = oracle.cpu_reduce_add(x, threads)
    assert r == int(np.sum(x, dtype=np.uint64) & 0xFFFFFFFF)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("size", [1, 7, 16, 200])
def test_texel_wrap_matches_integer_definition(co, mode, size):
    """Texture whose texel x holds R = x: over random coordinates up to |u| ~ 2^22, NEAREST selects
    wrap(floor(fl32(s * size))) exactly, with the wrap computed here on Python integers (REPEAT: mod size;
    MIRRORED_REPEAT: t = i mod 2 size, t < size ? t : 2 size - 1 - t; CLAMP_TO_EDGE: clamp)."""
    tex = np.zeros((1, size, 4), np.uint8)
    tex[0, :, 0] = np.arange(size) % 256
    tex[0, :, 1] = np.arange(size) // 256
    rng = np.random.default_rng(size * 3 + mode)
    s = np.concatenate([rng.uniform(-3, 3, 4000), rng.uniform(-2e4, 2e4, 4000) / size,
                        (np.arange(-50, 50) / size).astype(np.float64)]).astype(F32)
    mats = _materials((0, 0, 1.0, 1.0, (1.0, 1.0, 1.0, 1.0)))
    tc = np.stack([s, np.zeros_like(s)], axis=1)
    ar, _ = co.resolve_materials(tc, np.zeros(s.size, np.uint16), mats, [(tex, 0, mode, mode)])
    got = np.rint(ar[:, 0] * 255).astype(np.int64) + 256 * np.rint(ar[:, 3] * 255).astype(np.int64)
    for k in range(s.size):
        i = int(np.floor(np.float64(np.float32(s[k]) * np.float32(size))))
        if mode == 0:
            want = i % size
        elif mode == 1:
            t = i % (2 * size)
            want = t if t < size else 2 * size - 1 - t
        else:
            want = min(max(i, 0), size - 1)
        assert got[k] == want, (s[k], got[k], want)
