"""Pins of the oracle's clustered-shading restatement (CPU only). Each check is independent of the oracle's code:
known answers worked out by hand from the reference's shaders, brute force, closed forms of special cases, and
FP64 central differences for the adjoint. A plausible mistake in the oracle (a swapped Morton axis, a wrong log
base in the depth slice, a dropped BRDF term, a wrong sign in H, a missed light in the traversal) fails one of them.
"""
import math

import numpy as np
import pytest

import oracle_lib as ol


@pytest.fixture(scope="module")
def co(oracle):
    return ol.ClusteredOracle(oracle)


def _half(x):
    return np.array([x], np.float16).view(np.uint16)[0]


# ---- construct_point_light_bvh: Morton codes (discretize_point_light_positions.comp:26-43) -----------------------

def test_morton_known_answers(co):
    cam = ol.camera(64, 64)   # view = world - (0, 0, -50): yaw = pitch = 0
    view_pts = np.array([
        [0.0, 0.0, 0.0],      # AABB min corner -> q (0,0,0) -> 0
        [32.0, 32.0, 32.0],   # AABB max corner -> q (32,32,32): only the low 5 bits are interleaved -> 0
        [1.5, 0.5, 0.5],      # q (1,0,0) -> x bit 0 -> 1
        [0.5, 1.5, 0.5],      # q (0,1,0) -> y bit 0 -> 2
        [0.5, 0.5, 1.5],      # q (0,0,1) -> z bit 0 -> 4
        [3.5, 5.5, 6.5],      # q (3,5,6): triples (z y x) from bit 0: 011, 101, 110 -> 3 + 5*8 + 6*64 = 427
        [31.5, 31.5, 31.5],   # q (31,31,31) -> 2^15 - 1
        [16.0, 0.0, 0.0],     # q (16,0,0) -> x bit 4 at bit 12 -> 4096
    ], np.float32)
    want_code = [0, 0, 1, 2, 4, 427, 32767, 4096]
    L = len(view_pts)
    pos = np.zeros((L, 4), np.float32)
    pos[:, :3] = view_pts + np.array([0.0, 0.0, -50.0], np.float32)
    lig = np.zeros((L, 4), np.float32)
    lig[:, :3] = 1.0
    lig[:, 3] = np.arange(1, L + 1, dtype=np.float32) * 0.25
    vp, codes, bvh = co.light_bvh(pos, lig, cam, ol.cfg())
    assert np.array_equal(vp[:, :3], view_pts) and np.all(vp[:, 3] == 1.0)
    # sorted by code, ties by light index (stable)
    order = sorted(range(L), key=lambda i: (want_code[i], i))
    assert codes[:, 1].tolist() == order
    assert codes[:, 0].tolist() == [want_code[i] for i in order]
    # leaves: view position -+ radius, in sorted order; the rest of the padded leaves are invalid
    for s, i in enumerate(order):
        assert np.array_equal(bvh["min"][s], view_pts[i] - lig[i, 3])
        assert np.array_equal(bvh["max"][s], view_pts[i] + lig[i, 3])
        assert bvh["next"][s] == 0xFFFFFFFF
    assert np.all(bvh["next"][L:32] == 0xFFFFFFFE)


def test_morton_30_bit_given_aabb(co):
    cam = ol.camera(64, 64)
    # given AABB [0, 1024)^3 in view space, 10 bits per axis: q = floor(v)
    rng = np.random.default_rng(5)
    q = rng.integers(0, 1024, size=(200, 3))
    view_pts = q.astype(np.float32) + 0.5
    pos = np.zeros((200, 4), np.float32)
    pos[:, :3] = view_pts + np.array([0.0, 0.0, -50.0], np.float32)
    lig = np.ones((200, 4), np.float32)
    cf = ol.cfg(morton_bits=10, morton_aabb_mode=1, aabb=((0.0, 0.0, 0.0), (1024.0, 1024.0, 1024.0)))
    vp, codes, bvh = co.light_bvh(pos, lig, cam, cf)

    def spread(v):  # bit b of v -> bit 3b, by repeated shifting (not the oracle's loop)
        out = 0
        for b in range(10):
            out |= ((int(v) >> b) & 1) << (3 * b)
        return out
    want = np.array([spread(a) | (spread(b) << 1) | (spread(c) << 2) for a, b, c in q], np.uint64)
    order = np.lexsort((np.arange(200), want))
    assert np.array_equal(codes[:, 1], order.astype(np.uint32))
    assert np.array_equal(codes[:, 0].astype(np.uint64), want[order])


# ---- find_unique_cluster_list: keys (find_unique_clusters.comp:46-121, clustered_shading.glsl:7-43) ---------------

NORMAL_CODES = {  # cube face (n < 0: 0..2, n > 0: 3..5 by axis), centre cell (1, 1) of the 3x3 face grid
    (-1, 0, 0): 0 * 9 + 4, (0, -1, 0): 1 * 9 + 4, (0, 0, -1): 2 * 9 + 4,
    (1, 0, 0): 3 * 9 + 4, (0, 1, 0): 4 * 9 + 4, (0, 0, 1): 5 * 9 + 4,
}


def test_cluster_key_fields(co):
    W, H = 200, 120
    cam = ol.camera(W, H)
    depth, normals = co.synth_frame(4321, W, H, cam)
    raw = co.tile_keys(W, H, cam, ol.cfg(), depth, normals)
    Sx, Sy = (W + 31) // 32, (H + 31) // 32
    view, proj = co.camera_matrices(cam)
    p10, p14 = float(proj[10]), float(proj[14])
    log_a = math.log(1.0 + 2.0 * math.tan(float(np.float32(cam.fov_y)) / 2.0) / Sy)
    checked_slices = 0
    seen_codes = set()
    for t in range(Sx * Sy):
        i, j = t % Sx, t // Sx
        for l in range(1024):
            gx, gy = i * 32 + (l & 31), j * 32 + (l >> 5)
            p = (gy % H) * W + (gx % W)
            k = int(raw[t, l])
            assert (k & 0xFF) == i and ((k >> 8) & 0xFF) == j
            n = tuple(int(v) for v in normals[p, :3].view(np.float16).astype(np.float32))
            code = k >> 26
            if n == (0, 0, 0):
                assert code == 63          # no surface: all-ones code
            else:
                assert code == NORMAL_CODES[n], (n, code)
                seen_codes.add(code)
            # log depth slice: floor(log(z / near) / log(1 + 2 tan(fov/2) / Sy)), z from the D32 value in FP64
            z = p14 / (float(depth[p]) - p10)
            s = math.log(z / cam.near_plane) / log_a
            if abs(s - round(s)) > 1e-3:
                assert ((k >> 16) & 0x3FF) == int(math.floor(s)), (p, z, s, (k >> 16) & 0x3FF)
                checked_slices += 1
    assert checked_slices > 0.99 * Sx * Sy * 1024
    assert len(seen_codes) >= 3


def test_cluster_key_sliced_mode(co):
    W, H = 96, 64
    cam = ol.camera(W, H)
    depth, normals = co.synth_frame(99, W, H, cam)
    cf = ol.cfg(key_mode=1, depth_slices=32)
    raw = co.tile_keys(W, H, cam, cf, depth, normals)
    view, proj = co.camera_matrices(cam)
    p10, p14 = float(proj[10]), float(proj[14])
    faces = {(-1, 0, 0): 0, (0, -1, 0): 1, (0, 0, -1): 2, (1, 0, 0): 3, (0, 1, 0): 4, (0, 0, 1): 5, (0, 0, 0): 7}
    for t in range(raw.shape[0]):
        for l in range(0, 1024, 7):
            gx, gy = (t % 3) * 32 + (l & 31), (t // 3) * 32 + (l >> 5)
            p = (gy % H) * W + (gx % W)
            k = int(raw[t, l])
            n = tuple(int(v) for v in normals[p, :3].view(np.float16).astype(np.float32))
            assert (k >> 24) == faces[n]
            z = p14 / (float(depth[p]) - p10)
            s = math.log(z / cam.near_plane) / math.log(cam.far_plane / cam.near_plane) * 32
            if abs(s - round(s)) > 1e-3:
                assert ((k >> 16) & 0xFF) == min(int(math.floor(s)), 31)


def test_unique_clusters_sorted_unique_per_tile(co):
    W, H = 200, 120   # partial tiles on both axes (the reference's REPEAT wrap fills them)
    cam = ol.camera(W, H)
    depth, normals = co.synth_frame(4321, W, H, cam)
    cf = ol.cfg()
    raw = co.tile_keys(W, H, cam, cf, depth, normals)
    keys, U, refs = co.unique_clusters(W, H, cam, cf, depth, normals)
    want = np.concatenate([np.unique(raw[t]) for t in range(raw.shape[0])])
    assert U == want.size and np.array_equal(keys, want)
    Sx = (W + 31) // 32
    gy, gx = np.divmod(np.arange(W * H), W)
    t = (gy // 32) * Sx + gx // 32
    l = (gy % 32) * 32 + (gx % 32)
    assert np.array_equal(keys[refs], raw[t, l])   # keys[ref[p]] == key(p)


# ---- assign_lights (assign_lights.comp:77-241): BVH traversal vs brute force --------------------------------------

@pytest.mark.parametrize("aabb_mode", [0, 1])
def test_assign_lights_vs_brute_force(co, aabb_mode):
    W, H, L = 160, 96, 700
    cam = ol.camera(W, H)
    cf = ol.cfg(cluster_aabb_mode=aabb_mode)
    pos, lig = co.synth_lights(777, L)
    depth, normals = co.synth_frame(778, W, H, cam)
    vp, codes, bvh = co.light_bvh(pos, lig, cam, cf)
    keys, U, refs = co.unique_clusters(W, H, cam, cf, depth, normals)
    counts, offsets, idx, tot = co.assign(W, H, cam, cf, keys, U, bvh, codes, vp)
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32))
    # every leaf whose ancestors all pass the node test and which passes the leaf test (no traversal order)
    bc, bi = co.assign_brute(W, H, cam, cf, keys, bvh, codes, vp, lig, mode=0)
    assert np.array_equal(counts, bc) and np.array_equal(idx, bi)
    if aabb_mode == 1:
        # with a well-formed cluster AABB every passing leaf also passes its ancestors: plain brute force over all
        # lights gives the same lists
        pc, pi = co.assign_brute(W, H, cam, cf, keys, bvh, codes, vp, lig, mode=1)
        assert np.array_equal(counts, pc) and np.array_equal(idx, pi)
    assert tot > U  # non-trivial lists
    for c in range(U):  # ascending light index within each list
        seg = idx[offsets[c]:offsets[c] + counts[c]]
        assert np.all(seg[1:] > seg[:-1])


# ---- shade (shade.comp:69-141, pbr.glsl:6-112): closed forms on a one-pixel, on-axis scene -------------------------

F32 = lambda x: float(np.float32(x))  # the shaders' constants are GLSL floats
PI_REF = F32(3.14)  # common.glsl:7


def _shade_axis(co, z_view, light_offsets, colours, radii, albedo, rough, metal, atten):
    W = H = 2
    cam = ol.camera(W, H)
    _, proj = co.camera_matrices(cam)
    depth = np.ones(W * H, np.float32)
    depth[3] = np.float32(float(proj[10]) + float(proj[14]) / z_view)
    z = float(proj[14]) / (float(depth[3]) - float(proj[10]))   # the surface's view depth the shader sees
    normals = np.zeros((W * H, 4), np.uint16)
    normals[3, 2] = _half(-1.0)
    ar = np.zeros((W * H, 4), np.float32)
    ar[3, :3] = albedo
    ar[3, 3] = rough
    me = np.zeros(W * H, np.float32)
    me[3] = metal
    n = len(light_offsets)
    pos = np.zeros((n, 4), np.float32)
    lig = np.zeros((n, 4), np.float32)
    for i, d in enumerate(light_offsets):
        pos[i, 2] = np.float32(-50.0 + z - d)   # on the axis, d in front of the surface (world = view - 50)
        lig[i, :3] = colours[i]
        lig[i, 3] = radii[i]
    refs = np.zeros(W * H, np.uint32)
    counts = np.array([n], np.uint32)
    offsets = np.array([0], np.uint32)
    idx = np.arange(n, dtype=np.uint32)
    out, mask = co.shade(W, H, cam, ol.cfg(attenuation=atten), depth, normals, ar, me, refs, 1, counts, offsets, idx,
                         pos, lig)
    assert mask.tolist() == [False, False, False, True]
    dist = [abs((-50.0 + z) - float(pos[i, 2])) for i in range(n)]
    return out[3, :3], dist


def _closed_form_linear(albedo, rough, metal, colour, dist, radius, atten):
    """N = V = -L (light on the axis between camera and surface): H = N, all cosines 1, f5 = 0, G = 1."""
    a2 = rough * rough
    ndf = a2 / (PI_REF * a2 * a2)
    F = np.array([F32(0.04) * (1 - metal) + a * metal for a in albedo])
    kD = (1.0 - F) * (1.0 - metal)
    spec = ndf * 1.0 / (4.0 + F32(0.0001))
    x = dist / radius
    att = (max(0.0, 1.0 - x ** 4) ** 2) if atten == 1 else 1.0 - min(x, 1.0) ** 1.5
    return (kD * np.asarray(albedo) / PI_REF + F * spec) * np.asarray(colour) * att


def _gamma(lin):
    return np.power(lin, 1.0 / float(np.float32(2.2)))


@pytest.mark.parametrize("atten", [0, 1])
@pytest.mark.parametrize("metal", [0.0, 1.0])
def test_shade_single_light_closed_form(co, atten, metal):
    albedo = np.array([0.25, 0.5, 0.75], np.float32)
    rough = np.float32(0.6)
    colour = np.array([3.0, 2.0, 1.0], np.float32)
    got, dist = _shade_axis(co, 20.0, [2.5], [colour], [np.float32(6.0)], albedo, rough, metal, atten)
    lin = F32(0.03) * albedo.astype(np.float64) + _closed_form_linear(albedo.astype(np.float64), float(rough), metal,
                                                                 colour, dist[0], 6.0, atten)
    assert np.allclose(got, _gamma(lin), rtol=1e-12, atol=0)


def test_shade_light_beyond_radius_and_no_lights(co):
    albedo = np.array([0.2, 0.4, 0.6], np.float32)
    got, _ = _shade_axis(co, 30.0, [5.0], [np.ones(3, np.float32)], [np.float32(4.0)], albedo, np.float32(0.5), 0.0, 0)
    assert np.allclose(got, _gamma(F32(0.03) * albedo.astype(np.float64)), rtol=1e-12, atol=0)   # ambient only
    got0, _ = _shade_axis(co, 30.0, [], [], [], albedo, np.float32(0.5), 0.0, 0)
    assert np.array_equal(got, got0)


def test_shade_lights_add_in_linear_space(co):
    albedo = np.array([0.3, 0.3, 0.3], np.float32)
    args = dict(albedo=albedo, rough=np.float32(0.4), metal=0.0, atten=1)
    c1, c2 = np.array([1.0, 0.5, 0.2], np.float32), np.array([0.1, 0.9, 2.0], np.float32)
    both, _ = _shade_axis(co, 25.0, [1.0, 3.0], [c1, c2], [np.float32(5.0), np.float32(7.0)], **args)
    one, _ = _shade_axis(co, 25.0, [1.0], [c1], [np.float32(5.0)], **args)
    two, _ = _shade_axis(co, 25.0, [3.0], [c2], [np.float32(7.0)], **args)
    g = float(np.float32(2.2))
    amb = F32(0.03) * albedo.astype(np.float64)
    assert np.allclose(both ** g - amb, (one ** g - amb) + (two ** g - amb), rtol=1e-9, atol=1e-15)


# ---- shade adjoint: FP64 central differences of the oracle's loss --------------------------------------------------

def test_shade_backward_vs_central_differences(co):
    W, H, L = 64, 40, 48
    cam = ol.camera(W, H)
    cf_lists, cf = ol.cfg(), ol.cfg(attenuation=1)
    pos, lig = co.synth_lights(31, L)
    # pull the lights towards the visible scene so that most of them light something
    pos[:, 2] = pos[:, 2] * 0.3 - 20.0
    lig[:, 3] *= 3.0
    depth, normals = co.synth_frame(32, W, H, cam)
    ar, me = co.synth_materials(33, W * H)
    vp, codes, bvh = co.light_bvh(pos, lig, cam, cf_lists)
    keys, U, refs = co.unique_clusters(W, H, cam, cf_lists, depth, normals)
    counts, offsets, idx, tot = co.assign(W, H, cam, cf_lists, keys, U, bvh, codes, vp)
    assert tot > 0
    key_cap = max(U, 1)
    args = (W, H, cam, cf, depth, normals, ar, me, refs, key_cap, counts, offsets, idx)
    rng = np.random.default_rng(3)
    target = np.zeros((W * H, 4), np.float32)
    target[:, :3] = rng.uniform(0.0, 1.0, (W * H, 3)).astype(np.float32)
    pos_d, lig_d = pos.astype(np.float64), lig.astype(np.float64)
    out, mask = co.shade(W, H, cam, cf, depth, normals, ar, me, refs, key_cap, counts, offsets, idx, pos, lig)
    # dLoss/dout for the L2 loss on the FP64 forward
    gout = np.zeros((W * H, 4), np.float32)
    gout[mask, :3] = (2.0 * (out[mask, :3] - target[mask, :3])).astype(np.float32)
    gpos, glig, gar = co.shade_backward(*args, pos_d, lig_d, gout)
    pixels = np.nonzero(mask)[0].astype(np.uint32)
    h = 1e-6
    checked = 0
    for j in np.unique(idx)[:12]:
        for comp in range(7):
            def ev(delta):
                pd, ld = pos_d.copy(), lig_d.copy()
                if comp < 3:
                    pd[j, comp] += delta
                else:
                    ld[j, comp - 3] += delta
                return co.shade_loss(*args, pd, ld, target, pixels)
            fd = (ev(h) - ev(-h)) / (2 * h)
            ga = gpos[j, comp] if comp < 3 else glig[j, comp - 3]
            # dLoss/dout was rounded to float32 per pixel (~6e-8 relative each; the per-light sums over pixels of
            # mixed sign amplify it), plus O(h^2) truncation and the loss's rounding / h: 1e-4 relative. A dropped
            # term or a wrong sign is an O(1) error.
            assert abs(ga - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(gpos).max()), (j, comp, ga, fd)
            checked += 1
    for p in pixels[counts[refs[pixels]] > 0][:16]:
        for comp in range(4):
            def ev(delta):
                par = ar[p].astype(np.float64).copy()
                par[comp] += delta
                return co.shade_loss(*args, pos_d, lig_d, target, np.array([p], np.uint32), par[None, :])
            fd = (ev(h) - ev(-h)) / (2 * h)
            assert abs(gar[p, comp] - fd) <= 1e-4 * max(abs(fd), 1e-6), (p, comp, gar[p, comp], fd)
            checked += 1
    assert checked >= 100


# ---- Karras LBVH (Karras 2012): radix-tree property, exact unions, queries vs brute force ------------------------

def _lbvh_leaf_range(nodes, n, node, memo):
    if node >= n - 1:
        return node - (n - 1), node - (n - 1)
    if node in memo:
        return memo[node]
    a = _lbvh_leaf_range(nodes, n, int(nodes["left"][node]), memo)
    b = _lbvh_leaf_range(nodes, n, int(nodes["right"][node]), memo)
    assert a[1] + 1 == b[0], "children must cover adjacent key ranges"
    memo[node] = (a[0], b[1])
    return memo[node]


@pytest.mark.parametrize("L,bits", [(1, 5), (2, 5), (3, 5), (33, 5), (700, 5), (4096, 5), (4096, 10)])
def test_lbvh_radix_tree_property(co, L, bits):
    cam = ol.camera(64, 64)
    pos, lig = co.synth_lights(9000 + L, L)
    cf = ol.cfg(morton_bits=bits, morton_aabb_mode=1 if bits == 10 else 0,
                aabb=((-50.0, -50.0, 0.0), (50.0, 50.0, 100.0)) if bits == 10 else None)
    vp, codes, _ = co.light_bvh(pos, lig, cam, cf)
    nodes, parents = co.lbvh(codes, vp, lig)
    n = L
    keys = (codes[:, 0].astype(np.uint64) << np.uint64(32)) | codes[:, 1].astype(np.uint64)
    assert np.all(keys[1:] > keys[:-1])
    # leaves
    li = codes[:, 1]
    assert np.all(nodes["left"][n - 1:] == 0xFFFFFFFF) and np.array_equal(nodes["right"][n - 1:], li)
    assert np.array_equal(nodes["min"][n - 1:], vp[li, :3] - lig[li, 3:4])
    assert np.array_equal(nodes["max"][n - 1:], vp[li, :3] + lig[li, 3:4])
    if n == 1:
        assert parents[0] == 0xFFFFFFFF
        return
    memo = {}
    assert _lbvh_leaf_range(nodes, n, 0, memo) == (0, n - 1)   # the root covers every leaf
    assert len(memo) == n - 1                                   # every internal node is reached once
    reached = np.zeros(2 * n - 1, np.int64)
    for i in range(n - 1):
        l, r = int(nodes["left"][i]), int(nodes["right"][i])
        reached[l] += 1
        reached[r] += 1
        assert parents[l] == i and parents[r] == i
        a, b = memo[i]
        split = _lbvh_leaf_range(nodes, n, l, memo)[1]
        # the split is the first bit where the range's first and last keys differ: below it on the left,
        # set on the right (the binary radix tree over the keys is unique)
        x = int(keys[a]) ^ int(keys[b])
        top = 63 - (x.bit_length() - 1)
        bit = np.uint64(1) << np.uint64(63 - top)
        assert np.all((keys[a:split + 1] & bit) == 0) and np.all((keys[split + 1:b + 1] & bit) != 0)
        # exact union of the children
        assert np.array_equal(nodes["min"][i], np.minimum(nodes["min"][l], nodes["min"][r]))
        assert np.array_equal(nodes["max"][i], np.maximum(nodes["max"][l], nodes["max"][r]))
    assert reached[0] == 0 and np.all(reached[1:] == 1) and parents[0] == 0xFFFFFFFF


def test_lbvh_sphere_queries_match_brute_force(co):
    L = 3000
    cam = ol.camera(64, 64)
    pos, lig = co.synth_lights(4242, L)
    vp, codes, _ = co.light_bvh(pos, lig, cam, ol.cfg(morton_bits=10))
    nodes, _ = co.lbvh(codes, vp, lig)
    n = L
    lo, hi = nodes["min"].astype(np.float64), nodes["max"].astype(np.float64)
    rng = np.random.default_rng(8)
    for _ in range(60):
        c = rng.uniform(-50, 50, 3) + np.array([0.0, 0.0, 50.0])
        rad = rng.uniform(0.5, 10.0)

        def hit(node):
            d = np.maximum(lo[node] - c, 0.0) + np.maximum(c - hi[node], 0.0)
            return float(d @ d) <= rad * rad
        found, stack = [], [0]
        while stack:
            node = stack.pop()
            if not hit(node):
                continue
            if node >= n - 1:
                found.append(int(nodes["right"][node]))
            else:
                stack += [int(nodes["left"][node]), int(nodes["right"][node])]
        leaf_lo, leaf_hi = lo[n - 1:], hi[n - 1:]
        d = np.maximum(leaf_lo - c, 0.0) + np.maximum(c - leaf_hi, 0.0)
        brute = sorted(int(x) for x in nodes["right"][n - 1:][np.einsum("ij,ij->i", d, d) <= rad * rad])
        assert sorted(found) == brute


def test_shade_pixels_equals_whole_frame_shade(co):
    """The per-pixel FP64 shade (used to sample large frames) gives the whole-frame FP64 shade's values."""
    W, H, L = 96, 64, 300
    cam = ol.camera(W, H)
    cf = ol.cfg()
    pos, lig = co.synth_lights(61, L)
    depth, normals = co.synth_frame(62, W, H, cam)
    ar, me = co.synth_materials(63, W * H)
    vp, codes, bvh = co.light_bvh(pos, lig, cam, cf)
    keys, U, refs = co.unique_clusters(W, H, cam, cf, depth, normals)
    counts, offsets, idx, tot = co.assign(W, H, cam, cf, keys, U, bvh, codes, vp)
    full, mask = co.shade(W, H, cam, cf, depth, normals, ar, me, refs, U, counts, offsets, idx, pos, lig)
    pix = np.arange(0, W * H, 7, dtype=np.uint64)
    got, sh = co.shade_pixels(W, H, cam, cf, depth, normals, ar, me, refs, U, counts, offsets, idx, pos, lig, pix)
    assert np.array_equal(sh, mask[pix.astype(np.int64)]) and sh.sum() > 100
    assert np.array_equal(got[sh], full[pix.astype(np.int64)][sh, :3])
