"""GPU parity of the clustered-shading stages through the C ABI against the CPU oracle.

Integer outputs (Morton codes, BVH topology and AABBs, cluster keys and references, light lists) are compared bit
for bit; the shaded radiance is compared with the FP64 oracle at a relative tolerance of 1e-5 (absolute floor
1e-6), and the error distribution is printed per output channel.
"""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

W_FRAME, H_FRAME = 1920, 1080
L_FRAME = 4096
SEED_LIGHTS = 42 + 1000 * 1
SEED_FRAME = 42 + 1000 * 2
SHADE_RTOL, SHADE_ATOL = 1e-5, 1e-6


@pytest.fixture(scope="module")
def co(oracle):
    return ol.ClusteredOracle(oracle)


@pytest.fixture(scope="module")
def f1080(co):
    cam = ol.camera(W_FRAME, H_FRAME)
    pos, lig = co.synth_lights(SEED_LIGHTS, L_FRAME)
    depth, normals = co.synth_frame(SEED_FRAME, W_FRAME, H_FRAME, cam)
    ar, me = co.synth_materials(SEED_FRAME, W_FRAME * H_FRAME)
    return dict(cam=cam, pos=pos, lig=lig, depth=depth, normals=normals, ar=ar, me=me)


def gpu_light_bvh(vctx, co, pos, lig, cam, cf):
    L = pos.shape[0]
    vp = empty_dev(16 * L)
    codes = empty_dev(8 * L)
    nb = co.o.calc("buffer_length", L)
    bvh = empty_dev(32 * nb, fill=0xCD)
    vctx.construct_point_light_bvh(to_dev(pos), to_dev(lig), L, cam, cf, vp, codes, bvh)
    return (from_dev(vp, np.float32, (L, 4)), from_dev(codes, np.uint32, (L, 2)),
            from_dev(bvh, ol.BVH_DTYPE, (nb,)))


def test_synth_generators_match_oracle(vctx, co, f1080):
    cam = f1080["cam"]
    p, l = empty_dev(16 * L_FRAME), empty_dev(16 * L_FRAME)
    vctx.synth_lights(SEED_LIGHTS, L_FRAME, p, l)
    assert np.array_equal(from_dev(p, np.uint32, (L_FRAME, 4)), f1080["pos"].view(np.uint32))
    assert np.array_equal(from_dev(l, np.uint32, (L_FRAME, 4)), f1080["lig"].view(np.uint32))
    d, n = empty_dev(4 * W_FRAME * H_FRAME), empty_dev(8 * W_FRAME * H_FRAME)
    vctx.synth_frame(SEED_FRAME, W_FRAME, H_FRAME, cam, 64, d, n)
    assert np.array_equal(from_dev(d, np.uint32, (W_FRAME * H_FRAME,)), f1080["depth"].view(np.uint32))
    assert np.array_equal(from_dev(n, np.uint16, (W_FRAME * H_FRAME, 4)), f1080["normals"])
    a, m = empty_dev(16 * W_FRAME * H_FRAME), empty_dev(4 * W_FRAME * H_FRAME)
    vctx.synth_materials(SEED_FRAME, W_FRAME * H_FRAME, a, m)
    assert np.array_equal(from_dev(a, np.uint32, (W_FRAME * H_FRAME, 4)), f1080["ar"].view(np.uint32))
    assert np.array_equal(from_dev(m, np.uint32, (W_FRAME * H_FRAME,)), f1080["me"].view(np.uint32))


def test_camera_matrices_match_oracle(vctx, co):
    for cam in (ol.camera(W_FRAME, H_FRAME), ol.camera(640, 360, fov_deg=60.0, pos=(3.0, -2.0, 7.5))):
        v1, p1 = vctx.camera_matrices(cam)
        v2, p2 = co.camera_matrices(cam)
        assert np.array_equal(v1, v2) and np.array_equal(p1, p2)


@pytest.mark.parametrize("mode", ["reference15", "scene30"])
def test_light_bvh_4096_lights(vctx, co, f1080, mode):
    cam = f1080["cam"]
    if mode == "reference15":
        cf = ol.cfg()
    else:  # 30-bit codes on the scene AABB ([-50,50]^3 world = view z in [0,100] for this camera)
        cf = ol.cfg(morton_bits=10, morton_aabb_mode=1, aabb=((-50, -50, 0), (50, 50, 100)))
    vp, codes, bvh = gpu_light_bvh(vctx, co, f1080["pos"], f1080["lig"], cam, cf)
    ovp, ocodes, obvh = co.light_bvh(f1080["pos"], f1080["lig"], cam, cf)
    assert np.array_equal(vp.view(np.uint32), ovp.view(np.uint32))
    assert np.array_equal(codes, ocodes)
    assert np.array_equal(bvh.view(np.uint32).reshape(-1, 8)[:, :7], obvh.view(np.uint32).reshape(-1, 8)[:, :7])


def test_bucket_sort_light_morton_codes(vctx, co, f1080):
    # the BucketSort of 4096 lights into 64 buckets, on the lights' own 30-bit Morton codes (scene AABB)
    # rather than random keys: (code, light index) pairs in light order, bucket = the code's top 6 bits
    cam = f1080["cam"]
    cf = ol.cfg(morton_bits=10, morton_aabb_mode=1, aabb=((-50, -50, 0), (50, 50, 100)))
    _, codes, _ = gpu_light_bvh(vctx, co, f1080["pos"], f1080["lig"], cam, cf)
    L = codes.shape[0]
    pairs = np.zeros((L, 2), np.uint32)
    pairs[codes[:, 1]] = codes                       # back to light order
    assert np.array_equal(pairs[:, 1], np.arange(L, dtype=np.uint32))
    assert int(pairs[:, 0].max()) < (1 << 30)
    out = empty_dev(8 * L)
    offs = empty_dev(4 * 64)
    vctx.bucket_sort(to_dev(pairs), L, out, offs, 64, 24)
    opairs, ooffs = co.o.bucket_sort(pairs, 64, 24)
    got, goffs = from_dev(out, np.uint32, (L, 2)), from_dev(offs, np.uint32, (64,))
    assert np.array_equal(got, opairs) and np.array_equal(goffs, ooffs)
    # the buckets partition the Morton order: bucket b holds exactly the codes with top bits b, in light order
    for b in range(64):
        lo, hi = int(goffs[b]), int(goffs[b + 1]) if b < 63 else L
        seg = got[lo:hi]
        assert ((seg[:, 0] >> 24) == b).all() and (np.diff(seg[:, 1].astype(np.int64)) > 0).all()
    assert len(np.unique(pairs[:, 0] >> 24)) > 8   # the codes spread over many buckets


@pytest.mark.parametrize("L", [1, 31, 32, 33, 1000, 4097, 40000])
def test_light_bvh_sizes(vctx, co, L):
    cam = ol.camera(640, 360)
    pos, lig = co.synth_lights(777, L)
    vp, codes, bvh = gpu_light_bvh(vctx, co, pos, lig, cam, ol.cfg())
    ovp, ocodes, obvh = co.light_bvh(pos, lig, cam, ol.cfg())
    assert np.array_equal(codes, ocodes), L
    assert np.array_equal(bvh.view(np.uint32).reshape(-1, 8)[:, :7], obvh.view(np.uint32).reshape(-1, 8)[:, :7]), L


@pytest.mark.parametrize("L,bits", [(1, 5), (2, 5), (3, 5), (33, 5), (1000, 5), (4096, 5), (4096, 10),
                                    (40000, 10)])
def test_lbvh_matches_oracle(vctx, co, L, bits):
    """Karras LBVH over sorted light codes: topology, parents and refitted AABBs bit-exact against the oracle."""
    cam = ol.camera(64, 64)
    pos, lig = co.synth_lights(5000 + L, L)
    cf = ol.cfg(morton_bits=bits)
    vp, codes, _ = co.light_bvh(pos, lig, cam, cf)
    m = 2 * L - 1
    nodes, parents = empty_dev(32 * m, fill=0xCD), empty_dev(4 * m, fill=0xCD)
    vctx.build_lbvh(to_dev(codes), L, to_dev(vp), to_dev(lig), nodes, parents)
    onodes, oparents = co.lbvh(codes, vp, lig)
    assert np.array_equal(from_dev(parents, np.uint32, (m,)), oparents)
    assert np.array_equal(from_dev(nodes, np.uint32, (m, 8)), onodes.view(np.uint32).reshape(m, 8))


def gpu_clusters(vctx, W, H, cam, cf, depth, normals, key_cap=1 << 20):
    keys = empty_dev(4 * key_cap)
    cnt = empty_dev(4)
    refs = empty_dev(4 * W * H, fill=0xEE)
    vctx.find_unique_cluster_list(W, H, cam, cf, to_dev(depth), to_dev(normals), keys, key_cap, cnt, refs)
    U = int(from_dev(cnt, np.uint32)[0])
    return keys, from_dev(keys, np.uint32, (min(U, key_cap),)), U, from_dev(refs, np.uint32, (W * H,)), cnt, refs


@pytest.mark.parametrize("case", ["1080p_reference", "odd_sliced", "odd_reference"])
def test_unique_clusters(vctx, co, f1080, case):
    if case == "1080p_reference":
        W, H, cam, cf = W_FRAME, H_FRAME, f1080["cam"], ol.cfg()
        depth, normals = f1080["depth"], f1080["normals"]
    else:
        W, H = 1000, 700
        cam = ol.camera(W, H, fov_deg=75.0)
        depth, normals = co.synth_frame(555, W, H, cam)
        cf = ol.cfg(key_mode=1, depth_slices=32) if case == "odd_sliced" else ol.cfg()
    _, keys, U, refs, _, _ = gpu_clusters(vctx, W, H, cam, cf, depth, normals)
    okeys, oU, orefs = co.unique_clusters(W, H, cam, cf, depth, normals)
    assert U == oU
    assert np.array_equal(keys, okeys)
    assert np.array_equal(refs, orefs)


@pytest.mark.parametrize("mode", ["reference", "sliced"])
def test_unique_clusters_slice_boundaries(vctx, co, mode):
    """Depth buffers of float bit-pattern sweeps, keys and references bit for bit against the oracle's formula:
    a coarse sweep over (0, 1) reaches every depth slice, consecutive floats just below 1.0 cross the far slices'
    boundaries between neighbouring floats, depths above 1.0 drive the view depth past far, through +inf to
    negative values (beyond the threshold table), and NaN, +-inf, -0 and subnormal depths close the buffer."""
    W, H = 1024, 1024
    n = W * H
    cam = ol.camera(W, H)
    q = n // 4
    coarse = np.linspace(0x34000000, 0x3F7FFFFF, q).astype(np.uint32)
    below_one = np.uint32(0x3F7FFFFF) - np.arange(q, dtype=np.uint32)
    above_one = np.uint32(0x3F800000) + np.arange(q, dtype=np.uint32) * np.uint32(64)
    # the depth where the inverse projection's w row vanishes, and 1024 floats on each side: view depths of huge
    # magnitude up to +-inf (log slices >= 1023 wrap in the 10-bit field)
    inv, _ = co.frame_consts(W, H, cam)
    d0 = (-np.float32(inv[15]) / np.float32(inv[11])).astype(np.float32).view(np.uint32)
    pole = (d0.astype(np.int64) + np.arange(-1024, 1024)).astype(np.uint32)
    special = np.array([0x7FC00000, 0x7F800000, 0xFF800000, 0x80000000, 0x00000001, 0x007FFFFF, 0x00000000,
                        0x3F800000], np.uint32)
    rest = np.resize(np.concatenate([pole, special]), n - 3 * q)
    depth = np.concatenate([coarse, below_one, above_one, rest]).view(np.float32)
    _, normals = co.synth_frame(777, W, H, cam)
    cf = ol.cfg(key_mode=1, depth_slices=200) if mode == "sliced" else ol.cfg()
    _, keys, U, refs, _, _ = gpu_clusters(vctx, W, H, cam, cf, depth, normals)
    okeys, oU, orefs = co.unique_clusters(W, H, cam, cf, depth, normals)
    assert U == oU
    assert np.array_equal(keys, okeys)
    assert np.array_equal(refs, orefs)
    shift, width = (16, 8) if mode == "sliced" else (16, 10)
    slices = np.unique((okeys >> shift) & ((1 << width) - 1))
    assert slices.size == 200 if mode == "sliced" else slices.size > 150, slices.size


@pytest.mark.parametrize("mode", ["reference", "sliced"])
def test_unique_clusters_every_f16_normal_component(vctx, co, mode):
    """Normals holding every f16 bit pattern (zeros, subnormals, the largest finite, +-inf, NaNs) as one component,
    rotating over the three axes, with the other two components random and mostly smaller in magnitude (so that
    pattern is the dominant axis and the key divides by it), a quarter of them random over the whole f16 range;
    keys and references bit for bit against the oracle."""
    W = H = 256
    n = W * H
    cam = ol.camera(W, H)
    rng = np.random.default_rng(4321)
    h = np.arange(n, dtype=np.uint32).astype(np.uint16)
    hv = h.view(np.float16).astype(np.float32)
    finite = np.isfinite(hv)
    normals = np.zeros((n, 4), np.uint16)
    axis = np.arange(n) % 3
    for o in (1, 2):
        with np.errstate(all="ignore"):
            small = (rng.uniform(-1.0, 1.0, n).astype(np.float32) * np.where(finite, hv, 1.0)).astype(np.float16)
        wide = rng.integers(0, 1 << 16, n, dtype=np.uint32).astype(np.uint16)
        other = np.where(rng.uniform(size=n) < 0.25, wide, small.view(np.uint16))
        normals[np.arange(n), (axis + o) % 3] = other
    normals[np.arange(n), axis] = h
    depth, _ = co.synth_frame(778, W, H, cam)
    cf = ol.cfg(key_mode=1, depth_slices=24) if mode == "sliced" else ol.cfg()
    _, keys, U, refs, _, _ = gpu_clusters(vctx, W, H, cam, cf, depth, normals)
    okeys, oU, orefs = co.unique_clusters(W, H, cam, cf, depth, normals)
    assert U == oU
    assert np.array_equal(keys, okeys)
    assert np.array_equal(refs, orefs)


def test_unique_clusters_capacity_flag(vctx, co, f1080):
    from vren_amd import VrenError
    _, keys, U, refs, _, _ = gpu_clusters(vctx, W_FRAME, H_FRAME, f1080["cam"], ol.cfg(), f1080["depth"], f1080["normals"], key_cap=100)
    okeys, oU, orefs = co.unique_clusters(W_FRAME, H_FRAME, f1080["cam"], ol.cfg(), f1080["depth"], f1080["normals"], key_cap=100)
    assert U == oU and U > 100 and np.array_equal(keys, okeys) and np.array_equal(refs, orefs)
    with pytest.raises(VrenError) as e:
        vctx.sync_status()
    assert e.value.status == -2
    vctx.sync_status()  # flags were cleared


def run_assign(vctx, co, W, H, cam, cf, depth, normals, pos, lig, key_cap=1 << 20, index_cap=1 << 22):
    L = pos.shape[0]
    dpos, dlig = to_dev(pos), to_dev(lig)
    vp = empty_dev(16 * L)
    codes = empty_dev(8 * L)
    nb = co.o.calc("buffer_length", L)
    bvh = empty_dev(32 * nb)
    vctx.construct_point_light_bvh(dpos, dlig, L, cam, cf, vp, codes, bvh)
    dkeys, keys, U, refs, cnt, drefs = gpu_clusters(vctx, W, H, cam, cf, depth, normals, key_cap)
    counts, offsets = empty_dev(4 * key_cap), empty_dev(4 * key_cap)
    idx, tot = empty_dev(4 * index_cap), empty_dev(4)
    vctx.assign_lights(W, H, cam, cf, dkeys, cnt, key_cap, bvh, L, codes, vp, counts, offsets, idx, index_cap, tot)
    total = int(from_dev(tot, np.uint32)[0])
    n = min(U, key_cap)
    return dict(U=U, keys=keys, refs=refs, counts=from_dev(counts, np.uint32, (n,)),
                offsets=from_dev(offsets, np.uint32, (n,)), idx=from_dev(idx, np.uint32, (min(total, index_cap),)),
                total=total, dev=dict(pos=dpos, lig=dlig, refs=drefs, counts=counts, offsets=offsets, idx=idx),
                vp=from_dev(vp, np.float32, (L, 4)), codes=from_dev(codes, np.uint32, (L, 2)),
                bvh=from_dev(bvh, ol.BVH_DTYPE, (nb,)))


ENCODERS_1080P = {"reference": dict(), "sliced32": dict(key_mode=1, depth_slices=32)}


@pytest.mark.parametrize("encoder", ["reference", "sliced32"])
@pytest.mark.parametrize("aabb_mode", [0, 1])
def test_assign_lights_1080p(vctx, co, f1080, aabb_mode, encoder):
    """1920x1080 with 4096 lights, with the reference key encoder and with the sliced one (32x32 tiles x 32
    log-depth slices over [near, far] x 6 cube faces): counts, offsets and lists bit for bit, both AABB modes."""
    cam = f1080["cam"]
    cf = ol.cfg(cluster_aabb_mode=aabb_mode, **ENCODERS_1080P[encoder])
    g = run_assign(vctx, co, W_FRAME, H_FRAME, cam, cf, f1080["depth"], f1080["normals"], f1080["pos"], f1080["lig"])
    ovp, ocodes, obvh = co.light_bvh(f1080["pos"], f1080["lig"], cam, cf)
    okeys, oU, orefs = co.unique_clusters(W_FRAME, H_FRAME, cam, cf, f1080["depth"], f1080["normals"])
    oc, oo, oi, ot = co.assign(W_FRAME, H_FRAME, cam, cf, okeys, oU, obvh, ocodes, ovp)
    assert g["U"] == oU and np.array_equal(g["keys"], okeys) and np.array_equal(g["refs"], orefs)
    assert g["total"] == ot
    assert np.array_equal(g["counts"], oc)
    assert np.array_equal(g["offsets"], oo)
    assert np.array_equal(g["idx"], oi)
    print("\n[f1080 %s aabb=%d] %d clusters, %d assignments" % (encoder, aabb_mode, oU, ot))


def test_assign_lights_many_lights(vctx, co):
    # a frame with 40 000 lights: BVH of 2^20 leaf slots (4 levels), LDS bitmask of 1250 words per wave
    W, H = 640, 360
    cam = ol.camera(W, H)
    pos, lig = co.synth_lights(4242, 40000)
    depth, normals = co.synth_frame(4243, W, H, cam)
    cf = ol.cfg()
    g = run_assign(vctx, co, W, H, cam, cf, depth, normals, pos, lig)
    ovp, ocodes, obvh = co.light_bvh(pos, lig, cam, cf)
    okeys, oU, orefs = co.unique_clusters(W, H, cam, cf, depth, normals)
    oc, oo, oi, ot = co.assign(W, H, cam, cf, okeys, oU, obvh, ocodes, ovp)
    assert g["total"] == ot and np.array_equal(g["counts"], oc) and np.array_equal(g["idx"], oi)


def shade_error_report(got, ref, mask):
    g = got[mask, :3].astype(np.float64)
    r = ref[mask, :3]
    err = np.abs(g - r) / np.maximum(np.abs(r), SHADE_ATOL / SHADE_RTOL)
    return err


@pytest.mark.parametrize("atten,encoder", [(0, "reference"), (1, "reference"), (0, "sliced32"), (1, "sliced32")])
def test_shade_1080p_vs_fp64(vctx, co, f1080, atten, encoder):
    cam = f1080["cam"]
    cf = ol.cfg(attenuation=atten, **ENCODERS_1080P[encoder])
    g = run_assign(vctx, co, W_FRAME, H_FRAME, cam, ol.cfg(**ENCODERS_1080P[encoder]), f1080["depth"], f1080["normals"], f1080["pos"],
                   f1080["lig"])
    out = empty_dev(16 * W_FRAME * H_FRAME, fill=0xFF)  # NaN sentinel: background pixels must stay untouched
    key_cap = 1 << 20
    vctx.shade(W_FRAME, H_FRAME, cam, cf, to_dev(f1080["depth"]), to_dev(f1080["normals"]), to_dev(f1080["ar"]), to_dev(f1080["me"]),
               g["dev"]["refs"], key_cap, g["dev"]["counts"], g["dev"]["offsets"], g["dev"]["idx"], 1 << 22,
               g["dev"]["pos"], g["dev"]["lig"], L_FRAME, out)
    got = from_dev(out, np.float32, (W_FRAME * H_FRAME, 4))
    cap_counts = np.zeros(key_cap, np.uint32)
    cap_offs = np.zeros(key_cap, np.uint32)
    cap_counts[:g["counts"].size] = g["counts"]
    cap_offs[:g["offsets"].size] = g["offsets"]
    ref, mask = co.shade(W_FRAME, H_FRAME, cam, cf, f1080["depth"], f1080["normals"], f1080["ar"], f1080["me"], g["refs"], key_cap,
                         cap_counts, cap_offs, g["idx"], f1080["pos"], f1080["lig"])
    assert np.all(np.isnan(got[~mask]))               # background untouched (shade.comp:86)
    assert np.all(got[mask, 3] == 1.0)
    err = shade_error_report(got, ref, mask)
    print("\nshade %s atten=%d: %d shaded outputs; rel err max %.3e p99.9 %.3e p50 %.3e per channel max %s"
          % (encoder, atten, err.size, err.max(), np.quantile(err, 0.999), np.median(err),
             ["%.2e" % x for x in err.max(axis=0)]))
    assert err.max() <= SHADE_RTOL, err.max()


def make_frames_batch(vctx, co, specs):
    """specs: list of (W, H, L, seed). Inputs generated on the device; returns (frame descs, keep-alive tensors)."""
    import vren_amd
    keep, descs = [], []
    for (W, H, L, seed) in specs:
        cam = ol.camera(W, H)
        d, n = empty_dev(4 * W * H), empty_dev(8 * W * H)
        a, m = empty_dev(16 * W * H), empty_dev(4 * W * H)
        p, l = empty_dev(16 * max(L, 1)), empty_dev(16 * max(L, 1))
        o = empty_dev(16 * W * H, fill=0xFF)
        vctx.synth_frame(seed, W, H, cam, 64, d, n)
        vctx.synth_materials(seed, W * H, a, m)
        vctx.synth_lights(seed + 7, L, p, l)
        fd = vren_amd.FrameDesc()
        fd.width, fd.height, fd.camera = W, H, cam
        fd.depth, fd.normals, fd.albedo_roughness, fd.metallic = (d.data_ptr(), n.data_ptr(), a.data_ptr(),
                                                                   m.data_ptr())
        fd.light_positions, fd.point_lights, fd.light_count = p.data_ptr(), l.data_ptr(), L
        fd.directional_lights, fd.directional_light_count, fd.output = None, 0, o.data_ptr()
        fd.key_capacity, fd.index_capacity = 0, 0
        descs.append(fd)
        keep.append(dict(d=d, n=n, a=a, m=m, p=p, l=l, o=o, cam=cam, W=W, H=H, L=L))
    return descs, keep


def test_batch_modes_identical_and_match_oracle(vctx, co):
    import torch
    import vren_amd
    specs = [(640, 360, 256, 4000), (1280, 720, 5000, 4001), (333, 200, 1, 4002), (1920, 1080, 4096, 4003),
             (160, 90, 0, 4004)]
    descs, keep = make_frames_batch(vctx, co, specs)
    b = vren_amd.Batch(vctx, descs, ol.cfg())
    results = []
    for mode in (b.PER_FRAME, b.BATCHED, b.GRAPH, b.GRAPH):
        for k in keep:
            k["o"].fill_(0xFF)
        b.run(mode)
        torch.cuda.synchronize()
        results.append([from_dev(k["o"], np.uint32).copy() for k in keep])
    for r in results[1:]:
        for a, c in zip(results[0], r):
            assert np.array_equal(a, c)
    vctx.sync_status()
    # frame 0 intermediates and output against the oracle
    k = keep[0]
    W, H, L = k["W"], k["H"], k["L"]
    depth = from_dev(k["d"], np.float32, (W * H,))
    normals = from_dev(k["n"], np.uint16, (W * H, 4))
    pos = from_dev(k["p"], np.float32, (L, 4))
    lig = from_dev(k["l"], np.float32, (L, 4))
    ar = from_dev(k["a"], np.float32, (W * H, 4))
    me = from_dev(k["m"], np.float32, (W * H,))
    okeys, oU, orefs = co.unique_clusters(W, H, k["cam"], ol.cfg(), depth, normals)
    ovp, ocodes, obvh = co.light_bvh(pos, lig, k["cam"], ol.cfg())
    oc, oo, oi, ot = co.assign(W, H, k["cam"], ol.cfg(), okeys, oU, obvh, ocodes, ovp)
    o = b.outputs(0)
    dev_u = _read_ptr(o.cluster_count, 4, np.uint32)
    assert int(dev_u[0]) == oU
    assert np.array_equal(_read_ptr(o.cluster_refs, 4 * W * H, np.uint32), orefs)
    assert np.array_equal(_read_ptr(o.light_indices, 4 * int(ot), np.uint32), oi)
    cap_counts = np.zeros(o.key_capacity, np.uint32)
    cap_offs = np.zeros(o.key_capacity, np.uint32)
    cap_counts[:oc.size] = oc
    cap_offs[:oo.size] = oo
    ref, mask = co.shade(W, H, k["cam"], ol.cfg(), depth, normals, ar, me, orefs, o.key_capacity, cap_counts,
                         cap_offs, oi, pos, lig)
    got = results[0][0].view(np.float32).reshape(-1, 4)
    err = shade_error_report(got, ref, mask)
    assert err.max() <= SHADE_RTOL
    # frame 2 (333x200: 77 tiles, padded to a multiple of 8 in the batch's tile numbering): keys and references
    k = keep[2]
    W, H = k["W"], k["H"]
    depth = from_dev(k["d"], np.float32, (W * H,))
    normals = from_dev(k["n"], np.uint16, (W * H, 4))
    okeys, oU, orefs = co.unique_clusters(W, H, k["cam"], ol.cfg(), depth, normals)
    o = b.outputs(2)
    assert int(_read_ptr(o.cluster_count, 4, np.uint32)[0]) == oU
    assert np.array_equal(_read_ptr(o.cluster_keys, 4 * oU, np.uint32), okeys)
    assert np.array_equal(_read_ptr(o.cluster_refs, 4 * W * H, np.uint32), orefs)
    # frame 4 has no lights: the radiance is the ambient term alone, (0.03 albedo)^(1/2.2) on every shaded pixel,
    # and the background stays untouched
    k = keep[4]
    W, H = k["W"], k["H"]
    depth = from_dev(k["d"], np.float32, (W * H,))
    normals = from_dev(k["n"], np.uint16, (W * H, 4))
    ar = from_dev(k["a"], np.float32, (W * H, 4))
    me = from_dev(k["m"], np.float32, (W * H,))
    _, _, orefs = co.unique_clusters(W, H, k["cam"], ol.cfg(), depth, normals)
    kc = 1 << 20
    zero = np.zeros(kc, np.uint32)
    ref, mask = co.shade(W, H, k["cam"], ol.cfg(), depth, normals, ar, me, orefs, kc, zero, zero,
                         np.zeros(1, np.uint32), np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32))
    got = results[0][4].view(np.float32).reshape(-1, 4)
    assert np.all(np.isnan(got[~mask])) and mask.sum() > 0
    amb = (0.03 * ar[mask, :3].astype(np.float64)) ** (1.0 / 2.2)
    assert np.abs(got[mask, :3] - amb).max() <= SHADE_RTOL * amb.max()
    assert shade_error_report(got, ref, mask).max() <= SHADE_RTOL
    b.close()


def _batch_host_frame(b, keep, i):
    k = keep[i]
    W, H, L = k["W"], k["H"], k["L"]
    o = b.outputs(i)
    U = int(_read_ptr(o.cluster_count, 4, np.uint32)[0])
    tot = int(_read_ptr(o.assigned_total, 4, np.uint32)[0])
    return dict(W=W, H=H, L=L, cam=k["cam"], depth=from_dev(k["d"], np.float32, (W * H,)),
                normals=from_dev(k["n"], np.uint16, (W * H, 4)), pos=from_dev(k["p"], np.float32, (L, 4)),
                lig=from_dev(k["l"], np.float32, (L, 4)), ar=from_dev(k["a"], np.float32, (W * H, 4)),
                me=from_dev(k["m"], np.float32, (W * H,)), out=from_dev(k["o"], np.float32, (W * H, 4)),
                U=U, keys=_read_ptr(o.cluster_keys, 4 * U, np.uint32).copy(),
                refs=_read_ptr(o.cluster_refs, 4 * W * H, np.uint32).copy(),
                counts=_read_ptr(o.light_counts, 4 * U, np.uint32).copy(),
                offsets=_read_ptr(o.light_offsets, 4 * U, np.uint32).copy(), total=tot,
                idx=_read_ptr(o.light_indices, 4 * tot, np.uint32).copy(), key_capacity=o.key_capacity)


def _batch_check_frame(co, cf, i, f, n_shade=2048):
    """One frame of a batch against the oracle: keys, references, counts, offsets and lists bit for bit; the
    radiance of n_shade sampled pixels within 1e-5 of FP64 (background pixels untouched)."""
    W, H, L, cam = f["W"], f["H"], f["L"], f["cam"]
    okeys, oU, orefs = co.unique_clusters(W, H, cam, cf, f["depth"], f["normals"])
    assert f["U"] == oU and np.array_equal(f["keys"], okeys) and np.array_equal(f["refs"], orefs), i
    ovp, ocodes, obvh = co.light_bvh(f["pos"], f["lig"], cam, cf)
    oc, oo, oi, ot = co.assign(W, H, cam, cf, okeys, oU, obvh, ocodes, ovp)
    assert f["total"] == ot, i
    assert np.array_equal(f["counts"], oc) and np.array_equal(f["offsets"], oo), i
    assert np.array_equal(f["idx"], oi), i
    cc = np.zeros(f["key_capacity"], np.uint32)
    cco = np.zeros(f["key_capacity"], np.uint32)
    cc[:oU], cco[:oU] = oc, oo
    pix = np.random.default_rng(i).choice(W * H, min(n_shade, W * H), replace=False).astype(np.uint64)
    ref, sh = co.shade_pixels(W, H, cam, cf, f["depth"], f["normals"], f["ar"], f["me"], orefs, f["key_capacity"],
                              cc, cco, oi, f["pos"], f["lig"], pix)
    got = f["out"][pix.astype(np.int64)]
    assert np.all(np.isnan(got[~sh])), i
    err = shade_error_report(got[sh], ref[sh], np.ones(int(sh.sum()), bool))
    return oU, int(ot), int(sh.sum()), float(err.max()) if err.size else 0.0


def _batch_check_all(b, keep, co, cf, tag):
    import concurrent.futures
    results = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        for c0 in range(0, len(keep), 8):   # 8 frames on the host at a time
            frames = [(i, _batch_host_frame(b, keep, i)) for i in range(c0, min(c0 + 8, len(keep)))]
            results += list(ex.map(lambda a: _batch_check_frame(co, cf, *a), frames))
            del frames
    worst = max(r[3] for r in results)
    print("[%s] all %d frames bit-exact: %d clusters, %d assignments; shade sample %d px, rel err max %.2e"
          % (tag, len(keep), sum(r[0] for r in results), sum(r[1] for r in results), sum(r[2] for r in results),
             worst), flush=True)
    assert worst <= SHADE_RTOL, worst


def _batch_all_modes_identical(b, keep):
    import torch
    first = None
    for mode in (b.PER_FRAME, b.BATCHED, b.GRAPH, b.GRAPH):
        for k in keep:
            k["o"].fill_(0xFF)
        b.run(mode)
        torch.cuda.synchronize()
        if first is None:
            first = [k["o"].clone() for k in keep]
        else:
            for i, (a, k) in enumerate(zip(first, keep)):
                assert torch.equal(a, k["o"]), (mode, i)


def test_batch_64_heterogeneous_frames(vctx, co):
    """The batched configuration at full size: 64 frames, resolutions drawn from {640x360, 1280x720, 1920x1080,
    2560x1440, 3840x2160}, light counts log-uniform in [256, 65536]. Per-frame, batched and graph launches give the
    same bytes for every frame (compared on the device). Every frame is checked against the oracle: keys,
    references, light counts, offsets and lists bit for bit, and the radiance of 2048 sampled pixels per frame
    within 1e-5 of FP64."""
    import vren_amd
    rng = np.random.default_rng(4444)
    res = [(640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]
    specs = []
    for f in range(64):
        W, H = res[rng.integers(0, len(res))]
        L = int(round(float(np.exp(rng.uniform(np.log(256.0), np.log(65536.0))))))
        specs.append((W, H, L, 7000 + f))
    print("\n[batch64] %d frames, %.1f M pixels, lights %d..%d" % (
        len(specs), sum(w * h for w, h, _, _ in specs) / 1e6, min(s[2] for s in specs), max(s[2] for s in specs)),
          flush=True)
    descs, keep = make_frames_batch(vctx, co, specs)
    b = vren_amd.Batch(vctx, descs, ol.cfg())
    _batch_all_modes_identical(b, keep)
    vctx.sync_status()
    _batch_check_all(b, keep, co, ol.cfg(), "batch64")
    b.close()


def test_batch_long_lists_and_many_clusters(vctx, co):
    """The two ways a batch's writing pass gets a light list: copied from the slot the counting pass kept (lists of
    at most 192 lights, for the first 32 768 clusters of a frame) or traversed again. Frame 0 (640x360, 3000
    lights with five times their radius) has lists of 65-192 and of more than 192 lights; frame 1 (1920x1080 with
    its depth and normals shuffled over the pixels, 2000 lights) has more than 32 768 unique clusters, those past
    32 768 with lights. Counts, offsets and lists bit for bit against the oracle after per-frame and after graph
    launches."""
    import torch
    import vren_amd
    specs = [(640, 360, 3000, 5100), (1920, 1080, 2000, 5200)]
    descs, keep = make_frames_batch(vctx, co, specs)
    keep[0]["l"].view(torch.float32).view(3000, 4)[:, 3] *= 5.0
    W, H = 1920, 1080
    perm = torch.from_numpy(np.random.default_rng(5).permutation(W * H)).cuda()
    d = keep[1]["d"].view(torch.float32)
    d.copy_(d[perm])
    n = keep[1]["n"].view(torch.int16).view(W * H, 4)
    n.copy_(n[perm])
    torch.cuda.synchronize()
    b = vren_amd.Batch(vctx, descs, ol.cfg())
    for mode in (b.PER_FRAME, b.GRAPH):
        for k in keep:
            k["o"].fill_(0xFF)
        b.run(mode)
        torch.cuda.synchronize()
        vctx.sync_status()
        f0, f1 = _batch_host_frame(b, keep, 0), _batch_host_frame(b, keep, 1)
        c0 = f0["counts"]
        assert ((c0 > 64) & (c0 <= 192)).sum() > 0 and (c0 > 192).sum() > 0, c0.max()
        assert f1["U"] > 32768 and f1["counts"][32768:].sum() > 0, f1["U"]
        # frame 0's tiles stage up to ~1000 list entries, more than the shade's 256-entry stage (windows)
        _batch_check_frame(co, ol.cfg(), 0, f0, n_shade=16384)
        _batch_check_frame(co, ol.cfg(), 1, f1, n_shade=512)
        print("\n[long lists, mode %d] frame 0: %d clusters, %d of 65-192 lights, %d over 192 (max %d); frame 1: %d "
              "clusters" % (mode, f0["U"], ((c0 > 64) & (c0 <= 192)).sum(), (c0 > 192).sum(), c0.max(), f1["U"]),
              flush=True)
    b.close()


@pytest.mark.parametrize("aabb_mode", [0, 1])
def test_batch_sliced_encoder(vctx, co, aabb_mode):
    """The sliced key encoder (32 log-depth slices x 6 cube faces) through the batched chain: 1080p
    with 4096 lights and three other frames; every launch mode gives the same bytes and every frame matches the
    oracle (keys, references, lists; shade samples within 1e-5 of FP64)."""
    import vren_amd
    specs = [(1920, 1080, 4096, 8100), (640, 360, 700, 8101), (1280, 720, 20000, 8102), (333, 200, 3, 8103)]
    cf = ol.cfg(key_mode=1, depth_slices=32, cluster_aabb_mode=aabb_mode)
    descs, keep = make_frames_batch(vctx, co, specs)
    b = vren_amd.Batch(vctx, descs, cf)
    _batch_all_modes_identical(b, keep)
    vctx.sync_status()
    _batch_check_all(b, keep, co, cf, "batch sliced aabb=%d" % aabb_mode)
    b.close()


def _read_ptr(ptr, nbytes, dtype):
    """Copy nbytes from a raw device pointer to the host (hipMemcpy through the HIP runtime torch loaded)."""
    import ctypes
    import torch
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.zeros(max(nbytes, 1), np.uint8)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes),
                       ctypes.c_int(2))
    assert rc == 0
    return out[:nbytes].view(dtype)


def test_batch_light_count_sequence_and_animation(vctx, co):
    """A frame sequence through the batched chain. (a) Light counts that change between frames: a batch is
    specialised to its frames' light counts (its buffer sizes, BVH shapes, the BVH padding written at creation and
    the captured graph), so a new count takes a new batch. The sequence 4096 -> 1000 -> 1 -> 40000 -> 0 -> 4096
    rebuilds it each time, and every frame of every batch matches the oracle. (b) With the count fixed, the lights
    are moved in place by bounce_point_lights between graph replays (the reference demo's per-frame animation
    ahead of cluster_and_shade): every replay reads the moved lights and matches the oracle."""
    import torch
    import vren_amd
    cf = ol.cfg()
    for step, L in enumerate([4096, 1000, 1, 40000, 0, 4096]):
        specs = [(640, 360, L, 9000 + 10 * step), (1280, 720, L, 9001 + 10 * step)]
        descs, keep = make_frames_batch(vctx, co, specs)
        b = vren_amd.Batch(vctx, descs, cf)
        b.run(b.GRAPH)
        vctx.sync_status()
        _batch_check_all(b, keep, co, cf, "light count %d" % L)
        b.close()
    # (b) animation between replays of one captured graph
    specs = [(640, 360, 3000, 9100), (1920, 1080, 4096, 9101)]
    descs, keep = make_frames_batch(vctx, co, specs)
    b = vren_amd.Batch(vctx, descs, cf)
    rng = np.random.default_rng(9102)
    dirs = []
    for k in keep:
        d = rng.normal(size=(k["L"], 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        dirs.append(to_dev(np.concatenate([d, np.zeros((k["L"], 1))], 1).astype(np.float32)))
    for frame in range(3):
        if frame:
            for k, dd in zip(keep, dirs):
                vctx.bounce_point_lights(k["p"], dd, k["L"], (-50.0, -50.0, -50.0), (50.0, 50.0, 50.0), 10.0,
                                         1.0 / 60.0)
        for k in keep:
            k["o"].fill_(0xFF)
        b.run(b.GRAPH)
        torch.cuda.synchronize()
        vctx.sync_status()
        _batch_check_all(b, keep, co, cf, "animated frame %d" % frame)
    b.close()


def test_batch_chain_shapes_identical_and_autotuned(vctx, co, tmp_path):
    """Every chain shape (shade strips per workgroup, persistent assignment waves per CU, light-sort digit width)
    gives the same bytes as the default, in graph replays too; the autotuner keeps one of them for the batch and
    caches it for the size class (a new batch over the same frames starts with it), and the choice survives an
    autotune save / load into another context."""
    import torch
    import vren_amd
    specs = [(1280, 720, 5000, 9300), (640, 360, 300, 9301), (333, 200, 1, 9302)]
    descs, keep = make_frames_batch(vctx, co, specs)
    b = vren_amd.Batch(vctx, descs, ol.cfg())
    assert b.shape() == (0, 0, 0, 0)
    b.run(b.BATCHED)
    torch.cuda.synchronize()
    want = [k["o"].clone() for k in keep]
    for shape in [(1, 8, 4, 0), (2, 16, 8, 1), (2, 32, 4, 0), (1, 24, 8, 1), (2, 8, 8, 1), (1, 32, 8, 1)]:
        b.set_shape(*shape)
        assert b.shape() == shape
        for mode in (b.BATCHED, b.GRAPH, b.GRAPH):
            for k in keep:
                k["o"].fill_(0xFF)
            b.run(mode)
            torch.cuda.synchronize()
            for i, (a, k) in enumerate(zip(want, keep)):
                assert torch.equal(a, k["o"]), (shape, mode, i)
    with pytest.raises(vren_amd.VrenError):
        b.set_shape(3, 8, 4)
    with pytest.raises(vren_amd.VrenError):
        b.set_shape(1, 8, 4, 2)
    best, ms = b.autotune(reps=1)
    assert best[0] in (1, 2) and best[1] in (8, 16, 24, 32) and best[2] in (4, 8) and best[3] in (0, 1) and ms > 0
    assert b.shape() == best
    for k in keep:
        k["o"].fill_(0xFF)
    b.run(b.GRAPH)
    torch.cuda.synchronize()
    for i, (a, k) in enumerate(zip(want, keep)):
        assert torch.equal(a, k["o"]), ("tuned", i)
    b2 = vren_amd.Batch(vctx, descs, ol.cfg())
    assert b2.shape() == best
    path = str(tmp_path / "tuned.txt")
    vctx.check(vctx.lib.vren_autotune_save(vctx.h, path.encode()))
    other = vren_amd.Context()
    other.check(other.lib.vren_autotune_load(other.h, path.encode()))
    b3 = vren_amd.Batch(other, descs, ol.cfg())
    assert b3.shape() == best
    b3.close()
    other.close()
    b2.close()
    b.close()
