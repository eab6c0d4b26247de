"""GPU parity of the rows beside the clustered-shading chain, through the C ABI against the CPU oracle:

* the material path of shade (shade.comp:89-96): texcoord + material-index G-buffer, a material table and RGBA8
  textures in every filter and address mode, within 1e-5 of the FP64 oracle that resolves the materials the same way,
  as a single call and inside a batch that mixes material-path and resolved-material frames;
* the depth-buffer pyramid (depth_buffer_pyramid.cpp:177-299), bit for bit;
* show_clusters (show_clusters.comp:27-126) in every mode and visualize_bvh (show_bvh.comp), bit for bit.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

SHADE_RTOL, SHADE_ATOL = 1e-5, 1e-6


@pytest.fixture(scope="module")
def co(oracle):
    return ol.ClusteredOracle(oracle)


def _err(got, ref):
    return np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), SHADE_ATOL / SHADE_RTOL)


def make_materials(seed, n_materials=8, n_textures=6):
    """Textures of assorted sizes in every (filter, address mode) pairing and n_materials materials over them. The
    metallic-roughness textures keep green >= 40, so every in-table material has a roughness > 0. More textures
    than 6 repeat the sizes with other (filter, address) pairings."""
    import vren_amd
    rng = np.random.default_rng(seed)
    textures = []
    sizes = [(256, 256), (64, 128), (7, 5), (1, 1), (33, 17), (128, 64)]
    for k in range(n_textures):
        w, h = sizes[k % 6]
        t = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        if k % 2 == 1:
            t[..., 1] = rng.integers(40, 256, (h, w), dtype=np.uint8)
        j = k % 6 + k // 6
        textures.append((t, j % 2 if k % 6 < 4 else 1 - j % 2, j % 3, (j + 1) % 3))
    mats = (vren_amd.Material * n_materials)()
    for m in range(n_materials):
        mats[m].base_color_texture_idx = int(rng.integers(0, n_textures)) // 2 * 2      # an even texture
        mats[m].metallic_roughness_texture_idx = int(rng.integers(0, n_textures // 2)) * 2 + 1  # an odd one
        mats[m].metallic_factor = float(rng.uniform(0.0, 1.0))
        mats[m].roughness_factor = float(rng.uniform(0.3, 1.0))
        for c in range(3):
            mats[m].base_color_factor[c] = float(rng.uniform(0.2, 1.0))
        mats[m].base_color_factor[3] = 1.0
    return textures, mats


class DeviceMaterials:
    """Device copies of a frame's texcoords, material indices, material table and textures, and the MaterialSet
    pointing at them."""

    def __init__(self, vren_amd, texcoords, mat_idx, textures, mats):
        self.tc, self.mi = to_dev(texcoords), to_dev(mat_idx)
        self.tex_dev = [to_dev(t) for t, _, _, _ in textures]
        tab = (vren_amd.Texture * len(textures))()
        for i, ((t, f, au, av), d) in enumerate(zip(textures, self.tex_dev)):
            tab[i].texels, tab[i].width, tab[i].height = d.data_ptr(), t.shape[1], t.shape[0]
            tab[i].filter, tab[i].address_u, tab[i].address_v = f, au, av
        self.tab = to_dev(np.frombuffer(bytes(tab), np.uint8))
        self.mats = to_dev(np.frombuffer(bytes(mats), np.uint8))
        ms = vren_amd.MaterialSet()
        ms.texcoords, ms.material_indices = self.tc.data_ptr(), self.mi.data_ptr()
        ms.materials, ms.material_count = self.mats.data_ptr(), len(mats)
        ms.textures, ms.texture_count = self.tab.data_ptr(), len(textures)
        self.set = ms


def synth_texcoords(seed, W, H, n_materials, out_of_range=0.002):
    rng = np.random.default_rng(seed)
    gy, gx = np.divmod(np.arange(W * H), W)
    tc = np.stack([gx / W * 3.0 - 1.0 + rng.normal(0, 0.05, W * H),
                   gy / H * 2.5 - 0.7 + rng.normal(0, 0.05, W * H)], axis=1).astype(np.float32)
    mi = rng.integers(0, n_materials, W * H).astype(np.uint16)
    mi[rng.uniform(size=W * H) < out_of_range] = n_materials + 3   # past the table: the zero material
    return tc, mi


def _lists_on_gpu(vctx, co, W, H, cam, cf, depth, normals, pos, lig, key_cap=1 << 20, index_cap=1 << 22):
    from test_clustered_gpu import run_assign
    return run_assign(vctx, co, W, H, cam, cf, depth, normals, pos, lig, key_cap, index_cap)


@pytest.mark.parametrize("n_materials,n_textures", [(8, 6), (100, 70)])
def test_material_resolve_through_ambient_only_shade(vctx, co, n_materials, n_textures):
    """With no lights the radiance is the ambient term alone, (0.03 albedo)^(1/2.2) (shade.comp:133-137), so every
    shaded pixel's resolved albedo can be read back from the output: it must match the FP64 resolve of its
    texcoord and material index (texel selection, filtering, factors) to 1e-5. The larger tables run past the
    entries the kernel stages in LDS, so both the staged and the global lookups are checked."""
    import vren_amd
    W, H = 1024, 512
    cam = ol.camera(W, H)
    depth, normals = co.synth_frame(42 + 2000, W, H, cam)
    textures, mats = make_materials(11, n_materials, n_textures)
    tc, mi = synth_texcoords(12, W, H, len(mats))
    dm = DeviceMaterials(vren_amd, tc, mi, textures, mats)
    refs = np.zeros(W * H, np.uint32)
    out = empty_dev(16 * W * H, fill=0xFF)
    zero = to_dev(np.zeros(4, np.uint32))
    vctx.shade_materials(W, H, cam, ol.cfg(), to_dev(depth), to_dev(normals), dm.set, to_dev(refs), 1, zero, zero,
                         zero, 1, None, None, 0, out)
    got = from_dev(out, np.float32, (W * H, 4))
    ar, me = co.resolve_materials(tc, mi, mats, textures)
    _, mask = co.shade(W, H, cam, ol.cfg(), depth, normals, np.zeros((W * H, 4), np.float32),
                       np.zeros(W * H, np.float32), refs, 1, np.zeros(1, np.uint32), np.zeros(1, np.uint32),
                       np.zeros(1, np.uint32), np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32))
    assert np.all(np.isnan(got[~mask])) and mask.sum() > W * H // 4
    alb = np.power(got[mask, :3].astype(np.float64), 2.2) / 0.03
    want = ar[mask, :3]
    err = np.abs(alb - want) / np.maximum(want, 1e-2)
    worst = np.argsort(err.max(axis=1))[::-1][:5]
    pix = np.nonzero(mask)[0]
    for w in worst:
        p = pix[w]
        m = int(mi[p])
        tx = (mats[m].base_color_texture_idx, mats[m].metallic_roughness_texture_idx) if m < len(mats) else None
        print("pixel %d material %d textures %s texcoord %s: gpu albedo %s want %s"
              % (p, m, tx, tc[p].tolist(), alb[w].tolist(), want[w].tolist()))
    print("ambient-only albedo: %d pixels, rel err max %.3e, %d above 1e-5"
          % (err.shape[0], err.max(), int((err.max(axis=1) > 1e-5).sum())))
    assert err.max() <= 1e-5, err.max()


def test_shade_material_path_1080p_vs_fp64(vctx, co):
    """1920x1080 with 4096 lights, the reference encoder, both attenuations: the material path against the FP64
    oracle resolving the same materials (texel selection of the shared sampler rules, weights and c / 255 in
    double). Pixels of the zero material (an index past the table) can be 0/0 in the reference's Smith G at
    N.V = 0: where the oracle is NaN there is nothing to compare (fewer than 1 % of the pixels)."""
    import vren_amd
    W, H, L = 1920, 1080, 4096
    cam = ol.camera(W, H)
    pos, lig = co.synth_lights(42 + 1000, L)
    depth, normals = co.synth_frame(42 + 2000, W, H, cam)
    textures, mats = make_materials(11)
    tc, mi = synth_texcoords(12, W, H, len(mats))
    dm = DeviceMaterials(vren_amd, tc, mi, textures, mats)
    g = _lists_on_gpu(vctx, co, W, H, cam, ol.cfg(), depth, normals, pos, lig)
    key_cap = 1 << 20
    cc = np.zeros(key_cap, np.uint32)
    oo = np.zeros(key_cap, np.uint32)
    cc[:g["counts"].size], oo[:g["offsets"].size] = g["counts"], g["offsets"]
    for atten in (0, 1):
        cf = ol.cfg(attenuation=atten)
        out = empty_dev(16 * W * H, fill=0xFF)
        vctx.shade_materials(W, H, cam, cf, to_dev(depth), to_dev(normals), dm.set, g["dev"]["refs"], key_cap,
                             g["dev"]["counts"], g["dev"]["offsets"], g["dev"]["idx"], 1 << 22, g["dev"]["pos"],
                             g["dev"]["lig"], L, out)
        got = from_dev(out, np.float32, (W * H, 4))
        ref, mask = co.shade_materials(W, H, cam, cf, depth, normals, tc, mi, mats, textures, g["refs"], key_cap, cc,
                                       oo, g["idx"], pos, lig)
        assert np.all(np.isnan(got[~mask]))
        r = ref[mask, :3]
        fin = ~np.isnan(r).any(axis=1)
        assert fin.mean() > 0.99
        err = _err(got[mask, :3][fin], r[fin])
        print("\nshade material path atten=%d: %d outputs, rel err max %.3e p99.9 %.3e"
              % (atten, err.size, err.max(), np.quantile(err, 0.999)))
        assert err.max() <= SHADE_RTOL, err.max()


def test_batch_mixes_material_and_resolved_frames(vctx, co):
    """A batch of three frames, the first and last on the material path: every launch mode gives the same bytes,
    and each frame's radiance matches the FP64 oracle of its own path."""
    import vren_amd
    from test_clustered_gpu import _batch_all_modes_identical, _batch_host_frame, make_frames_batch
    specs = [(1280, 720, 3000, 8200), (640, 360, 500, 8201), (960, 540, 1500, 8202)]
    descs, keep = make_frames_batch(vctx, co, specs)
    textures, mats = make_materials(21)
    dms = {}
    for i in (0, 2):
        W, H = specs[i][0], specs[i][1]
        tc, mi = synth_texcoords(30 + i, W, H, len(mats))
        dms[i] = (DeviceMaterials(vren_amd, tc, mi, textures, mats), tc, mi)
        descs[i].materials = ctypes.pointer(dms[i][0].set)
        descs[i].albedo_roughness, descs[i].metallic = None, None
    cf = ol.cfg()
    b = vren_amd.Batch(vctx, descs, cf)
    _batch_all_modes_identical(b, keep)
    vctx.sync_status()
    for i in range(3):
        f = _batch_host_frame(b, keep, i)
        W, H = f["W"], f["H"]
        okeys, oU, orefs = co.unique_clusters(W, H, f["cam"], cf, f["depth"], f["normals"])
        assert f["U"] == oU and np.array_equal(f["refs"], orefs)
        ovp, ocodes, obvh = co.light_bvh(f["pos"], f["lig"], f["cam"], cf)
        oc, oo, oi, ot = co.assign(W, H, f["cam"], cf, okeys, oU, obvh, ocodes, ovp)
        assert np.array_equal(f["idx"], oi)
        cc = np.zeros(f["key_capacity"], np.uint32)
        cco = np.zeros(f["key_capacity"], np.uint32)
        cc[:oU], cco[:oU] = oc, oo
        if i in dms:
            _, tc, mi = dms[i]
            ref, mask = co.shade_materials(W, H, f["cam"], cf, f["depth"], f["normals"], tc, mi, mats, textures,
                                           orefs, f["key_capacity"], cc, cco, oi, f["pos"], f["lig"])
        else:
            ref, mask = co.shade(W, H, f["cam"], cf, f["depth"], f["normals"], f["ar"], f["me"], orefs,
                                 f["key_capacity"], cc, cco, oi, f["pos"], f["lig"])
        got = f["out"]
        assert np.all(np.isnan(got[~mask]))
        r = ref[mask, :3]
        fin = ~np.isnan(r).any(axis=1)
        err = _err(got[mask, :3][fin], r[fin])
        assert fin.mean() > 0.99 and err.max() <= SHADE_RTOL, (i, err.max())
    b.close()


def test_batch_material_path_long_lists(vctx, co):
    """The material path when a tile's light-list range is longer than the shade's 256-entry stage: 640x360 with 3000
    lights at five times their radius (most tiles stage their range in windows). Every shaded pixel against the FP64
    oracle of the material path."""
    import torch
    import vren_amd
    from test_clustered_gpu import _batch_host_frame, make_frames_batch
    specs = [(640, 360, 3000, 5100)]
    descs, keep = make_frames_batch(vctx, co, specs)
    keep[0]["l"].view(torch.float32).view(3000, 4)[:, 3] *= 5.0
    textures, mats = make_materials(23)
    W, H = 640, 360
    tc, mi = synth_texcoords(33, W, H, len(mats))
    dm = DeviceMaterials(vren_amd, tc, mi, textures, mats)
    descs[0].materials = ctypes.pointer(dm.set)
    descs[0].albedo_roughness, descs[0].metallic = None, None
    torch.cuda.synchronize()
    cf = ol.cfg()
    b = vren_amd.Batch(vctx, descs, cf)
    for mode in (b.PER_FRAME, b.GRAPH):
        keep[0]["o"].fill_(0xFF)
        b.run(mode)
        torch.cuda.synchronize()
        vctx.sync_status()
        f = _batch_host_frame(b, keep, 0)
        okeys, oU, orefs = co.unique_clusters(W, H, f["cam"], cf, f["depth"], f["normals"])
        ovp, ocodes, obvh = co.light_bvh(f["pos"], f["lig"], f["cam"], cf)
        oc, oo, oi, ot = co.assign(W, H, f["cam"], cf, okeys, oU, obvh, ocodes, ovp)
        assert f["U"] == oU and np.array_equal(f["idx"], oi)
        tile_sums = np.bincount(np.unique(okeys & 0xFFFF, return_inverse=True)[1], weights=oc.astype(np.int64))
        assert (tile_sums > 256).sum() > 0
        cc = np.zeros(f["key_capacity"], np.uint32)
        cco = np.zeros(f["key_capacity"], np.uint32)
        cc[:oU], cco[:oU] = oc, oo
        ref, mask = co.shade_materials(W, H, f["cam"], cf, f["depth"], f["normals"], tc, mi, mats, textures, orefs,
                                       f["key_capacity"], cc, cco, oi, f["pos"], f["lig"])
        got = f["out"]
        assert np.all(np.isnan(got[~mask]))
        r = ref[mask, :3]
        fin = ~np.isnan(r).any(axis=1)
        err = _err(got[mask, :3][fin], r[fin])
        assert fin.mean() > 0.99 and err.max() <= SHADE_RTOL, (mode, err.max())
    b.close()


# ---- depth pyramid -------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("W,H,images", [(1, 1, 1), (2, 1, 3), (5, 100, 2), (64, 64, 1), (333, 200, 4),
                                        (1920, 1080, 1), (1920, 1080, 3), (4096, 3, 1), (640, 360, 64),
                                        (3840, 2160, 1), (65535, 2, 1)])
def test_depth_pyramid_bit_exact(vctx, co, W, H, images):
    """Level 0 is the depth copy, every level the GLSL max over its in-bounds 2x2 texels; odd sizes drop the last
    row / column; NaN, negative and -0 depths in the input. `images` stacked depth buffers in one call."""
    rng = np.random.default_rng(W + 7 * H + images)
    d = rng.uniform(0.0, 1.0, images * W * H).astype(np.float32)
    k = max(1, d.size // 100)
    d[rng.integers(0, d.size, k)] = np.nan
    d[rng.integers(0, d.size, k)] = -0.25
    d[rng.integers(0, d.size, k)] = -0.0
    n = vctx.lib.vren_depth_pyramid_length(W, H)
    out = empty_dev(4 * n * images, fill=0xCD)
    vctx.depth_pyramid(to_dev(d), W, H, images, out)
    got = from_dev(out, np.uint32, (images, n))
    for i in range(images):
        want, levels = co.depth_pyramid(d[i * W * H:(i + 1) * W * H], W, H)
        assert want.size == n and levels == vctx.lib.vren_depth_pyramid_level_count(W, H)
        assert np.array_equal(got[i], want.view(np.uint32)), i


# ---- debug consumers -----------------------------------------------------------------------------------------------

def test_show_clusters_every_mode_1080p(vctx, co):
    from test_clustered_gpu import run_assign
    W, H, L = 1920, 1080, 4096
    cam = ol.camera(W, H)
    pos, lig = co.synth_lights(42 + 1000, L)
    depth, normals = co.synth_frame(42 + 2000, W, H, cam)
    g = run_assign(vctx, co, W, H, cam, ol.cfg(), depth, normals, pos, lig)
    dkeys = to_dev(np.concatenate([g["keys"], np.zeros(1, np.uint32)]))
    for mode in range(7):
        out = empty_dev(16 * W * H)
        vctx.show_clusters(W, H, mode, g["dev"]["refs"], dkeys, g["U"], g["dev"]["counts"], g["dev"]["offsets"],
                           g["dev"]["idx"], g["total"], out)
        got = from_dev(out, np.uint32, (W * H, 4))
        want = co.show_clusters(W, H, mode, g["refs"], g["keys"], g["counts"], g["offsets"], g["idx"])
        assert np.array_equal(got, want.view(np.uint32)), mode


@pytest.mark.parametrize("L", [1, 700, 4096, 40000])
def test_visualize_bvh_bit_exact(vctx, co, L):
    cam = ol.camera(640, 360)
    pos, lig = co.synth_lights(900 + L, L)
    _, _, bvh = co.light_bvh(pos, lig, cam, ol.cfg())
    levels = co.o.calc("level_count", L)
    n = sum(32 ** l for l in range(levels + 1))
    assert n == bvh.size
    out = empty_dev(16 * 24 * n, fill=0xAB)
    vctx.visualize_bvh(to_dev(bvh), levels, out)
    got = from_dev(out, np.uint32, (24 * n, 4))
    assert np.array_equal(got, co.visualize_bvh(bvh, levels))
