"""Pins of the oracle functions whose earlier checks only covered special cases (CPU only). Every expected value here
is worked out from the reference's formulas by hand or by plain FP64 geometry written for this file, never by
calling the oracle function under test:

* the cluster AABB (clustered_shading.glsl:72-110) against the frustum-slice corners in closed form, in both AABB
  modes and both key encoders, plus the invariant that every pixel's FP64 view-space point lies in its key's AABB;
* off-axis cube-face normal codes (clustered_shading.glsl:7-43), hand-worked, and a round trip through the
  reference's own decoder (clustered_shading.glsl:45-54) over random directions;
* the GGX / Smith / Schlick BRDF (pbr.glsl:6-73) at angles where every cosine differs from 1;
* camera view matrices with yaw and pitch (camera.cpp:9-38) against the camera basis in closed form.

A dropped or altered term in any of these (far plane one slice off, u and v swapped, a Fresnel exponent of 4,
N.V used twice in Smith G, a rotation applied in the wrong order or sign) fails one of the tests below.
"""
import math

import numpy as np
import pytest

import oracle_lib as ol

F32 = np.float32


def _half(x):
    return np.array([x], np.float16).view(np.uint16)[0]


@pytest.fixture(scope="module")
def co(oracle):
    return ol.ClusteredOracle(oracle)


# ---- cluster AABB (clustered_shading.glsl:72-110) -------------------------------------------------------------------

C3_W, C3_H = 1920, 1080


def _proj_scales(cam, W, H):
    """x and y scales of camera.cpp:40-50 in float32: ndc_x = p0 * x / z, ndc_y = p5 * y / z."""
    t = F32(math.tan(float(F32(cam.fov_y) / F32(2))))
    aspect = F32(W) / F32(H)
    return float(F32(1) / (t * aspect)), float(F32(1) / t)


def _slice_planes(cam, W, H, k, cf):
    """view-space z of the near and far planes of depth slice k, FP64 closed forms of the two encoders."""
    near, far = float(F32(cam.near_plane)), float(F32(cam.far_plane))
    if cf.key_mode == 1:
        S = cf.depth_slices
        return near * (far / near) ** (k / S), near * (far / near) ** ((k + 1) / S)
    Sy = (H + 31) // 32
    t = F32(math.tan(float(F32(cam.fov_y) / F32(2))))
    a = float(F32(2) * t / F32(Sy) + F32(1))   # GLSL float: 2 tan(half fov) / Sy + 1
    return near * a ** k, near * a ** (k + 1)


def _tile_ndc(i, j, W, H):
    Sx, Sy = (W + 31) // 32, (H + 31) // 32
    x0, x1 = i / Sx * 2 - 1, (i + 1) / Sx * 2 - 1
    y0, y1 = (1 - j / Sy) * 2 - 1, (1 - (j + 1) / Sy) * 2 - 1   # image rows grow downwards
    return x0, x1, y0, y1


def _want_aabb(cam, W, H, cf, i, j, k):
    p0, p5 = _proj_scales(cam, W, H)
    zn, zf = _slice_planes(cam, W, H, k, cf)
    x0, x1, y0, y1 = _tile_ndc(i, j, W, H)
    if cf.cluster_aabb_mode == 1:
        # tight: min/max over the 8 corners of the frustum cell (tile rectangle x [zn, zf])
        pts = np.array([[x * z / p0, y * z / p5, z] for x in (x0, x1) for y in (y0, y1) for z in (zn, zf)])
        return pts.min(0), pts.max(0)
    # reference: the near-plane point of corner (i, j) and the far-plane point of corner (i + 1, j + 1)
    return (np.array([x0 * zn / p0, y0 * zn / p5, zn]), np.array([x1 * zf / p0, y1 * zf / p5, zf]))


def _key(i, j, k, cf):
    if cf.key_mode == 1:
        return i | (j << 8) | (k << 16) | (5 << 24)
    return i | (j << 8) | (k << 16) | (4 << 26)


def _assert_close(got, want, scale, what):
    err = np.abs(np.asarray(got, np.float64) - want)
    assert np.all(err <= 2e-6 * scale), (what, got, want, err / scale)


def test_cluster_aabb_tile0_slice0_hand_values(co):
    """Tile (0, 0), slice 0 of the 1920x1080 camera (90 deg, near 0.1, far 100): tan = 1, aspect = 16/9,
    Sx = 60, Sy = 34, a = 1 + 2/34. Reference two-corner box: min = (-aspect * 0.1, 0.1, 0.1) (NDC (-1, 1) on the
    near plane), max = NDC (-1 + 2/60, 1 - 2/34) on the plane 0.1 a."""
    cam = ol.camera(C3_W, C3_H)
    a = 1.0 + 2.0 / 34.0
    aspect = 1920.0 / 1080.0
    want_min = np.array([-aspect * 0.1, 0.1, 0.1])
    want_max = np.array([(-1 + 2 / 60) * aspect * 0.1 * a, (1 - 2 / 34) * 0.1 * a, 0.1 * a])
    mn, mx = co.cluster_aabb(C3_W, C3_H, cam, ol.cfg(), _key(0, 0, 0, ol.cfg()))
    _assert_close(mn, want_min, 0.2, "min")
    _assert_close(mx, want_max, 0.2, "max")
    # tight box of the same cell: x from -aspect * far_0 (left edge, far plane) to the right edge on the near plane
    mn, mx = co.cluster_aabb(C3_W, C3_H, cam, ol.cfg(cluster_aabb_mode=1), _key(0, 0, 0, ol.cfg()))
    _assert_close(mn, [-aspect * 0.1 * a, (1 - 2 / 34) * 0.1, 0.1], 0.2, "tight min")
    _assert_close(mx, [(-1 + 2 / 60) * aspect * 0.1, 0.1 * a, 0.1 * a], 0.2, "tight max")


@pytest.mark.parametrize("aabb_mode", [0, 1])
@pytest.mark.parametrize("key_mode", [0, 1])
def test_cluster_aabb_closed_form(co, aabb_mode, key_mode):
    """Every tile of a 1920x1080 and an off-size frame at a spread of slices, both encoders and both AABB modes, against
    the frustum-slice corners in FP64 (slice planes near a^k and near (far/near)^(k/S))."""
    cf = ol.cfg(key_mode=key_mode, depth_slices=32, cluster_aabb_mode=aabb_mode)
    slices = range(0, 32) if key_mode == 1 else [0, 1, 2, 7, 30, 64, 100, 119]
    for (W, H, fov) in [(C3_W, C3_H, 90.0), (1000, 700, 60.0)]:
        cam = ol.camera(W, H, fov_deg=fov, pos=(3.0, -1.0, 2.0))   # the view translation does not enter the AABB
        Sx, Sy = (W + 31) // 32, (H + 31) // 32
        for k in slices:
            for (i, j) in [(0, 0), (Sx - 1, Sy - 1), (Sx // 2, Sy // 3), (7, Sy - 2), (Sx - 3, 1)]:
                mn, mx = co.cluster_aabb(W, H, cam, cf, _key(i, j, k, cf))
                wmn, wmx = _want_aabb(cam, W, H, cf, i, j, k)
                scale = max(np.abs(wmn).max(), np.abs(wmx).max())
                _assert_close(mn, wmn, scale, (W, i, j, k, "min"))
                _assert_close(mx, wmx, scale, (W, i, j, k, "max"))


@pytest.mark.parametrize("key_mode", [0, 1])
def test_every_pixel_lies_in_its_cluster_aabb(co, key_mode):
    """Tight mode: the view-space point of every pixel (unprojected in FP64 from its D32 depth and its tile
    coordinate gid / size) lies inside the AABB of the key the encoder gave it; and the slice of the key brackets
    the pixel's depth in FP64 (near a^k <= z <= near a^(k+1)). The frame is a multiple of 32 pixels on both axes:
    the AABB splits the screen into Sx x Sy equal parts (clustered_shading.glsl:80-84) while keys use 32-pixel
    tiles, so for other sizes the last rows and columns of pixels fall outside their key's AABB, in the reference
    as here."""
    W, H = 416, 256
    cam = ol.camera(W, H)
    cf = ol.cfg(key_mode=key_mode, depth_slices=32, cluster_aabb_mode=1)
    depth, normals = co.synth_frame(1234 + key_mode, W, H, cam)
    keys, U, refs = co.unique_clusters(W, H, cam, cf, depth, normals)
    p0, p5 = _proj_scales(cam, W, H)
    near, far = float(F32(cam.near_plane)), float(F32(cam.far_plane))
    f = float(F32(far) / (F32(far) - F32(near)))
    g = float(-(F32(far) * F32(near)) / (F32(far) - F32(near)))
    gy, gx = np.divmod(np.arange(W * H), W)
    d = depth.astype(np.float64)
    z = g / (d - f)                                   # ndc depth d = f + g / z  (camera.cpp:46-48)
    x = (gx / W * 2 - 1) * z / p0
    y = ((1 - gy / H) * 2 - 1) * z / p5
    fg = (z >= near) & (z <= far)                     # pixels in front of the far plane
    boxes = {}
    for key in np.unique(keys):
        boxes[int(key)] = co.cluster_aabb(W, H, cam, cf, int(key))
    bad = 0
    checked = 0
    for p in np.nonzero(fg)[0][::3]:
        key = int(keys[refs[p]])
        mn, mx = boxes[key]
        pt = np.array([x[p], y[p], z[p]])
        tol = 2e-6 * np.abs(pt).max()
        if np.any(pt < mn - tol) or np.any(pt > mx + tol):
            bad += 1
        k = (key >> 16) & (0xFF if key_mode == 1 else 0x3FF)
        zn, zf = _slice_planes(cam, W, H, k, cf)
        assert zn * (1 - 2e-6) <= z[p] <= zf * (1 + 2e-6), (p, z[p], k, zn, zf)
        checked += 1
    assert checked > 10000 and bad == 0, bad


# ---- off-axis normal codes (clustered_shading.glsl:7-43) ------------------------------------------------------------

# (normal, code) worked by hand: the face is the axis of the smallest t = sign(n_i) / n_i (ties keep the lower
# axis; 0/0 is NaN and never smaller), index (n_i > 0) * 3 + i; p = n * t; u = p[(axis + 1) % 3],
# v = p[(axis + 2) % 3]; cell = floor((u + 1) / 2 * 3), floor((v + 1) / 2 * 3); code = face * 9 + cu * 3 + cv.
HAND_CODES = [
    ((0.6, 0.8, 0.0), 41),      # face y+ (4); p = (0.75, 1, 0): u = p.z = 0 -> 1, v = p.x = 0.75 -> 2
    ((0.8, -0.6, 0.0), 28),     # face x+ (3); p = (1, -0.75, 0): u = p.y -> 0, v = p.z = 0 -> 1
    ((-0.3, 0.2, 0.93), 49),    # face z+ (5); u = -0.3/0.93 -> floor(1.016) = 1, v = 0.2/0.93 -> 1
    ((0.1, -0.9, -0.42), 10),   # face y- (1); u = p.z = -0.467 -> 0, v = p.x = 0.111 -> 1
    ((-0.7, -0.2, 0.68), 5),    # face x- (0); u = p.y = -0.286 -> 1, v = p.z = 0.971 -> floor(2.957) = 2
    ((0.2, 0.5, -0.85), 23),    # face z- (2); u = p.x = 0.235 -> 1, v = p.y = 0.588 -> 2
    ((0.5, -0.5, 0.1), 28),     # tie x/y: face x+ (3); u = p.y = -1 exactly -> 0 (the u = -1 edge), v = 0.2 -> 1
    ((0.5, 0.5, -0.25), 36),    # tie x/y: face x+; u = +1 exactly -> floor(3) = 3 (the reference's overflow), v -> 0
    ((0.0, 0.5, 0.5), 46),      # tie y/z: face y+ (4); u = p.z = 1 -> 3, v = p.x = 0 -> 1
    ((0.5, 0.5, 0.5), 39),      # three-way tie: face x+; u = v = 1 -> 3, 3
    ((-0.5, 0.0, -0.5), 3),     # tie x/z: face x- (0); u = p.y = 0 -> 1, v = p.z = -1 -> 0
    ((0.0, 0.0, -3.0), 22),     # not unit length: face z- (2), centre cell
    ((1e-3, -2e-3, 0.0), 14),   # tiny: face y- (1); p = n * 500 = (0.5, -1, 0): u = p.z = 0 -> 1, v = 0.5 -> 2
]


def _codes_of(co, ns, key_mode=0):
    """The 6-bit normal codes (or 3-bit faces in the sliced encoder) the key encoder gives pixels with normals
    `ns` (one 1-pixel-high frame, depth mid-range)."""
    n = len(ns)
    W, H = n, 1
    cam = ol.camera(W, H)
    depth = np.full(n, 0.999, np.float32)
    normals = np.zeros((n, 4), np.uint16)
    for p, v in enumerate(ns):
        normals[p, :3] = [_half(c) for c in v]
    raw = co.tile_keys(W, H, cam, ol.cfg(key_mode=key_mode), depth, normals)
    Sx = (W + 31) // 32
    keys = np.array([raw[gx // 32, gx % 32] for gx in range(n)], np.uint32)
    return (keys >> 24) if key_mode == 1 else (keys >> 26)


def test_normal_codes_hand_worked(co):
    got = _codes_of(co, [n for n, _ in HAND_CODES])
    assert got.tolist() == [c for _, c in HAND_CODES]


def test_normal_codes_decode_round_trip(co):
    """clustered_shading_decode_normal_idx (clustered_shading.glsl:45-54) maps a code back to (face, u cell, v cell);
    the face's axis is (face % 3), its sign (face / 3), and clustered_shading_get_light_intensity
    (clustered_shading.glsl:127-131) puts the u cell on axis (face + 1) % 3 and the v cell on (face + 2) % 3. Every
    random direction must fall in the cell its code names."""
    rng = np.random.default_rng(11)
    v = rng.normal(size=(4000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    nh = v.astype(np.float16)
    n = nh.astype(np.float64)
    codes = _codes_of(co, [tuple(map(float, r)) for r in nh])
    checked = 0
    for c, d in zip(codes.tolist(), n):
        top = np.sort(np.abs(d))
        if top[2] == top[1]:
            continue   # a tie puts u or v at exactly +-1: the reference's cell 3 overflows (hand-worked above)
        face, cu, cv = c // 9, (c // 3) % 3, c % 3
        axis = face % 3
        m = abs(d[axis])
        assert m >= np.abs(d).max() * (1 - 1e-12), (d, c)          # the dominant axis
        assert (d[axis] > 0) == (face // 3 == 1), (d, c)
        u, w = d[(axis + 1) % 3] / m, d[(axis + 2) % 3] / m
        eps = 1e-6
        assert -1 + 2 * cu / 3 - eps <= u <= -1 + 2 * (cu + 1) / 3 + eps, (d, c, u)
        assert -1 + 2 * cv / 3 - eps <= w <= -1 + 2 * (cv + 1) / 3 + eps, (d, c, w)
        checked += 1
    assert checked > 3900
    # the sliced encoder keeps the face only
    faces = _codes_of(co, [tuple(map(float, r)) for r in nh[:500]], key_mode=1)
    assert np.array_equal(faces, codes[:500] // 9)


# ---- BRDF off the axis (pbr.glsl:6-73, shade.comp:69-141) -----------------------------------------------------------

PI_REF = float(F32(3.14))   # common.glsl:7


def _brdf_by_hand(P, eye, N, light_pos, colour, radius, albedo, rough, metal, atten):
    """One point light, FP64, written out term by term from pbr.glsl in textbook form: GGX NDF with a = roughness,
    Smith G with Schlick-GGX k = roughness on N.V and N.L, Schlick F with exponent 5 on H.V."""
    V = (eye - P) / np.linalg.norm(eye - P)
    d = P - light_pos
    dist = np.linalg.norm(d)
    Ldir = d / dist
    toL = -Ldir
    Hh = (V + toL) / np.linalg.norm(V + toL)
    n_v, n_l, n_h, h_v = max(N @ V, 0.0), max(N @ toL, 0.0), max(N @ Hh, 0.0), max(Hh @ V, 0.0)
    a2 = rough * rough
    ndf = a2 / (PI_REF * (n_h * n_h * (a2 - 1) + 1) ** 2)
    g_v = n_v / (n_v * (1 - rough) + rough)
    g_l = n_l / (n_l * (1 - rough) + rough)
    F0 = float(F32(0.04)) * (1 - metal) + albedo * metal
    F = F0 + (1 - F0) * (1 - h_v) ** 5
    spec = ndf * g_v * g_l / (4 * n_v * n_l + float(F32(0.0001)))
    x = dist / radius
    att = max(0.0, 1 - x ** 4) ** 2 if atten == 1 else 1 - min(x, 1.0) ** 1.5
    lin = ((1 - F) * (1 - metal) * albedo / PI_REF + F * spec) * colour * att * n_l
    return np.maximum(lin, 0.0), dict(n_v=n_v, n_l=n_l, n_h=n_h, h_v=h_v)


def _shade_centre_pixel(co, z_view, normal, lights, albedo, rough, metal, atten):
    """2x2 frame; pixel (1, 1) sits at NDC (0, 0), i.e. on the camera axis at view depth z_view. `lights` holds
    (world position, colour, radius). Returns the oracle's gamma-corrected RGB, the FP64 world point and N."""
    W = H = 2
    cam = ol.camera(W, H)
    near, far = F32(cam.near_plane), F32(cam.far_plane)
    f = F32(far / (far - near))
    g = F32(-(far * near) / (far - near))
    depth = np.ones(W * H, np.float32)
    depth[3] = F32(float(f) + float(g) / z_view)
    z = float(g) / (float(depth[3]) - float(f))            # the view depth the D32 value encodes
    P = np.array([0.0, 0.0, z - 50.0])                      # camera at (0, 0, -50), looking down +z
    normals = np.zeros((W * H, 4), np.uint16)
    normals[3, :3] = [_half(c) for c in normal]
    N = np.array([np.float16(c) for c in normal], np.float64)
    ar = np.zeros((W * H, 4), np.float32)
    ar[3, :3] = albedo
    ar[3, 3] = rough
    me = np.zeros(W * H, np.float32)
    me[3] = metal
    n = len(lights)
    pos = np.zeros((n, 4), np.float32)
    lig = np.zeros((n, 4), np.float32)
    for i, (lp, col, r) in enumerate(lights):
        pos[i, :3] = lp
        lig[i, :3] = col
        lig[i, 3] = r
    out, mask = co.shade(W, H, cam, ol.cfg(attenuation=atten), depth, normals, ar, me, np.zeros(W * H, np.uint32), 1,
                         np.array([n], np.uint32), np.array([0], np.uint32), np.arange(n, dtype=np.uint32), pos, lig)
    assert mask.tolist() == [False, False, False, True]
    return out[3, :3], P, N, pos[:, :3].astype(np.float64), lig.astype(np.float64)


@pytest.mark.parametrize("atten", [0, 1])
@pytest.mark.parametrize("metal,rough", [(0.0, 0.35), (1.0, 0.6), (0.0, 0.9)])
def test_brdf_off_axis_one_light(co, atten, metal, rough):
    """N tilted 70 deg from the view direction, the light 40 deg from N on the other side: N.V = cos 70,
    N.L = cos 40, H halfway between V and L (55 deg from each), so N.H = cos 15 and H.V = cos 55. Every BRDF
    factor is away from its on-axis value (1 - H.V = 0.43, so Schlick's f5 = 0.014 against f4 = 0.033)."""
    alpha, beta = math.radians(70.0), math.radians(40.0)
    normal = (math.sin(alpha), 0.0, -math.cos(alpha))          # V = (0, 0, -1); N rotated 70 deg towards +x
    w = np.array([math.sin(alpha + beta), 0.0, -math.cos(alpha + beta)])   # towards the light, 40 deg past N
    z_view = 20.0
    P_est = np.array([0.0, 0.0, z_view - 50.0])
    lp = (P_est + 3.0 * w).astype(np.float32)
    albedo = np.array([0.25, 0.5, 0.75], np.float32)
    colour = np.array([3.0, 2.0, 1.0], np.float32)
    got, P, N, pos, lig = _shade_centre_pixel(co, z_view, normal, [(lp, colour, 6.0)], albedo, F32(rough), metal,
                                              atten)
    eye = np.array([0.0, 0.0, -50.0])
    lin, cos = _brdf_by_hand(P, eye, N, pos[0], lig[0, :3], lig[0, 3], albedo.astype(np.float64), float(F32(rough)),
                             metal, atten)
    # the geometry is what the docstring says (to the f16 / f32 rounding of N and the light position)
    assert abs(cos["n_v"] - math.cos(alpha)) < 2e-3 and abs(cos["n_l"] - math.cos(beta)) < 2e-3
    assert abs(cos["n_h"] - math.cos(math.radians(15))) < 2e-3 and abs(cos["h_v"] - math.cos(math.radians(55))) < 2e-3
    want = np.power(float(F32(0.03)) * albedo.astype(np.float64) + lin, 1.0 / float(F32(2.2)))
    assert np.allclose(got, want, rtol=1e-10, atol=0), (got, want)


def test_brdf_off_plane_lights_sum(co):
    """Three lights at generic 3-D directions (one below the horizon, which adds nothing but can still reach the
    specular term's max(.., 0)), rough 0.2 and a half-metal: the oracle equals the hand sum."""
    normal = (0.3, -0.4, -0.866)
    z_view = 35.0
    P_est = np.array([0.0, 0.0, z_view - 50.0])
    dirs = [np.array([0.5, 0.3, -0.6]), np.array([-0.2, -0.9, -0.3]), np.array([-0.3, 0.4, 0.9])]
    lights = []
    for k, dvec in enumerate(dirs):
        dvec = dvec / np.linalg.norm(dvec)
        lights.append(((P_est + (2.0 + k) * dvec).astype(np.float32), np.array([1.0 + k, 0.5, 2.0 - k], np.float32),
                       F32(7.0)))
    albedo = np.array([0.6, 0.3, 0.1], np.float32)
    got, P, N, pos, lig = _shade_centre_pixel(co, z_view, normal, lights, albedo, F32(0.2), 0.5, 1)
    eye = np.array([0.0, 0.0, -50.0])
    lin = np.zeros(3)
    for i in range(len(lights)):
        li, _ = _brdf_by_hand(P, eye, N, pos[i], lig[i, :3], lig[i, 3], albedo.astype(np.float64), float(F32(0.2)),
                              0.5, 1)
        lin += li
    want = np.power(float(F32(0.03)) * albedo.astype(np.float64) + lin, 1.0 / float(F32(2.2)))
    assert np.allclose(got, want, rtol=1e-10, atol=0), (got, want)


# ---- camera with yaw and pitch (camera.cpp:9-38) --------------------------------------------------------------------

@pytest.mark.parametrize("yaw,pitch", [(0.3, -0.2), (-1.1, 0.7), (2.5, 0.0), (0.0, -1.2), (3.0, 1.4)])
def test_rotated_camera_basis(co, yaw, pitch):
    """orientation = Ry(yaw) Rx(-pitch) (glm::rotate right-handed), view = orientation^-1 T(-position). In
    closed form the camera's right, up and forward axes are
        right   = (cos y, 0, -sin y)
        up      = (-sin y sin p, cos p, -cos y sin p)
        forward = (sin y cos p, sin p, cos y cos p),
    so the view matrix must send position + a right + b up + c forward to (a, b, c)."""
    pos = (3.0, -2.0, 7.5)
    cam = ol.camera(640, 360, pos=pos, yaw=yaw, pitch=pitch)
    y, p = float(F32(yaw)), float(F32(pitch))
    right = np.array([math.cos(y), 0.0, -math.sin(y)])
    up = np.array([-math.sin(y) * math.sin(p), math.cos(p), -math.cos(y) * math.sin(p)])
    fwd = np.array([math.sin(y) * math.cos(p), math.sin(p), math.cos(y) * math.cos(p)])
    view, _ = co.camera_matrices(cam)
    Vm = view.reshape(4, 4).T.astype(np.float64)   # column-major -> row-major
    rng = np.random.default_rng(1)
    for abc in rng.uniform(-20, 20, size=(20, 3)):
        world = np.array(pos, np.float64) + abc[0] * right + abc[1] * up + abc[2] * fwd
        got = Vm @ np.append(world, 1.0)
        assert np.allclose(got[:3], abc, rtol=0, atol=5e-5 * (1 + np.abs(world).max())), (abc, got)
        assert got[3] == 1.0
