"""Differentiable shade: the hand-written adjoint (vren_shade_backward) through the torch.autograd.Function harness,
checked against FP64 central differences of the oracle's loss (h = 1e-4, 256 sampled parameters) and against the
oracle's FP64 analytic adjoint on every parameter.

The workload: a 1920x1080 G-buffer (albedo U[0,1]^3, roughness U[0.05,1], metallic Bernoulli(0.3), depth/normals
of the procedural scene), 4096 lights; smooth windowed attenuation; L2 loss against a target rendered from lights
perturbed by N(0, 0.5) in position and 10% in intensity, with the culled light lists held fixed.
"""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

W_GRAD, H_GRAD, L_GRAD = 1920, 1080, 4096
SEED_GRAD = 42 + 1000 * 4
H_FD = 1e-4
N_PARAMS = 256
# tolerances (stated per check below)
FD_RTOL = 2e-3          # |g_gpu - g_fd| <= FD_RTOL * max(|g_fd|, floor)
ADJ_REL_L2 = 1e-3       # ||g_gpu - g_fp64|| / ||g_fp64|| per parameter group


@pytest.fixture(scope="module")
def grad_frame(oracle, vctx):
    import torch
    import vren_amd
    from vren_amd.autograd import ShadeFrame

    co = ol.ClusteredOracle(oracle)
    cam = ol.camera(W_GRAD, H_GRAD)
    pos, lig = co.synth_lights(SEED_GRAD, L_GRAD)
    depth, normals = co.synth_frame(SEED_GRAD, W_GRAD, H_GRAD, cam)
    ar, me = co.synth_materials(SEED_GRAD, W_GRAD * H_GRAD)
    cf_lists = ol.cfg()
    cf = ol.cfg(attenuation=1)
    # light lists from the GPU chain (held fixed from here on)
    key_cap, index_cap = 1 << 20, 1 << 22
    dpos, dlig = to_dev(pos), to_dev(lig)
    vp, codes = empty_dev(16 * L_GRAD), empty_dev(8 * L_GRAD)
    bvh = empty_dev(32 * oracle.calc("buffer_length", L_GRAD))
    vctx.construct_point_light_bvh(dpos, dlig, L_GRAD, cam, cf_lists, vp, codes, bvh)
    keys, cnt, refs = empty_dev(4 * key_cap), empty_dev(4), empty_dev(4 * W_GRAD * H_GRAD)
    ddepth, dnorm = to_dev(depth), to_dev(normals)
    vctx.find_unique_cluster_list(W_GRAD, H_GRAD, cam, cf_lists, ddepth, dnorm, keys, key_cap, cnt, refs)
    counts, offsets, idx, tot = empty_dev(4 * key_cap), empty_dev(4 * key_cap), empty_dev(4 * index_cap), empty_dev(4)
    vctx.assign_lights(W_GRAD, H_GRAD, cam, cf_lists, keys, cnt, key_cap, bvh, L_GRAD, codes, vp, counts, offsets, idx, index_cap,
                       tot)
    total = int(from_dev(tot, np.uint32)[0])
    frame = ShadeFrame(vctx, W_GRAD, H_GRAD, cam, cf, ddepth, dnorm, to_dev(me), refs, key_cap, counts, offsets, idx,
                       index_cap)
    host = dict(U=int(from_dev(cnt, np.uint32)[0]), refs=from_dev(refs, np.uint32, (W_GRAD * H_GRAD,)),
                counts=from_dev(counts, np.uint32, (key_cap,)),
                offsets=from_dev(offsets, np.uint32, (key_cap,)), idx=from_dev(idx, np.uint32, (total,)))
    # target: lights perturbed by N(0, 0.5) in position and x(1 + 0.1 N(0, 1)) in intensity
    rng = np.random.default_rng(2024)
    pos_t = pos.copy()
    pos_t[:, :3] += rng.normal(0.0, 0.5, (L_GRAD, 3)).astype(np.float32)
    lig_t = lig.copy()
    lig_t[:, :3] *= (1.0 + 0.1 * rng.normal(0.0, 1.0, (L_GRAD, 1))).astype(np.float32)
    target = torch.zeros((W_GRAD * H_GRAD, 4), dtype=torch.float32, device="cuda")
    vctx.shade(W_GRAD, H_GRAD, cam, cf, ddepth, dnorm, to_dev(ar), frame.metallic, refs, key_cap, counts,
               offsets, idx, index_cap, to_dev(pos_t), to_dev(lig_t), L_GRAD, target)
    torch.cuda.synchronize()
    return dict(co=co, cam=cam, cf=cf, pos=pos, lig=lig, depth=depth, normals=normals, ar=ar, me=me, frame=frame,
                host=host, target=target, key_cap=key_cap, vren_amd=vren_amd)


def gpu_grads(grad_frame):
    import torch
    from vren_amd.autograd import shade
    L = L_GRAD
    lp = torch.from_numpy(grad_frame["pos"]).cuda().requires_grad_(True)
    colour = torch.from_numpy(grad_frame["lig"][:, :3].copy()).cuda()
    intensity = colour.max(dim=1, keepdim=True).values.clamp_min(1e-6)
    colour = (colour / intensity).detach().requires_grad_(True)
    intensity = intensity.detach().requires_grad_(True)
    radius = torch.from_numpy(grad_frame["lig"][:, 3:].copy()).cuda().requires_grad_(True)
    lights = torch.cat([colour * intensity, radius], dim=1)
    lights.retain_grad()
    ar = torch.from_numpy(grad_frame["ar"]).cuda().requires_grad_(True)
    out = shade(lp, lights, ar, grad_frame["frame"])
    shaded = (out[:, 3] == 1.0)
    diff = (out[:, :3] - grad_frame["target"][:, :3]) * shaded[:, None]
    loss = (diff * diff).sum()
    loss.backward()
    torch.cuda.synchronize()
    return dict(loss=float(loss), out=out.detach().cpu().numpy(), pos=lp.grad.cpu().numpy(),
                lights=lights.grad.cpu().numpy(), colour=colour.grad.cpu().numpy(),
                intensity=intensity.grad.cpu().numpy(), ar=ar.grad.cpu().numpy(),
                shaded=shaded.cpu().numpy(), lights_value=lights.detach().cpu().numpy())


def test_colour_intensity_chain_rule(vctx, grad_frame):
    # d/dcolour and d/dintensity follow from d/d(rgb = colour * intensity) by the chain rule torch applies
    g = gpu_grads(grad_frame)
    lv = g["lights_value"]
    intensity = np.maximum(lv[:, :3].max(axis=1, keepdims=True), 1e-6)
    colour = lv[:, :3] / intensity
    np.testing.assert_allclose(g["colour"], g["lights"][:, :3] * intensity, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g["intensity"][:, 0], (g["lights"][:, :3] * colour).sum(axis=1), rtol=1e-4, atol=1e-5)


def test_adjoint_vs_fp64_adjoint(vctx, grad_frame):
    co, host = grad_frame["co"], grad_frame["host"]
    g = gpu_grads(grad_frame)
    # FP64 forward and adjoint of the same loss (dLoss/dout from the FP64 forward)
    ref, mask = co.shade(W_GRAD, H_GRAD, grad_frame["cam"], grad_frame["cf"], grad_frame["depth"], grad_frame["normals"], grad_frame["ar"], grad_frame["me"], host["refs"],
                         grad_frame["key_cap"], host["counts"], host["offsets"], host["idx"], grad_frame["pos"], grad_frame["lig"])
    target = grad_frame["target"].cpu().numpy()
    gout = np.zeros((W_GRAD * H_GRAD, 4), np.float32)
    gout[mask, :3] = (2.0 * (ref[mask, :3] - target[mask, :3])).astype(np.float32)
    gp, gl, gar = co.shade_backward(W_GRAD, H_GRAD, grad_frame["cam"], grad_frame["cf"], grad_frame["depth"], grad_frame["normals"], grad_frame["ar"], grad_frame["me"],
                                    host["refs"], grad_frame["key_cap"], host["counts"], host["offsets"], host["idx"],
                                    grad_frame["pos"].astype(np.float64), grad_frame["lig"].astype(np.float64), gout)
    report = {}
    for name, a, b in (("position", g["pos"][:, :3], gp[:, :3]), ("rgb", g["lights"][:, :3], gl[:, :3]),
                       ("radius", g["lights"][:, 3], gl[:, 3]), ("albedo", g["ar"][:, :3], gar[:, :3]),
                       ("roughness", g["ar"][:, 3], gar[:, 3])):
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        report[name] = rel
    print("\nadjoint vs FP64 adjoint, relative L2 error per group:",
          {k: "%.2e" % v for k, v in report.items()})
    for k, v in report.items():
        assert v < ADJ_REL_L2, (k, v)


def fp64_residual_gradient(grad_frame):
    """dLoss/dout of the FP64 forward (the function the central differences differentiate)."""
    co, host = grad_frame["co"], grad_frame["host"]
    ref, mask = co.shade(W_GRAD, H_GRAD, grad_frame["cam"], grad_frame["cf"], grad_frame["depth"], grad_frame["normals"], grad_frame["ar"], grad_frame["me"], host["refs"],
                         grad_frame["key_cap"], host["counts"], host["offsets"], host["idx"], grad_frame["pos"], grad_frame["lig"])
    target = grad_frame["target"].cpu().numpy()
    gout = np.zeros((W_GRAD * H_GRAD, 4), np.float32)
    gout[mask, :3] = (2.0 * (ref[mask, :3] - target[mask, :3])).astype(np.float32)
    return gout, mask


def gpu_backward(grad_frame, gout):
    """vren_shade_backward for a given dLoss/dout (isolates the adjoint from the forward's rounding)."""
    import torch
    f = grad_frame["frame"]
    lp = torch.from_numpy(grad_frame["pos"]).cuda()
    pl = torch.from_numpy(grad_frame["lig"]).cuda()
    ar = torch.from_numpy(grad_frame["ar"]).cuda()
    g_pos, g_pl, g_ar = torch.zeros_like(lp), torch.zeros_like(pl), torch.zeros_like(ar)
    f.ctx.shade_backward(f.W, f.H, f.cam, f.cfg, f.depth, f.normals, ar, f.metallic, f.refs, f.key_capacity, f.counts,
                         f.offsets, f.indices, f.index_capacity, lp, pl, L_GRAD, torch.from_numpy(gout).cuda(), g_pos, g_pl,
                         g_ar)
    torch.cuda.synchronize()
    return g_pos.cpu().numpy(), g_pl.cpu().numpy(), g_ar.cpu().numpy()


def test_adjoint_vs_central_differences(vctx, grad_frame):
    co, host = grad_frame["co"], grad_frame["host"]
    gout, mask = fp64_residual_gradient(grad_frame)
    gpos, glig, gar = gpu_backward(grad_frame, gout)
    target = grad_frame["target"].cpu().numpy()
    refs, counts, offsets, idx = host["refs"], host["counts"], host["offsets"], host["idx"]
    U = host["U"]
    cluster_of_entry = np.repeat(np.arange(U, dtype=np.int64), counts[:U])
    assert cluster_of_entry.size == idx.size
    shaded_px = np.nonzero(mask)[0]
    pos_d = grad_frame["pos"].astype(np.float64)
    lig_d = grad_frame["lig"].astype(np.float64)
    rng = np.random.default_rng(77)
    used_lights = np.unique(idx)
    lights_pick = rng.choice(used_lights, size=N_PARAMS // 2, replace=False)
    px_with_lists = shaded_px[counts[refs[shaded_px]] > 0]
    pixels_pick = rng.choice(px_with_lists, size=N_PARAMS - N_PARAMS // 2, replace=False)
    args = (W_GRAD, H_GRAD, grad_frame["cam"], grad_frame["cf"], grad_frame["depth"], grad_frame["normals"], grad_frame["ar"], grad_frame["me"], refs, grad_frame["key_cap"],
            counts, offsets, idx)
    errs = []

    def central(evaluate):
        # D(h) and D(h/2): the reported derivative is D(h) (h = 1e-4); |D(h) - D(h/2)| ~ 3/4 of D(h)'s truncation
        # error, which bounds how far the exact derivative can be from D(h) (large for dark pixels, where the gamma
        # curve 1/2.2 has a large second derivative)
        d = []
        for hh in (H_FD, H_FD / 2):
            d.append((evaluate(hh) - evaluate(-hh)) / (2 * hh))
        return d[0], abs(d[0] - d[1])

    for n, j in enumerate(lights_pick):
        comp = n % 7  # x, y, z, r, g, b, radius
        clusters = np.unique(cluster_of_entry[idx == j])
        affected = shaded_px[np.isin(refs[shaded_px], clusters)].astype(np.uint32)

        def ev(delta, j=j, comp=comp, affected=affected):
            pd, ld = pos_d.copy(), lig_d.copy()
            if comp < 3:
                pd[j, comp] += delta
            else:
                ld[j, comp - 3] += delta
            return co.shade_loss(*args, pd, ld, target, affected)
        fd, trunc = central(ev)
        ga = gpos[j, comp] if comp < 3 else glig[j, comp - 3]
        errs.append(("light", int(j), comp, float(ga), fd, trunc))
    for n, p in enumerate(pixels_pick):
        comp = n % 4  # albedo r, g, b, roughness

        def ev(delta, p=p, comp=comp):
            par = grad_frame["ar"][p].astype(np.float64).copy()
            par[comp] += delta
            return co.shade_loss(*args, pos_d, lig_d, target, np.array([p], np.uint32), par[None, :])
        fd, trunc = central(ev)
        errs.append(("pixel", int(p), comp, float(gar[p, comp]), fd, trunc))
    # |g_adjoint - D(h)| <= FD_RTOL |D(h)| + 2 |D(h) - D(h/2)| + 1e-12
    ratio = [abs(e[3] - e[4]) / (FD_RTOL * abs(e[4]) + 2.0 * e[5] + 1e-12) for e in errs]
    rel = [abs(e[3] - e[4]) / max(abs(e[4]), 1e-300) for e in errs]
    order = np.argsort(ratio)[::-1]
    print("\ncentral differences (h=%g, %d params): median relative error %.2e, worst tolerance ratio %.3f; "
          "worst entries (kind, id, component, adjoint, D(h), |D(h)-D(h/2)|):"
          % (H_FD, len(errs), float(np.median(rel)), ratio[order[0]]))
    for i in order[:4]:
        print("   ", errs[i], "rel %.2e" % rel[i])
    assert ratio[order[0]] <= 1.0, errs[order[0]]
