"""torch.autograd.Function around vren_shade / vren_shade_backward (the differentiable shade).

The light lists (cluster keys, references, counts, offsets, indices) are held fixed: they come from a forward run
of the clustered chain and are not differentiated. The differentiable inputs are
  light_positions  float32 [L, 4]  (xyz, w ignored)
  point_lights     float32 [L, 4]  (rgb = colour * intensity, radius)
  albedo_roughness float32 [W*H, 4]
Background pixels (view depth >= far - 0.001) are not shaded: the forward leaves them at zero and they receive no
gradient. Use the smooth attenuation (ATTEN_SMOOTH) so the loss is differentiable at the light radius.
"""
import torch


class ShadeFrame:
    """The fixed (non-differentiated) part of one frame: G-buffer, camera and the culled light lists."""

    def __init__(self, ctx, W, H, cam, cfg, depth, normals, metallic, refs, key_capacity, counts, offsets, indices,
                 index_capacity):
        self.ctx, self.W, self.H, self.cam, self.cfg = ctx, W, H, cam, cfg
        self.depth, self.normals, self.metallic, self.refs = depth, normals, metallic, refs
        self.key_capacity, self.counts, self.offsets = key_capacity, counts, offsets
        self.indices, self.index_capacity = indices, index_capacity


class ShadeFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, light_positions, point_lights, albedo_roughness, frame):
        f = frame
        lp, pl, ar = light_positions.contiguous(), point_lights.contiguous(), albedo_roughness.contiguous()
        out = torch.zeros((f.W * f.H, 4), dtype=torch.float32, device=lp.device)
        f.ctx.shade(f.W, f.H, f.cam, f.cfg, f.depth, f.normals, ar, f.metallic, f.refs, f.key_capacity, f.counts,
                    f.offsets, f.indices, f.index_capacity, lp, pl, lp.shape[0], out)
        ctx.save_for_backward(lp, pl, ar)
        ctx.frame = f
        return out

    @staticmethod
    def backward(ctx, grad_out):
        lp, pl, ar = ctx.saved_tensors
        f = ctx.frame
        g_pos = torch.zeros_like(lp)
        g_pl = torch.zeros_like(pl)
        g_ar = torch.zeros_like(ar)
        f.ctx.shade_backward(f.W, f.H, f.cam, f.cfg, f.depth, f.normals, ar, f.metallic, f.refs, f.key_capacity,
                             f.counts, f.offsets, f.indices, f.index_capacity, lp, pl, lp.shape[0],
                             grad_out.contiguous(), g_pos, g_pl, g_ar)
        return g_pos, g_pl, g_ar, None


def shade(light_positions, point_lights, albedo_roughness, frame):
    return ShadeFunction.apply(light_positions, point_lights, albedo_roughness, frame)
