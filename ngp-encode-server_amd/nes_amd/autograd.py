"""Thin torch.autograd.Function over the trainer's raw forward/backward.

The GPU work happens inside libnes_render.so (nes_trainer_forward /
nes_trainer_backward); tensors cross the boundary as CPU arrays.  For tests:
it lets torch build losses on rendered colours and checks that gradients
reach the hash table and MLP weights.
"""
import numpy as np
import torch


class RenderRays(torch.autograd.Function):
    """rgb = render(table, weights; rays) for explicit rays."""

    @staticmethod
    def forward(ctx, table, weights, trainer, ray_o, ray_d):
        t = table.detach().cpu().numpy()
        w = weights.detach().cpu().numpy()
        o = np.ascontiguousarray(ray_o, np.float32)
        d = np.ascontiguousarray(ray_d, np.float32)
        trainer.set_params(t, w)
        rgb = trainer.forward(o, d)
        # everything the backward needs travels with this node: the trainer
        # is shared by every call, so no per-call state may live in it
        ctx.trainer, ctx.params, ctx.rays = trainer, (t, w), (o, d)
        return torch.from_numpy(rgb)

    @staticmethod
    def backward(ctx, grad_rgb):
        ctx.trainer.set_params(*ctx.params)
        gt, gw = ctx.trainer.backward(*ctx.rays,
                                      grad_rgb.detach().cpu().numpy().astype(np.float32))
        return torch.from_numpy(gt), torch.from_numpy(gw), None, None, None


def render_rays(table, weights, trainer, ray_o, ray_d):
    return RenderRays.apply(table, weights, trainer, ray_o, ray_d)
