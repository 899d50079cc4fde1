"""torch.autograd.Function over the renderer's forward/backward C ABI.

Test-harness only: it lets the parity tests drive the hand-written backward
kernels through torch's autograd (loss.backward()).
"""
import numpy as np
import torch

from gsr import lib


class SplatRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, gsr_ctx, uniform, tile):
        n = params.shape[1]
        W, H = int(uniform[40]), int(uniform[41])
        out = torch.empty((3, H, W), dtype=torch.float32, device=params.device)
        torch.cuda.synchronize()
        gsr_ctx.forward(params.detach().contiguous(), n, uniform, lib.options(tile_size=tile, save_state=True), out)
        gsr_ctx.synchronize()
        ctx.gsr_ctx = gsr_ctx
        ctx.uniform = uniform
        ctx.save_for_backward(params)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (params,) = ctx.saved_tensors
        n = params.shape[1]
        grad = torch.empty_like(params)
        g = grad_out.detach().contiguous().float()
        torch.cuda.synchronize()
        ctx.gsr_ctx.backward(params.detach().contiguous(), n, ctx.uniform, g, grad)
        ctx.gsr_ctx.synchronize()
        return grad, None, None, None


def render(params, gsr_ctx, uniform, tile=16):
    return SplatRender.apply(params, gsr_ctx, np.asarray(uniform, np.float32), tile)
