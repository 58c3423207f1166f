"""torch.autograd.Function over the differentiable HQS passes.

Parameters, images and gradients stay in device memory: the function hands
torch tensors' device addresses to scr_diff_render / scr_diff_backward.  The
library works on its own HIP stream, so each call synchronises torch's device
work first and the library's stream after.

    render = HQSRender(DifferentiableHQS(ctx, cloud, sizes, cfg), cameras)
    images = render(params)            # (V, H, W, 3) float32, params (poolPoints, 4)
    loss = ((images - targets) ** 2).mean(); loss.backward()

The gradient of all views is accumulated in FP64 by the library (views in
order) and cast to float32 once.

torch bundles a HIP runtime with the same soname as the system one and only
sees the GPU when its own copy serves the process: import torch before the
first Context is created.
"""
from __future__ import annotations

import torch

if not torch.cuda.is_available():
    raise ImportError("torch sees no GPU (import torch before creating the first scr Context, so that torch's "
                      "HIP runtime is the one loaded)")


class DiffHQSFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, renderer, cameras):
        if not params.is_cuda or params.dtype != torch.float32 or params.dim() != 2 or params.shape[1] != 4:
            raise ValueError("params must be a (poolPoints, 4) float32 device tensor")
        if params.shape[0] < renderer.poolPoints:
            raise ValueError(f"params has {params.shape[0]} rows, the pool holds {renderer.poolPoints} points")
        params = params.contiguous()
        w, h = renderer.frames[0].width, renderer.frames[0].height
        if any((f.width, f.height) != (w, h) for f in renderer.frames):
            raise ValueError("all views of one call share a resolution")
        images = torch.empty((len(cameras), h, w, 3), dtype=torch.float32, device=params.device)
        torch.cuda.synchronize(params.device)
        renderer.forward_all(cameras, params.data_ptr(), [images[v].data_ptr() for v in range(len(cameras))])
        renderer.ctx.sync()
        ctx.renderer = renderer
        ctx.cameras = cameras
        ctx.n = params.shape[0]
        return images

    @staticmethod
    def backward(ctx, grad_images):
        g = grad_images.contiguous().to(torch.float32)
        grad = torch.zeros((ctx.n, 4), dtype=torch.float64, device=g.device)
        torch.cuda.synchronize(g.device)
        ctx.renderer.backward_all(ctx.cameras, [g[v].data_ptr() for v in range(g.shape[0])], grad.data_ptr())
        ctx.renderer.ctx.sync()
        return grad.to(torch.float32), None, None


class HQSRender(torch.nn.Module):
    """Module form: images = HQSRender(renderer, cameras)(params)."""

    def __init__(self, renderer, cameras):
        super().__init__()
        self.renderer = renderer
        self.cameras = list(cameras)

    def forward(self, params):
        return DiffHQSFunction.apply(params, self.renderer, self.cameras)
