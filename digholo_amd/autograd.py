"""Check-only torch binding of the alignment-loss gradient.

``AlignmentLoss.apply(params, lib, h, hold)`` sets a [pol, 6] parameter tensor
(AutoAlign order: waist, defocus, tilt x, centre x, tilt y, centre y; API
units) on a configured handle, processes its batch on the GPU and returns the
loss -(1/B) sum |c|^2; ``backward`` hands back the gradient the library's
adjoint kernels computed (``digHoloAutoAlignGradient``).  It exists so the
adjoint can be checked with torch's own tooling (finite differences,
``torch.autograd.gradcheck``-style comparisons); the library's alignment search
does not use torch.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api


class AlignmentLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params: torch.Tensor, lib, h: int, hold: bool = False) -> torch.Tensor:
        P = params.shape[0]
        api.set_align_params(lib, h, params.detach().cpu().numpy())
        loss, grad = api.alignment_gradient(lib, h, hold=hold)
        ctx.save_for_backward(torch.as_tensor(grad[:P], dtype=params.dtype))
        return torch.as_tensor(loss, dtype=params.dtype)

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None


def alignment_loss(params: torch.Tensor, lib, h: int, hold: bool = False) -> torch.Tensor:
    return AlignmentLoss.apply(params, lib, h, hold)
