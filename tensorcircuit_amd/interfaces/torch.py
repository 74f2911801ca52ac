"""torch_interface (interfaces/torch.py:18-110 of the reference).

Every function of this package already takes and returns torch tensors and
differentiates through torch autograd, so the interface is a
torch.autograd.Function whose backward is the backend's vjp of ``fun`` --
the same contract as the reference, which converts between frameworks at
this boundary.
"""
from __future__ import annotations

from typing import Any, Callable

import torch

from ..backend import backend


def torch_interface(fun: Callable[..., Any], jit: bool = False, enable_dlpack: bool = False) -> Callable[..., Any]:
    """Wrap ``fun`` (tensors in, tensors out) as a torch function with a
    vjp-based backward.  ``enable_dlpack`` is accepted for signature parity;
    no conversion is needed here."""
    del enable_dlpack
    f = backend.jit(fun) if jit else fun

    class Fun(torch.autograd.Function):
        @staticmethod
        def forward(ctx: Any, *x: torch.Tensor) -> Any:
            ctx.save_for_backward(*x)
            return f(*[xi.detach() for xi in x])

        @staticmethod
        def backward(ctx: Any, *grad_y: torch.Tensor) -> Any:
            x = ctx.saved_tensors
            g = grad_y[0] if len(grad_y) == 1 else tuple(grad_y)
            _, gx = backend.vjp(f, tuple(xi.detach() for xi in x), g)
            if not isinstance(gx, (tuple, list)):
                gx = (gx,)
            return tuple(gx)

    return Fun.apply


pytorch_interface = torch_interface
