"""Parameter-shift gradients evaluated as one batched run of the circuit
(experimental.py:215-269 of the reference)."""
from __future__ import annotations

from typing import Any, Callable, Sequence, Tuple, Union

import numpy as np
import torch

from .backend import backend

Tensor = Any


def parameter_shift_grad(f: Callable[..., Tensor], argnums: Union[int, Sequence[int]] = 0, jit: bool = False,
                         shifts: Tuple[float, float] = (np.pi / 2, 2)) -> Callable[..., Tensor]:
    """Gradient of ``f`` by the shift rule: for every entry k of each
    differentiated argument, ``(f(x + s0 e_k) - f(x - s0 e_k)) / s1``.

    All 2 x size shifted copies of an argument go through ``backend.vmap``,
    so the circuit runs once as a batch of 2 x size parameter sets on the
    device.  Exact for gates exp(-i a/2 P) with the default shifts.
    """
    if jit:
        f = backend.jit(f)
    nums = [argnums] if isinstance(argnums, int) else list(argnums)
    vfs = {i: backend.vmap(f, vectorized_argnums=i) for i in nums}

    def grad_f(*args: Any, **kws: Any) -> Any:
        out = []
        for i in nums:
            x = args[i]
            x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
            shape = tuple(x.shape)
            size = int(x.numel())
            onehot = shifts[0] * torch.eye(size, dtype=x.dtype, device=x.device).reshape((size,) + shape)
            base = x.reshape((1,) + shape).expand((size,) + shape)
            plus = list(args)
            plus[i] = base + onehot
            minus = list(args)
            minus[i] = base - onehot
            r = (vfs[i](*plus, **kws) - vfs[i](*minus, **kws)) / shifts[1]
            out.append(r.reshape(shape))
        return tuple(out) if len(nums) > 1 else out[0]

    return grad_f
