"""TorchLayer: a torch.nn.Module around a quantum function (torchnn.py:16-90
of the reference).  Weights are module parameters; inputs are batched
with ``backend.vmap`` so a batch of inputs runs as one batched circuit
evaluation on the device."""
from __future__ import annotations

from typing import Any, Callable, Sequence, Tuple, Union

import torch

from .backend import backend
from .interfaces import torch_interface

Tensor = Any


class QuantumNet(torch.nn.Module):
    def __init__(self, f: Callable[..., Any], weights_shape: Sequence[Tuple[int, ...]],
                 initializer: Union[Any, Sequence[Any]] = None, use_vmap: bool = True,
                 vectorized_argnums: Union[int, Sequence[int]] = 0, use_interface: bool = False,
                 use_jit: bool = False, enable_dlpack: bool = False) -> None:
        """``f(*inputs, *weights)``: inputs first, then one tensor per entry of
        ``weights_shape``.  ``use_interface`` defaults to False: ``f`` is a
        torch function already (the reference makes the same choice for its
        pytorch backend, tests/test_torchnn.py:37-40)."""
        super().__init__()
        if use_vmap:
            f = backend.vmap(f, vectorized_argnums=vectorized_argnums)
        if use_interface:
            f = torch_interface(f, jit=use_jit, enable_dlpack=enable_dlpack)
        elif use_jit:
            f = backend.jit(f)
        self.f = f
        self.q_weights = torch.nn.ParameterList()
        if isinstance(weights_shape[0], int):
            weights_shape = [weights_shape]  # type: ignore[list-item]
        if not isinstance(initializer, (list, tuple)):
            initializer = [initializer] * len(weights_shape)
        for ws, initf in zip(weights_shape, initializer):
            initf = torch.randn if initf is None else initf
            self.q_weights.append(torch.nn.Parameter(initf(tuple(ws))))

    def forward(self, *inputs: Tensor) -> Tensor:
        return self.f(*inputs, *self.q_weights)


TorchLayer = QuantumNet
