"""The pytorch backend object: the transform and array methods the
statevector path uses, with the semantics of the reference's
PyTorchBackend (backends/pytorch_backend.py:194-711)."""
from __future__ import annotations

from functools import partial
from typing import Any, Callable, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import cons

Tensor = Any


class PyTorchBackend:
    name = "pytorch"

    # ---------------------------------------------------------- transforms --
    def grad(self, f: Callable[..., Any], argnums: Union[int, Sequence[int]] = 0,
             has_aux: bool = False) -> Callable[..., Any]:
        """pytorch_backend.py:497-509."""

        def wrapper(*args: Any, **kws: Any) -> Any:
            y, gr = self.value_and_grad(f, argnums, has_aux)(*args, **kws)
            if has_aux:
                return gr, y[1:]
            return gr

        return wrapper

    def value_and_grad(self, f: Callable[..., Any], argnums: Union[int, Sequence[int]] = 0,
                       has_aux: bool = False) -> Callable[..., Tuple[Any, Any]]:
        """pytorch_backend.py:533-544 (torch.func.grad_and_value)."""

        def wrapper(*args: Any, **kws: Any) -> Any:
            gavf = torch.func.grad_and_value(f, argnums=argnums, has_aux=has_aux)
            g, v = gavf(*args, **kws)
            return v, g

        return wrapper

    def vmap(self, f: Callable[..., Any], vectorized_argnums: Union[int, Sequence[int]] = 0) -> Any:
        """pytorch_backend.py:605-617."""
        if isinstance(vectorized_argnums, int):
            vectorized_argnums = (vectorized_argnums,)

        def wrapper(*args: Any, **kws: Any) -> Tensor:
            in_axes = tuple([0 if i in vectorized_argnums else None for i in range(len(args))])
            return torch.vmap(f, in_axes, 0)(*args, **kws)

        return wrapper

    def vectorized_value_and_grad(self, f: Callable[..., Any], argnums: Union[int, Sequence[int]] = 0,
                                  vectorized_argnums: Union[int, Sequence[int]] = 0,
                                  has_aux: bool = False) -> Callable[..., Tuple[Any, Any]]:
        """vmap(value_and_grad); grads of non-vectorized args are summed over
        the batch (pytorch_backend.py:673-709)."""
        if isinstance(vectorized_argnums, int):
            vectorized_argnums = (vectorized_argnums,)

        def wrapper(*args: Any, **kws: Any) -> Tuple[Tensor, Union[Tensor, Tuple[Tensor, ...]]]:
            jf = self.value_and_grad(f, argnums=argnums, has_aux=has_aux)
            jf = self.vmap(jf, vectorized_argnums=vectorized_argnums)
            vs, gs = jf(*args, **kws)
            if isinstance(argnums, int):
                argnums_list = [argnums]
                gs = [gs]
            else:
                argnums_list = list(argnums)
                gs = list(gs)
            for i, (j, g) in enumerate(zip(argnums_list, gs)):
                if j not in vectorized_argnums:
                    gs[i] = self.tree_map(partial(torch.sum, dim=0), g)
            if isinstance(argnums, int):
                gs = gs[0]
            else:
                gs = tuple(gs)
            return vs, gs

        return wrapper

    vvag = vectorized_value_and_grad

    def jit(self, f: Callable[..., Any], static_argnums: Optional[Union[int, Sequence[int]]] = None,
            jit_compile: Optional[bool] = None, **kws: Any) -> Any:
        """No tracing compiler: programs are built once and cached by the engine
        (pytorch_backend.py:659-669 returns f unchanged by default too)."""
        return f

    def vjp(self, f: Callable[..., Any], inputs: Any, v: Any) -> Tuple[Any, Any]:
        if isinstance(inputs, list):
            inputs = tuple(inputs)
        if isinstance(v, list):
            v = tuple(v)
        return torch.autograd.functional.vjp(f, inputs, v)

    def tree_map(self, f: Callable[..., Any], *pytrees: Any) -> Any:
        return torch.utils._pytree.tree_map(f, *pytrees)

    # --------------------------------------------------------------- arrays --
    def convert_to_tensor(self, a: Any, dtype: Optional[str] = None) -> Tensor:
        t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
        return t if dtype is None else self.cast(t, dtype)

    def cast(self, a: Tensor, dtype: str) -> Tensor:
        if isinstance(dtype, str):
            dtype = getattr(torch, dtype)
        if not isinstance(a, torch.Tensor):
            a = torch.as_tensor(np.asarray(a))
        return a.to(dtype)

    def numpy(self, a: Tensor) -> np.ndarray:
        if isinstance(a, torch.Tensor):
            return a.detach().cpu().numpy()
        return np.asarray(a)

    def real(self, a: Tensor) -> Tensor:
        return torch.real(a) if a.is_complex() else a

    def imag(self, a: Tensor) -> Tensor:
        return torch.imag(a) if a.is_complex() else torch.zeros_like(a)

    def conj(self, a: Tensor) -> Tensor:
        return torch.conj(a)

    def i(self, dtype: Any = None) -> Tensor:
        dtype = getattr(torch, dtype or cons.dtypestr) if isinstance(dtype, (str, type(None))) else dtype
        return torch.tensor(1j, dtype=dtype)

    def ones(self, shape: Sequence[int], dtype: Optional[str] = None) -> Tensor:
        return torch.ones(tuple(shape), dtype=getattr(torch, dtype or cons.dtypestr))

    def zeros(self, shape: Sequence[int], dtype: Optional[str] = None) -> Tensor:
        return torch.zeros(tuple(shape), dtype=getattr(torch, dtype or cons.dtypestr))

    def eye(self, n: int, m: Optional[int] = None, dtype: Optional[str] = None) -> Tensor:
        return torch.eye(n, m if m is not None else n, dtype=getattr(torch, dtype or cons.dtypestr))

    def reshape(self, a: Tensor, shape: Sequence[int]) -> Tensor:
        return torch.reshape(a, tuple(shape))

    def reshape2(self, a: Tensor) -> Tensor:
        """backends/abstract_backend.py:390-400."""
        n = int(round(np.log2(int(np.prod(a.shape)))))
        return torch.reshape(a, [2] * n)

    def reshapem(self, a: Tensor) -> Tensor:
        """backends/abstract_backend.py:403-414."""
        d = int(round(np.sqrt(int(np.prod(a.shape)))))
        return torch.reshape(a, (d, d))

    def stack(self, a: Sequence[Tensor], axis: int = 0) -> Tensor:
        return torch.stack(list(a), dim=axis)

    def concat(self, a: Sequence[Tensor], axis: int = 0) -> Tensor:
        return torch.cat(list(a), dim=axis)

    def sum(self, a: Tensor, axis: Optional[int] = None, keepdims: bool = False) -> Tensor:
        return torch.sum(a) if axis is None else torch.sum(a, dim=axis, keepdim=keepdims)

    def mean(self, a: Tensor, axis: Optional[int] = None, keepdims: bool = False) -> Tensor:
        return torch.mean(a) if axis is None else torch.mean(a, dim=axis, keepdim=keepdims)

    def abs(self, a: Tensor) -> Tensor:
        return torch.abs(a)

    def sqrt(self, a: Tensor) -> Tensor:
        return torch.sqrt(a)

    def cos(self, a: Tensor) -> Tensor:
        return torch.cos(a)

    def sin(self, a: Tensor) -> Tensor:
        return torch.sin(a)

    def exp(self, a: Tensor) -> Tensor:
        return torch.exp(a)

    def norm(self, a: Tensor) -> Tensor:
        return torch.linalg.norm(a)

    def kron(self, a: Tensor, b: Tensor) -> Tensor:
        return torch.kron(a, b)

    def adjoint(self, a: Tensor) -> Tensor:
        return torch.conj(torch.transpose(a, -2, -1))

    def arange(self, start: int, stop: Optional[int] = None, step: int = 1) -> Tensor:
        return torch.arange(start, stop, step) if stop is not None else torch.arange(start)

    def is_tensor(self, a: Any) -> bool:
        return isinstance(a, torch.Tensor)

    def shape_tuple(self, a: Tensor) -> Tuple[int, ...]:
        return tuple(a.shape)

    def sizen(self, a: Tensor) -> int:
        return int(a.numel())

    def implicit_randn(self, shape: Union[int, Sequence[int]] = 1, mean: float = 0.0, stddev: float = 1.0,
                       dtype: str = "float32") -> Tensor:
        if isinstance(shape, int):
            shape = (shape,)
        return mean + stddev * torch.randn(tuple(shape), dtype=getattr(torch, dtype))

    def implicit_randu(self, shape: Union[int, Sequence[int]] = 1, low: float = 0.0, high: float = 1.0,
                       dtype: str = "float32") -> Tensor:
        if isinstance(shape, int):
            shape = (shape,)
        return low + (high - low) * torch.rand(tuple(shape), dtype=getattr(torch, dtype))

    # ---- randomness and sampling (abstract_backend.py:1124-1157) ----------
    def get_random_state(self, seed: Optional[int] = None) -> torch.Generator:
        g = torch.Generator()
        if seed is not None:
            g.manual_seed(int(seed))
        else:
            g.seed()
        return g

    def stateful_randu(self, g: torch.Generator, shape: Union[int, Sequence[int]] = 1, low: float = 0.0,
                       high: float = 1.0, dtype: str = "float32") -> Tensor:
        if isinstance(shape, int):
            shape = (shape,)
        return low + (high - low) * torch.rand(tuple(shape), generator=g, dtype=getattr(torch, dtype))

    def probability_sample(self, shots: int, p: Tensor, status: Optional[Tensor] = None, g: Any = None) -> Tensor:
        """``shots`` indices drawn from p (normalised) by inverse CDF on the
        device: r = total * (1 - status), index = smallest i whose prefix sum
        reaches r.  p of any length (zero-padded to a power of two)."""
        from . import engine

        pt = torch.as_tensor(p).reshape(-1)
        if pt.is_complex():
            pt = torch.real(pt)
        d = int(pt.shape[0])
        n = max(1, (d - 1).bit_length())
        if d < (1 << n):
            pt = torch.cat([pt, torch.zeros((1 << n) - d, dtype=pt.dtype, device=pt.device)])
        if status is not None:
            u = torch.as_tensor(status, dtype=torch.float64).reshape(1, -1)
        else:
            u = torch.rand((1, int(shots)), dtype=torch.float64, generator=g)
        return engine.sample_prob_indices(pt, n, u)[0]

    def __repr__(self) -> str:
        return "pytorch_backend"


backend = PyTorchBackend()
