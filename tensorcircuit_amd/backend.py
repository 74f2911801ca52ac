"""The pytorch backend object: the transform and array methods the
statevector path uses, with the semantics of the reference's
PyTorchBackend (backends/pytorch_backend.py:194-711)."""
from __future__ import annotations

from typing import Any, Callable, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import cons

Tensor = Any


class PyTorchBackend:
    name = "pytorch"

    # ---------------------------------------------------------- transforms --
    # All four transforms are views of one builder, ``_diff``: it turns f into
    # a function returning (value, aux, grads) with grads in the caller's
    # argnums shape (one tensor for an int, a tuple for a sequence).  The
    # public semantics are those of backends/pytorch_backend.py:497-709:
    #   grad(f)(x)                 -> grads            (has_aux: (grads, (aux,)))
    #   value_and_grad(f)(x)       -> (value, grads)   (has_aux: ((value, aux), grads))
    #   vmap(f, vec)(x)            -> f mapped over axis 0 of the args in vec
    #   vvag(f, argnums, vec)(x)   -> vmap of value_and_grad, with the grads of
    #                                 arguments outside vec summed over the batch

    @staticmethod
    def _positions(argnums: Union[int, Sequence[int]]) -> Tuple[Tuple[int, ...], bool]:
        """(argument positions, True when a single int was given)."""
        if isinstance(argnums, int):
            return (argnums,), True
        return tuple(int(a) for a in argnums), False

    def _diff(self, f: Callable[..., Any], argnums: Union[int, Sequence[int]],
              has_aux: bool) -> Callable[..., Tuple[Any, Any, Any]]:
        pos, single = self._positions(argnums)
        # torch.func differentiates with respect to the given positions; it
        # returns one gradient for an int argnums and a tuple otherwise
        tf_arg = pos[0] if single else pos
        inner = torch.func.grad_and_value(f, argnums=tf_arg, has_aux=has_aux)

        def run(*args: Any, **kws: Any) -> Tuple[Any, Any, Any]:
            grads, out = inner(*args, **kws)
            if has_aux:
                value, aux = out
                return value, aux, grads
            return out, None, grads

        return run

    def grad(self, f: Callable[..., Any], argnums: Union[int, Sequence[int]] = 0,
             has_aux: bool = False) -> Callable[..., Any]:
        """Gradient of a scalar f (pytorch_backend.py:497-509); with has_aux
        the auxiliary output comes back as a 1-tuple beside the grads."""
        run = self._diff(f, argnums, has_aux)

        def grad_f(*args: Any, **kws: Any) -> Any:
            _, aux, grads = run(*args, **kws)
            return (grads, (aux,)) if has_aux else grads

        return grad_f

    def value_and_grad(self, f: Callable[..., Any], argnums: Union[int, Sequence[int]] = 0,
                       has_aux: bool = False) -> Callable[..., Tuple[Any, Any]]:
        """(value, grads) of a scalar f (pytorch_backend.py:533-544); with
        has_aux the value slot holds (value, aux)."""
        run = self._diff(f, argnums, has_aux)

        def vag_f(*args: Any, **kws: Any) -> Tuple[Any, Any]:
            value, aux, grads = run(*args, **kws)
            return ((value, aux) if has_aux else value), grads

        return vag_f

    def vmap(self, f: Callable[..., Any], vectorized_argnums: Union[int, Sequence[int]] = 0) -> Any:
        """Map f over the leading axis of the vectorised positional arguments
        (pytorch_backend.py:605-617); the others are broadcast."""
        vec = frozenset(self._positions(vectorized_argnums)[0])

        def mapped(*args: Any, **kws: Any) -> Any:
            dims = tuple(0 if k in vec else None for k in range(len(args)))
            return torch.vmap(f, in_dims=dims, out_dims=0)(*args, **kws)

        return mapped

    def vectorized_value_and_grad(self, f: Callable[..., Any], argnums: Union[int, Sequence[int]] = 0,
                                  vectorized_argnums: Union[int, Sequence[int]] = 0,
                                  has_aux: bool = False) -> Callable[..., Tuple[Any, Any]]:
        """vmap of value_and_grad; the gradient of an argument that is not
        vectorised is the sum of its per-sample gradients
        (pytorch_backend.py:673-709)."""
        pos, single = self._positions(argnums)
        vec = self._positions(vectorized_argnums)[0]
        batched = self.vmap(self.value_and_grad(f, argnums, has_aux), vectorized_argnums=vec)
        # which of the returned gradients need the batch axis summed away
        fold = tuple(p not in vec for p in pos)

        def sum_batch(g: Any) -> Any:
            return self.tree_map(lambda t: torch.sum(t, dim=0), g)

        def vvag_f(*args: Any, **kws: Any) -> Tuple[Any, Any]:
            values, grads = batched(*args, **kws)
            if single:
                return values, (sum_batch(grads) if fold[0] else grads)
            return values, tuple(sum_batch(g) if s else g for g, s in zip(grads, fold))

        return vvag_f

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
