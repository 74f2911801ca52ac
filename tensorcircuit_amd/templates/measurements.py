"""Measurement shortcuts (the reference's templates/measurements.py) on the
fused fp64 Pauli-sum kernel: all terms of a Hamiltonian in one call instead
of one contraction per term."""
from __future__ import annotations

from typing import Any, Sequence

import numpy as np
import torch

from .. import engine


def _codes_of(structures: Any, n: int) -> np.ndarray:
    s = structures.detach().cpu().numpy() if isinstance(structures, torch.Tensor) else np.asarray(structures)
    s = np.asarray(s)
    if s.ndim == 3:  # onehot [nterms, n, 4]
        s = np.argmax(s, axis=-1)
    return s.reshape(-1, n).astype(np.int8)


def pauli_sum_expectation(c: Any, structures: Any, weights: Any = None, reuse: bool = True) -> Any:
    """sum_k w_k <P_k> as a complex scalar, one fused fp64 reduction; the
    gradient flows through one H|psi> application."""
    n = c._nstate
    codes = _codes_of(structures, n)
    w = np.ones(codes.shape[0]) if weights is None else np.asarray(
        weights.detach().cpu().numpy() if isinstance(weights, torch.Tensor) else weights, dtype=np.float64)
    return c._pauli_expectation(codes, w.reshape(-1), reuse)


def any_measurements(c: Any, structures: Any, onehot: bool = False, reuse: bool = True) -> Any:
    """Real expectation of each Pauli structure (templates/measurements.py:18-83);
    a single structure returns a scalar."""
    n = c._nstate
    codes = _codes_of(structures, n)
    vals = [torch.real(c._pauli_expectation(codes[k:k + 1], np.ones(1), reuse)) for k in range(codes.shape[0])]
    if np.asarray(structures).ndim == (2 if onehot else 1):
        return vals[0]
    return torch.stack(vals)


def parameterized_local_measurements(c: Any, structures: Any, onehot: bool = False) -> Any:
    """<P_i> of the single-qubit Pauli on every qubit (templates/measurements.py:86-120)."""
    n = c._nstate
    s = _codes_of(structures, n).reshape(-1)
    vals = []
    for i in range(n):
        codes = np.zeros((1, n), dtype=np.int8)
        codes[0, i] = s[i]
        vals.append(torch.real(c._pauli_expectation(codes, np.ones(1))))
    return torch.stack(vals)


def operator_expectation(c: Any, h: Any) -> Any:
    """Real <psi|H|psi> for H a PauliSum (fused Pauli kernel), a sparse
    matrix (CSR kernel) or a dense matrix (templates/measurements.py:156-174)."""
    if isinstance(h, engine.PauliSum):
        return torch.real(engine.PauliExpFn.apply(c._state(), h))
    if isinstance(h, torch.Tensor) and h.layout in (torch.sparse_coo, torch.sparse_csr):
        return sparse_expectation(c, h)
    w = c._state().reshape(-1, 1)
    hd = torch.as_tensor(h).to(device=w.device, dtype=w.dtype)
    return torch.real((torch.conj(w).T @ (hd @ w))[0, 0])


def sparse_expectation(c: Any, hamiltonian: Any) -> Any:
    """Real <psi|H|psi> for a sparse H (COO or CSR, e.g. from
    quantum.PauliStringSum2COO) by the CSR kernel with fp64 accumulation
    (templates/measurements.py:177-191)."""
    return torch.real(engine.csr_expectation(hamiltonian, c._state(), c._nstate))
