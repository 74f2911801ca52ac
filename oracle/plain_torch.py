"""The reference's plain method on the CPU, timed as the baseline only.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): used by the
``cpu_baseline`` leg of ``bench.py`` and by tests, never by the library.

Restates what tensorcircuit's pytorch backend computes for the HEA + TFIM
energy: gate-by-gate ``tensordot`` on the ``[2]*n`` state in insertion order
(``plain_contractor``, cons.py:334-370; leg wiring basecircuit.py:213-215),
energy as a sum of per-term ``expectation_ps`` contractions
(abstractcircuit.py:1208-1288) accumulated in the state dtype, and the
gradient by torch autograd (``value_and_grad`` = torch.func.grad_and_value,
backends/pytorch_backend.py:533-544).
"""
from __future__ import annotations

import time
from typing import List, Sequence, Tuple

import numpy as np
import torch


def _rot(pauli: torch.Tensor, th: torch.Tensor) -> torch.Tensor:
    c = torch.cos(th / 2).to(pauli.dtype)
    s = torch.sin(th / 2).to(pauli.dtype)
    return c * torch.eye(2, dtype=pauli.dtype) - 1j * s * pauli


def _apply(psi: torch.Tensor, g: torch.Tensor, qubits: Sequence[int], n: int) -> torch.Tensor:
    k = len(qubits)
    out = torch.tensordot(g.reshape([2] * (2 * k)), psi, dims=(list(range(k, 2 * k)), list(qubits)))
    rest = [q for q in range(n) if q not in qubits]
    perm = [0] * n
    for i, q in enumerate(qubits):
        perm[q] = i
    for j, q in enumerate(rest):
        perm[q] = k + j
    return out.permute(perm)


def hea_tfim_energy(params: torch.Tensor, n: int, nlayers: int, h: float = 1.0,
                    dtype=torch.complex64) -> torch.Tensor:
    """E(params) for params [2L, n]: ry/rz layers + CNOT chain, open-chain TFIM."""
    X = torch.tensor([[0, 1], [1, 0]], dtype=dtype)
    Y = torch.tensor([[0, -1j], [1j, 0]], dtype=dtype)
    Z = torch.tensor([[1, 0], [0, -1]], dtype=dtype)
    CX = torch.tensor([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=dtype)
    psi = torch.zeros([2] * n, dtype=dtype)
    psi.view(-1)[0] = 1.0
    for j in range(nlayers):
        for i in range(n):
            psi = _apply(psi, _rot(Y, params[2 * j, i]), (i,), n)
        for i in range(n):
            psi = _apply(psi, _rot(Z, params[2 * j + 1, i]), (i,), n)
        for i in range(n - 1):
            psi = _apply(psi, CX, (i, i + 1), n)
    conj = torch.conj(psi)
    e = torch.zeros((), dtype=dtype)
    for i in range(n - 1):
        phi = _apply(_apply(psi, Z, (i,), n), Z, (i + 1,), n)
        e = e - torch.sum(conj * phi)
    for i in range(n):
        e = e - h * torch.sum(conj * _apply(psi, X, (i,), n))
    return torch.real(e)


def time_energy_grad(n: int, nlayers: int, seed: int = 0, threads: int = 0) -> Tuple[float, float, List[float]]:
    """Seconds for one value_and_grad of hea_tfim_energy (after one warm-up
    at the same size if it is cheap); returns (seconds, energy, grad[:4])."""
    if threads > 0:
        torch.set_num_threads(threads)
    rng = np.random.default_rng(seed)
    p = torch.tensor(rng.uniform(0, 2 * np.pi, size=(2 * nlayers, n)).astype(np.float32))
    f = torch.func.grad_and_value(lambda q: hea_tfim_energy(q, n, nlayers))
    t0 = time.perf_counter()
    g, v = f(p)
    dt = time.perf_counter() - t0
    return dt, float(v), [float(x) for x in g.reshape(-1)[:4]]
