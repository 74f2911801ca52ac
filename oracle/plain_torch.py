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


def _layer(psi: torch.Tensor, ry: torch.Tensor, rz: torch.Tensor, n: int) -> torch.Tensor:
    """One HEA layer: ry on every qubit, rz on every qubit, CNOT chain."""
    dtype = psi.dtype
    Y = torch.tensor([[0, -1j], [1j, 0]], dtype=dtype)
    Z = torch.tensor([[1, 0], [0, -1]], dtype=dtype)
    CX = torch.tensor([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=dtype)
    for i in range(n):
        psi = _apply(psi, _rot(Y, ry[i]), (i,), n)
    for i in range(n):
        psi = _apply(psi, _rot(Z, rz[i]), (i,), n)
    for i in range(n - 1):
        psi = _apply(psi, CX, (i, i + 1), n)
    return psi


def hea_tfim_energy(params: torch.Tensor, n: int, nlayers: int, h: float = 1.0,
                    dtype=torch.complex64, remat: bool = False) -> torch.Tensor:
    """E(params) for params [2L, n]: ry/rz layers + CNOT chain, open-chain
    TFIM.  ``remat``: each layer under torch.utils.checkpoint (its
    intermediate states are recomputed in the backward pass instead of
    stored), the reference's per-layer memory-saving pattern
    (examples/checkpoint_memsave.py:41-104)."""
    X = torch.tensor([[0, 1], [1, 0]], dtype=dtype)
    Z = torch.tensor([[1, 0], [0, -1]], dtype=dtype)
    psi = torch.zeros([2] * n, dtype=dtype)
    psi.view(-1)[0] = 1.0
    for j in range(nlayers):
        if remat:
            psi = torch.utils.checkpoint.checkpoint(_layer, psi, params[2 * j], params[2 * j + 1], n,
                                                    use_reentrant=False)
        else:
            psi = _layer(psi, params[2 * j], params[2 * j + 1], n)
    conj = torch.conj(psi)
    e = torch.zeros((), dtype=dtype)
    for i in range(n - 1):
        phi = _apply(_apply(psi, Z, (i,), n), Z, (i + 1,), n)
        e = e - torch.sum(conj * phi)
    for i in range(n):
        e = e - h * torch.sum(conj * _apply(psi, X, (i,), n))
    return torch.real(e)


def time_energy_grad(n: int, nlayers: int, seed: int = 0, threads: int = 0,
                     remat: bool = False) -> Tuple[float, float, List[float]]:
    """Seconds for one energy + gradient (reverse-mode autograd, as
    value_and_grad) of hea_tfim_energy; returns (seconds, energy, grad[:4])."""
    import torch.utils.checkpoint  # noqa: F401

    if threads > 0:
        torch.set_num_threads(threads)
    rng = np.random.default_rng(seed)
    p = torch.tensor(rng.uniform(0, 2 * np.pi, size=(2 * nlayers, n)).astype(np.float32), requires_grad=True)
    t0 = time.perf_counter()
    v = hea_tfim_energy(p, n, nlayers, remat=remat)
    v.backward()
    dt = time.perf_counter() - t0
    return dt, float(v.detach()), [float(x) for x in p.grad.reshape(-1)[:4]]


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"
