"""Measurement oracle -- TEST INFRASTRUCTURE ONLY (see the package docstring).

numpy restatements, in float64, of the reference's measurement algorithms:

* ``probability_sample``: backend.probability_sample
  (backends/abstract_backend.py:1124-1157): normalise p, take its cumulative
  sum, r = cum[-1] * (1 - status), draw index = searchsorted(cum, r) (side
  "left": the smallest i with cum[i] >= r).
* ``measure_jit``: BaseCircuit.measure_jit (basecircuit.py:371-445): qubits
  measured one after another; pu = P(qubit = 0, earlier outcomes) / p,
  outcome 1 iff status - pu + 0.31415926e-12 > 0, p *= pu or 1 - pu.
* ``correlation_from_counts`` / ``correlation_from_samples``
  (quantum.py:2400-2450): the parity prod (1 - 2 b_j) averaged over a
  probability vector or over samples.
* ``readout``: readouterror_bs (basecircuit.py:760-803), the per-qubit
  readout matrices [[e0, 1 - e1], [1 - e0, e1]] applied to p.

Pinned in tests/test_sampling.py by: the reference's KATs
(correlation_from_counts([0, 1], [0.6, 0.4, 0, 0]) = 0.2 and [1] -> 0.2,
quantum.py:2432-2436; searchsorted(edges=[-1, 3.3, 9.1, 10], [0, 4.1, 12])
= [1, 2, 4], tests/test_backends.py:278-281), brute-force cumulative
sums on tiny vectors, the closed form P(record) = |psi_idx|^2 of a full
sequential measurement, basis states measuring to themselves, and the
frequency of many draws against p.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np


def searchsorted_left(edges: np.ndarray, values: np.ndarray) -> np.ndarray:
    """Smallest i with edges[i] >= v (numpy's side='left'), by a scalar scan."""
    edges = np.asarray(edges, dtype=np.float64)
    out = []
    for v in np.asarray(values, dtype=np.float64).reshape(-1):
        i = 0
        while i < len(edges) and edges[i] < v:
            i += 1
        out.append(i)
    return np.array(out, dtype=np.int64)


def probability_sample(p: np.ndarray, status: np.ndarray) -> np.ndarray:
    """abstract_backend.py:1151-1157 in float64."""
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    p = p / p.sum()
    cum = np.cumsum(p)
    r = cum[-1] * (1.0 - np.asarray(status, dtype=np.float64).reshape(-1))
    return np.searchsorted(cum, r, side="left").astype(np.int64)


def boundary_margin(p: np.ndarray, status: np.ndarray) -> np.ndarray:
    """Distance of each r from the nearest cumulative-sum boundary, relative
    to the total: draws closer than rounding can move are excluded from
    exact comparisons."""
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    cum = np.cumsum(p)
    r = cum[-1] * (1.0 - np.asarray(status, dtype=np.float64).reshape(-1))
    j = np.searchsorted(cum, r, side="left")
    lo = np.where(j > 0, cum[np.maximum(j - 1, 0)], 0.0)
    hi = cum[np.minimum(j, len(cum) - 1)]
    return np.minimum(np.abs(r - lo), np.abs(hi - r)) / cum[-1]


def measure_jit(psi: np.ndarray, qubits: Sequence[int], status: Sequence[float]) -> Tuple[np.ndarray, float]:
    """basecircuit.py:371-445 on a dense state (qubit 0 = most significant bit)."""
    psi = np.asarray(psi, dtype=np.complex128).reshape(-1)
    n = int(round(np.log2(psi.size)))
    prob = np.abs(psi) ** 2
    idx = np.arange(psi.size)
    mask = np.ones(psi.size, dtype=bool)
    p = 1.0
    out = []
    eps = 0.31415926 * 1e-12
    for k, q in enumerate(qubits):
        bit = (idx >> (n - 1 - q)) & 1
        pu = prob[mask & (bit == 0)].sum() / p
        o = 1 if status[k] - pu + eps > 0 else 0
        p = p * (pu if o == 0 else 1.0 - pu)
        mask &= bit == o
        out.append(float(o))
    return np.array(out), p


def correlation_from_counts(index: Sequence[int], p: np.ndarray) -> float:
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    p = p / p.sum()
    n = int(round(np.log2(p.size)))
    i = np.arange(p.size)
    s = np.ones(p.size)
    for j in index:
        s = s * (1 - 2 * ((i >> (n - 1 - j)) & 1))
    return float((p * s).sum())


def correlation_from_samples(index: Sequence[int], samples: np.ndarray, n: int) -> float:
    s = np.asarray(samples, dtype=np.int64).reshape(-1)
    r = np.ones(s.size)
    for j in index:
        r = r * (1 - 2 * ((s >> (n - 1 - j)) & 1))
    return float(r.mean())


def readout(p: np.ndarray, readout_error: Sequence[Sequence[float]]) -> np.ndarray:
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    n = len(readout_error)
    t = p.reshape([2] * n)
    for i, (e0, e1) in enumerate(readout_error):
        m = np.array([[e0, 1 - e1], [1 - e0, e1]])
        t = np.moveaxis(np.tensordot(m, t, axes=([1], [i])), 0, i)
    return t.reshape(-1)
