"""Host-side gate fusion: constant gates are multiplied into dense blocks of
at most ``kmax`` qubits (5 fills a 32 x 32 block, the unit the FP32 matrix
cores apply), reordering only gates on disjoint qubits.

The reference contracts gates by an opt_einsum path and merges single-qubit
nodes into neighbours (cons.py:236-279); a dense statevector engine instead
wants few, large, dense unitaries -- one HBM pass each.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from . import _lib
from .engine import GateSpec


def _apply_local(m: np.ndarray, u: np.ndarray, pos: Sequence[int], k: int) -> np.ndarray:
    """Left-multiply the 2^k x 2^k matrix ``m`` by gate ``u`` acting on local
    qubits ``pos`` (local qubit 0 = most significant)."""
    g = len(pos)
    t = m.reshape([2] * k + [m.shape[1]])
    ug = u.reshape([2] * (2 * g))
    out = np.tensordot(ug, t, axes=(list(range(g, 2 * g)), list(pos)))
    rest = [q for q in range(k) if q not in pos]
    perm = [0] * (k + 1)
    for i, q in enumerate(pos):
        perm[q] = i
    for j, q in enumerate(rest):
        perm[q] = g + j
    perm[k] = k
    return np.transpose(out, perm).reshape(2**k, m.shape[1])


def fuse(specs: Sequence[GateSpec], n: int, kmax: int = 5, pad: bool = True) -> List[GateSpec]:
    """Greedy fusion of constant gates (kind MAT) into dense blocks.  A gate
    joins the current block if it touches no gate deferred before it and the
    block still spans at most ``kmax`` qubits.  With ``pad`` every block is
    widened to exactly ``kmax`` qubits (identity on the extras)."""
    for s in specs:
        if s.kind != _lib.G_MAT:
            raise ValueError("fuse() takes constant (MAT) gates only")
    remaining = list(specs)
    out: List[GateSpec] = []
    while remaining:
        qs: set = set()
        block: List[GateSpec] = []
        blocked: set = set()
        deferred: List[GateSpec] = []
        for g in remaining:
            q = set(g.qubits)
            if q & blocked:
                deferred.append(g)
                blocked |= q
            elif len(qs | q) <= kmax:
                qs |= q
                block.append(g)
            else:
                deferred.append(g)
                blocked |= q
        if pad and len(qs) < kmax <= n:
            for q in range(n):
                if len(qs) >= kmax:
                    break
                qs.add(q)
        order = sorted(qs)
        k = len(order)
        loc = {q: i for i, q in enumerate(order)}
        m = np.eye(2**k, dtype=np.complex128)
        for g in block:
            m = _apply_local(m, np.asarray(g.mat, dtype=np.complex128), [loc[q] for q in g.qubits], k)
        out.append(GateSpec(_lib.G_MAT, tuple(order), -1, 0.0, 1.0, m))
        remaining = deferred
    return out
