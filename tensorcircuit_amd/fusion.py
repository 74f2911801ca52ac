"""Host-side gate fusion: constant gates are multiplied into dense blocks of
at most ``kmax`` qubits (5 fills a 32 x 32 block, the unit the FP32 matrix
cores apply), reordering only gates on disjoint qubits.

The reference contracts gates by an opt_einsum path and merges single-qubit
nodes into neighbours (cons.py:236-279); a dense statevector engine instead
wants few, large, dense unitaries -- one HBM pass each.
"""
from __future__ import annotations

import itertools
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


def _first_fit(qsets: Sequence[Tuple[int, ...]], kmax: int) -> List[Tuple[List[int], List[int]]]:
    """One block per sweep over the remaining gates: a gate joins the block
    if it touches no gate deferred before it and the block still spans at
    most ``kmax`` qubits."""
    remaining = list(range(len(qsets)))
    out = []
    while remaining:
        qs: set = set()
        block: List[int] = []
        blocked: set = set()
        deferred: List[int] = []
        for i in remaining:
            q = set(qsets[i])
            if q & blocked or len(qs | q) > kmax:
                deferred.append(i)
                blocked |= q
            else:
                qs |= q
                block.append(i)
        out.append((sorted(qs), block))
        remaining = deferred
    return out


def _window(qsets: Sequence[Tuple[int, ...]], n: int, kmax: int) -> List[Tuple[List[int], List[int]]]:
    """Blocks chosen one at a time by the gates they absorb.  A gate is ready
    when every earlier gate on its qubits is in an emitted block.  For each
    ready gate the qubit set starts as its qubits and grows by one or two
    qubits at a time (taken from the next gates on the set's qubits), each
    time by the growth that absorbs the most two-qubit gates; the set whose
    closure absorbs the most is emitted as the next block.  A qubit set S
    absorbs every gate on qubits within S whose earlier gates are all
    emitted or absorbed.  Single-qubit gates weigh 1/100 of a two-qubit gate
    (any block can take them)."""
    queues: List[List[int]] = [[] for _ in range(n)]
    for i, qs in enumerate(qsets):
        for q in qs:
            queues[q].append(i)
    head = [0] * n
    left = len(qsets)
    out = []

    def front(q: int, hd: List[int]) -> int:
        return queues[q][hd[q]] if hd[q] < len(queues[q]) else -1

    def absorb(S: set, hd0: List[int]) -> Tuple[List[int], List[int]]:
        hd = list(hd0)
        taken: List[int] = []
        grew = True
        while grew:
            grew = False
            for q in sorted(S):
                while hd[q] < len(queues[q]):
                    g = queues[q][hd[q]]
                    if not all(x in S and front(x, hd) == g for x in qsets[g]):
                        break
                    for x in qsets[g]:
                        hd[x] += 1
                    taken.append(g)
                    grew = True
        return taken, hd

    def weight(taken: List[int]) -> float:
        return sum(1.0 if len(qsets[g]) > 1 else 0.01 for g in taken)

    while left:
        ready = sorted({front(q, head) for q in range(n)} - {-1})
        ready = [g for g in ready if all(front(x, head) == g for x in qsets[g])]
        best = None
        for g in ready:
            S = set(qsets[g])
            while len(S) < kmax:
                _, hd = absorb(S, head)
                near: set = set()
                for q in S:
                    for h in queues[q][hd[q]:hd[q] + 3]:
                        near |= set(qsets[h])
                near = sorted(near - S)
                if not near:
                    break
                grow = [(q,) for q in near]
                if len(S) + 2 <= kmax:
                    grow += list(itertools.combinations(near, 2))
                # ties go to the smaller growth, then to the lower qubits
                S |= set(max(grow, key=lambda o: (weight(absorb(S | set(o), head)[0]), -len(o))))
            taken, hd = absorb(S, head)
            w = weight(taken)
            if best is None or w > best[0]:
                best = (w, S, taken, hd)
        _, S, taken, head = best
        left -= len(taken)
        out.append((sorted(S), taken))
    return out


def fuse(specs: Sequence[GateSpec], n: int, kmax: int = 5, pad: bool = True,
         method: str = "window") -> List[GateSpec]:
    """Fusion of constant gates (kind MAT) into dense blocks of at most
    ``kmax`` qubits, in an order that keeps every gate after the gates it
    follows on shared qubits.  ``method``: "window" (default) picks each
    block's qubit set by the number of gates it absorbs; "first_fit" sweeps
    the gates in order and closes a block at the first gate that does not
    fit.  With ``pad`` every block is widened to exactly ``kmax`` qubits
    (identity on the extras)."""
    for s in specs:
        if s.kind != _lib.G_MAT:
            raise ValueError("fuse() takes constant (MAT) gates only")
        qs = tuple(s.qubits)
        if not qs or len(set(qs)) != len(qs) or min(qs) < 0 or max(qs) >= n:
            raise ValueError("gate qubits %r: need distinct qubits in [0, %d)" % (qs, n))
        if len(qs) > kmax:
            raise ValueError("a gate on %d qubits does not fit a %d-qubit block" % (len(qs), kmax))
    qsets = [tuple(s.qubits) for s in specs]
    if method == "window":
        groups = _window(qsets, n, kmax)
    elif method == "first_fit":
        groups = _first_fit(qsets, kmax)
    else:
        raise ValueError("unknown fusion method %r" % (method,))
    out: List[GateSpec] = []
    for qs_sorted, block in groups:
        qs = set(qs_sorted)
        if pad and len(qs) < kmax <= n:
            for q in range(n):
                if len(qs) >= kmax:
                    break
                qs.add(q)
        order = sorted(qs)
        k = len(order)
        loc = {q: i for i, q in enumerate(order)}
        m = np.eye(2**k, dtype=np.complex128)
        for i in block:
            g = specs[i]
            m = _apply_local(m, np.asarray(g.mat, dtype=np.complex128), [loc[q] for q in g.qubits], k)
        out.append(GateSpec(_lib.G_MAT, tuple(order), -1, 0.0, 1.0, m))
    return out
