"""Host-side gate fusion: constant gates are multiplied into dense blocks of
at most ``kmax`` qubits (5 fills a 32 x 32 block, the unit the FP32 matrix
cores apply), reordering only gates on disjoint qubits.

The reference contracts gates by an opt_einsum path and merges single-qubit
nodes into neighbours (cons.py:236-279); a dense statevector engine instead
wants few, large, dense unitaries -- one HBM pass each.
"""
from __future__ import annotations

import itertools
import random
from typing import List, Optional, Sequence, Tuple

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


def _window(qsets: Sequence[Tuple[int, ...]], n: int, kmax: int,
            rng: Optional[random.Random] = None) -> List[Tuple[List[int], List[int]]]:
    """Blocks chosen one at a time by the gates they absorb.  A gate is ready
    when every earlier gate on its qubits is in an emitted block.  For each
    ready gate the qubit set starts as its qubits and grows by one or two
    qubits at a time (taken from the next gates on the set's qubits), each
    time by the growth that absorbs the most two-qubit gates; the set whose
    closure absorbs the most is emitted as the next block.  A qubit set S
    absorbs every gate on qubits within S whose earlier gates are all
    emitted or absorbed.  Single-qubit gates weigh 1/100 of a two-qubit gate
    (any block can take them).  With ``rng``, one block in five is the
    second best instead (a randomised variant for the search below)."""
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
        best = second = None
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
                second, best = best, (w, S, taken, hd)
            elif second is None or w > second[0]:
                second = (w, S, taken, hd)
        if rng is not None and second is not None and rng.random() < 0.2:
            best = second
        _, S, taken, head = best
        left -= len(taken)
        out.append((sorted(S), taken))
    return out


# Time of a matrix-core tile pass on a 30-qubit complex64 state: 1.52 ms per
# pass plus 1.88 ms per block in it (k_mfma_tile<13> on MI355X, least squares
# over the passes of the 30-qubit depth-20 brickwork).
_PASS_MS, _BLOCK_MS = 1.52, 1.88


def _search(qsets: Sequence[Tuple[int, ...]], n: int, kmax: int, tries: int,
            seed: int) -> List[Tuple[List[int], List[int]]]:
    """The window choice and ``tries`` - 1 seeded randomised variants of it,
    each scored by the tile passes the matrix-core scheduler makes of its
    blocks (``_PASS_MS`` per pass + ``_BLOCK_MS`` per block); the cheapest
    wins, the first on ties.  The scheduler sees placeholder matrices: the
    pass grouping depends on the blocks' qubits only."""
    from .engine import Program

    rng = random.Random(seed)
    dummy = np.linalg.qr(np.random.default_rng(seed).normal(size=(32, 32)))[0].astype(np.complex128)
    best = None
    for t in range(max(1, tries)):
        groups = _window(qsets, n, kmax, rng if t else None)
        if kmax == 5 and n >= 10:
            blocks = [GateSpec(_lib.G_MAT, tuple(_padded(S, n, kmax)), -1, 0.0, 1.0, dummy) for S, _ in groups]
            cost = _PASS_MS * Program(n, blocks, 0, fuse=1, mfma=2).npasses + _BLOCK_MS * len(blocks)
        else:
            cost = float(len(groups))
        if best is None or cost < best[0]:
            best = (cost, groups)
    return best[1]


def _padded(qs: Sequence[int], n: int, kmax: int) -> List[int]:
    """``qs`` widened to ``kmax`` qubits with the lowest unused ones."""
    out = set(qs)
    for q in range(n):
        if len(out) >= kmax:
            break
        out.add(q)
    return sorted(out)


def fuse(specs: Sequence[GateSpec], n: int, kmax: int = 5, pad: bool = True,
         method: str = "window", tries: int = 64, seed: int = 0) -> List[GateSpec]:
    """Fusion of constant gates (kind MAT) into dense blocks of at most
    ``kmax`` qubits, in an order that keeps every gate after the gates it
    follows on shared qubits.  ``method``: "window" (default) picks each
    block's qubit set by the number of gates it absorbs; "search" also tries
    ``tries`` - 1 randomised variants of that choice (seeded by ``seed``) and
    keeps the one whose blocks the matrix-core tile passes run fastest
    (fewest blocks and passes); "first_fit" sweeps the gates in order and
    closes a block at the first gate that does not fit.  With ``pad`` every
    block is widened to exactly ``kmax`` qubits (identity on the extras)."""
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
    elif method == "search":
        groups = _search(qsets, n, kmax, tries, seed)
    elif method == "first_fit":
        groups = _first_fit(qsets, kmax)
    else:
        raise ValueError("unknown fusion method %r" % (method,))
    out: List[GateSpec] = []
    for qs_sorted, block in groups:
        order = _padded(qs_sorted, n, kmax) if pad and len(qs_sorted) < kmax <= n else sorted(qs_sorted)
        k = len(order)
        loc = {q: i for i, q in enumerate(order)}
        m = np.eye(2**k, dtype=np.complex128)
        for i in block:
            g = specs[i]
            m = _apply_local(m, np.asarray(g.mat, dtype=np.complex128), [loc[q] for q in g.qubits], k)
        out.append(GateSpec(_lib.G_MAT, tuple(order), -1, 0.0, 1.0, m))
    return out
