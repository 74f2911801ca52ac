"""Seeded synthetic inputs of the benchmark configurations (no datasets):
HEA parameters, TFIM terms, and the Haar-SU(2) + CZ brickwork."""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from . import _lib
from .engine import GateSpec

CZ = np.diag([1.0, 1.0, 1.0, -1.0]).astype(np.complex128)


def hea_params(seed: int, nlayers: int, n: int, batch: int = 0) -> np.ndarray:
    """uniform[0, 2pi) angles drawn in float64 and rounded to float32, laid out
    [2L, n] (or [batch, 2L, n])."""
    rng = np.random.default_rng(seed)
    shape = (2 * nlayers, n) if batch == 0 else (batch, 2 * nlayers, n)
    return rng.uniform(0, 2 * np.pi, size=shape).astype(np.float32)


def tfim_terms(n: int, h: float = 1.0, j: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Open-chain H = -j sum Z_i Z_{i+1} - h sum X_i as (codes [2n-1, n], weights)."""
    codes, w = [], []
    for i in range(n - 1):
        c = [0] * n
        c[i] = c[i + 1] = 3
        codes.append(c)
        w.append(-j)
    for i in range(n):
        c = [0] * n
        c[i] = 1
        codes.append(c)
        w.append(-h)
    return np.array(codes, dtype=np.int8), np.array(w, dtype=np.float64)


def haar_su2(rng: np.random.Generator) -> np.ndarray:
    """Haar-random SU(2): QR of a complex Ginibre matrix, column phases fixed
    by the signs of diag(R), divided by sqrt(det) (Mezzadri 2007)."""
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))
    return q / np.sqrt(np.linalg.det(q))


def brickwork(n: int, depth: int, seed: int) -> List[GateSpec]:
    """depth layers of Haar SU(2) on every qubit, then CZ on (2i, 2i+1) for
    even layers and (2i+1, 2i+2) for odd layers."""
    rng = np.random.default_rng(seed)
    out: List[GateSpec] = []
    for layer in range(depth):
        for q in range(n):
            out.append(GateSpec(_lib.G_MAT, (q,), -1, 0.0, 1.0, haar_su2(rng)))
        start = 0 if layer % 2 == 0 else 1
        for a in range(start, n - 1, 2):
            out.append(GateSpec(_lib.G_MAT, (a, a + 1), -1, 0.0, 1.0, CZ))
    return out
