"""Ansatz blocks (the reference's templates/blocks.py) and the HEA layer used
by the statevector benchmarks."""
from __future__ import annotations

from typing import Any

import numpy as np

from .. import gates as G


def example_block(c: Any, param: Any, nlayers: int = 2, is_split: bool = False) -> Any:
    """H on every qubit, then ``nlayers`` of exp1(ZZ) ladders and rx
    (templates/blocks.py:113-152); ``param`` reshaped to [2*nlayers, n]."""
    n = c._nqubits
    param = param.reshape(2 * nlayers, n) if hasattr(param, "reshape") else np.reshape(param, (2 * nlayers, n))
    for i in range(n):
        c.H(i)
    for j in range(nlayers):
        for i in range(n - 1):
            c.exp1(i, i + 1, unitary=G._zz_matrix, theta=param[2 * j, i])
        for i in range(n):
            c.rx(i, theta=param[2 * j + 1, i])
    return c


def hea_layers(c: Any, param: Any, nlayers: int, barrier: bool = True) -> Any:
    """Per layer: ry on every qubit, rz on every qubit, CNOT chain i -> i+1;
    ``param`` laid out [2*nlayers, n] (row 2j = ry of layer j, row 2j+1 = rz),
    the indexing idiom of tests/test_backends.py:426-442."""
    n = c._nqubits
    for j in range(nlayers):
        for i in range(n):
            c.ry(i, theta=param[2 * j, i])
        for i in range(n):
            c.rz(i, theta=param[2 * j + 1, i])
        for i in range(n - 1):
            c.cnot(i, i + 1)
        if barrier:
            c.barrier_instruction()
    return c
