"""Global configuration (the reference's cons.py:28-222 subset for this path):
the complex dtype, the backend handle and the pass-compiler policy."""
from __future__ import annotations

import contextlib
import os
from typing import Any, Iterator, Optional

dtypestr = "complex64"
rdtypestr = "float32"

# Pass compiler policy for programs built by the Circuit API: None = compile
# the tile passes of states with at least JIT_MIN_QUBITS qubits (their
# passes run for milliseconds, so the one-off compile is amortised and
# cached on disk), True / False = always / never.  TCM_JIT=0/1 sets it.
jit: Optional[bool] = {"0": False, "1": True}.get(os.environ.get("TCM_JIT", ""), None)
JIT_MIN_QUBITS = 20


# Constant circuits on the FP32 matrix cores: None = automatic (complex64,
# at least MFMA_MIN_QUBITS qubits, no tensor-valued parameter: gates fused
# into dense 5-qubit blocks applied as real GEMMs), True / False = always
# (when the circuit qualifies) / never.  TCM_MFMA=0/1 sets it.
mfma: Optional[bool] = {"0": False, "1": True}.get(os.environ.get("TCM_MFMA", ""), None)
MFMA_MIN_QUBITS = 20


def set_mfma(flag: Optional[bool]) -> None:
    """Choose the matrix-core policy for constant circuits (None = by size)."""
    global mfma
    mfma = flag


def mfma_for(nqubits: int) -> bool:
    if dtypestr != "complex64" or mfma is False:
        return False
    return bool(mfma) or nqubits >= MFMA_MIN_QUBITS


def set_jit(flag: Optional[bool]) -> None:
    """Choose the pass-compiler policy (None = by state size)."""
    global jit
    jit = flag


def jit_for(nqubits: int) -> int:
    if jit is None:
        return 1 if nqubits >= JIT_MIN_QUBITS else 0
    return 1 if jit else 0


def set_dtype(dtype: Optional[str] = None) -> str:
    """Switch the module-wide precision: "complex64" or "complex128"
    (cons.py:129-182, also accepting "float32"/"float64" and "32"/"64")."""
    global dtypestr, rdtypestr
    if dtype is None:
        dtype = "complex64"
    d = str(dtype)
    if d in ("complex64", "float32", "32"):
        dtypestr, rdtypestr = "complex64", "float32"
    elif d in ("complex128", "float64", "64"):
        dtypestr, rdtypestr = "complex128", "float64"
    else:
        raise ValueError("unsupported dtype %r" % dtype)
    return dtypestr


@contextlib.contextmanager
def runtime_dtype(dtype: str) -> Iterator[str]:
    """cons.py:185-222."""
    old = dtypestr
    set_dtype(dtype)
    try:
        yield dtypestr
    finally:
        set_dtype(old)


def set_backend(backend: Optional[str] = None):
    """Only the pytorch backend exists in this build (the reference's
    backend registry, backends/backend_factory.py:27-59)."""
    from .backend import backend as _b

    if backend not in (None, "pytorch", "torch"):
        raise ValueError("Backend '%s' doesn't exist in this build (only 'pytorch')" % backend)
    return _b


get_backend = set_backend

_CONTRACTOR = {"method": "greedy", "optimizer": None}


def set_contractor(method: Optional[str] = None, optimizer: Any = None, memory_limit: Any = None,
                   opt_conf: Any = None, set_global: bool = True, contraction_info: bool = False,
                   debug_level: int = 0, preprocessing: bool = False) -> Any:
    """Accepted for compatibility with the reference's contractor hook
    (cons.py:732-830); the statevector engine applies gates in program
    order through fused passes and needs no contraction path, so the choice
    is recorded and has no effect.  Returns the recorded setting."""
    if set_global:
        _CONTRACTOR.update({"method": method or "greedy", "optimizer": optimizer})
    return dict(_CONTRACTOR)


def get_contractor(method: Optional[str] = None, optimizer: Any = None, memory_limit: Any = None,
                   opt_conf: Any = None, contraction_info: bool = False, debug_level: int = 0,
                   preprocessing: bool = False) -> Any:
    return set_contractor(method, optimizer, memory_limit, opt_conf, False, contraction_info, debug_level,
                          preprocessing)
