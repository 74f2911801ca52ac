"""Global configuration (the reference's cons.py:28-222 subset for this path):
the complex dtype and the backend handle."""
from __future__ import annotations

import contextlib
from typing import Iterator, Optional

dtypestr = "complex64"
rdtypestr = "float32"


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
