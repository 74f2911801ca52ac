"""Measurement shortcuts (the reference's templates/measurements.py) on the
device Pauli kernels: every structure of a call is evaluated from one
launch (tcm_expect_pauli_terms) instead of one contraction per structure.

Structure tables follow the reference (templates/measurements.py:18-153):
with ``onehot=True`` a structure is an integer code per wire
(0=I, 1=X, 2=Y, 3=Z), shape [n]; with ``onehot=False`` it is a coefficient
table over [I, X, Y, Z] per wire, shape [n, 4], and wire i carries the
operator sum_k s[i, k] P_k.  A leading axis stacks several structures (the
reference's vmap over structures); ``torch.vmap`` over the structures
argument is supported too.
"""
from __future__ import annotations

import collections
import itertools
import threading
from typing import Any, List, Tuple

import numpy as np
import torch

from .. import engine

# most Pauli strings one structure may expand into (a coefficient table with
# several non-zero entries on many wires multiplies out exponentially)
MAX_STRINGS_PER_STRUCTURE = 4096


def _host_table(structures: Any) -> np.ndarray:
    """The table's values on the host.  Read with torch.func's transforms
    off, so that a plain tensor used inside grad / vmap is not wrapped by the
    read itself; a tensor batched by vmap raises RuntimeError (no storage)."""
    if isinstance(structures, torch.Tensor):
        if structures.requires_grad:
            raise ValueError("structure tables are data here: gradients with respect to them are not supported")
        with torch._C._DisableFuncTorch():
            return structures.detach().cpu().numpy()
    return np.asarray(structures)


def _wire_factors(s: np.ndarray, onehot: bool, n: int) -> List[List[Tuple[int, complex]]]:
    """Per wire, the (Pauli code, coefficient) pairs of its operator."""
    if onehot:
        codes = np.asarray(s).reshape(-1)
        if codes.shape[0] != n:
            raise ValueError("onehot structure of length %d for %d qubits" % (codes.shape[0], n))
        out = []
        for c in codes:
            c = int(c)
            if c < 0 or c > 3:
                raise ValueError("Pauli code %d not in 0..3" % c)
            out.append([(c, 1.0)])
        return out
    tab = np.asarray(s).reshape(-1, 4)
    if tab.shape[0] != n:
        raise ValueError("structure table of shape %s for %d qubits (expected [%d, 4])" % (np.shape(s), n, n))
    return [[(k, complex(tab[i, k])) for k in range(4) if tab[i, k] != 0] for i in range(n)]


def _strings(factors: List[List[Tuple[int, complex]]]) -> List[Tuple[Tuple[int, ...], complex]]:
    """Expand prod_i (sum_k c_ik P_k) into (codes, weight) Pauli strings."""
    count = 1
    for f in factors:
        count *= len(f)
    if count > MAX_STRINGS_PER_STRUCTURE:
        raise ValueError("structure expands into %d Pauli strings (limit %d)" % (count, MAX_STRINGS_PER_STRUCTURE))
    out = []
    for combo in itertools.product(*factors):
        w = complex(1.0)
        for _, c in combo:
            w *= c
        out.append((tuple(k for k, _ in combo), w))
    return out


class _Lowered:
    """One Pauli sum over the distinct strings of a set of outputs, and the
    real matrix W [outputs, strings] with output_s = sum_t W[s, t] <P_t>
    (<P_t> is real for a Hermitian string, so only Re of each string's
    weight contributes to the real expectation)."""

    def __init__(self, n: int, rows: List[List[Tuple[Tuple[int, ...], complex]]]):
        index = {}
        for row in rows:
            for codes, _ in row:
                index.setdefault(codes, len(index))
        W = np.zeros((len(rows), max(1, len(index))))
        for s, row in enumerate(rows):
            for codes, w in row:
                W[s, index[codes]] += w.real
        codes = np.zeros((max(1, len(index)), n), dtype=np.int8)
        for c, t in index.items():
            codes[t] = c
        self.n = n
        self.ham = engine.PauliSum(n, codes, np.ones(codes.shape[0]), jit=0)
        self.W = W
        self._wdev = {}

    def w_dev(self, dev: torch.device) -> torch.Tensor:
        w = self._wdev.get(dev)
        if w is None:
            w = self._wdev[dev] = torch.as_tensor(self.W, dtype=torch.float64).to(dev)
        return w


_LOWERED: "collections.OrderedDict[Tuple, _Lowered]" = collections.OrderedDict()
_LOWER_LOCK = threading.Lock()


def _lowered(tab: np.ndarray, onehot: bool, local: bool, n: int) -> Tuple[_Lowered, Tuple[int, ...]]:
    """Lower a structure table (or a stack of them) to a Pauli sum and its
    output matrix; returns it with the output shape."""
    single_ndim = 1 if onehot else 2
    if tab.ndim not in (single_ndim, single_ndim + 1):
        raise ValueError("structures of shape %s: expected [n]%s or a stack of them"
                         % (tab.shape, "" if onehot else "[4]"))
    stacked = tab.ndim == single_ndim + 1
    key = (tab.tobytes(), str(tab.dtype), tab.shape, onehot, local, n)
    with _LOWER_LOCK:
        hit = _LOWERED.get(key)
        if hit is not None:
            _LOWERED.move_to_end(key)
    structs = list(tab) if stacked else [tab]
    if local:
        shape = (len(structs), n) if stacked else (n,)
    else:
        shape = (len(structs),) if stacked else ()
    if hit is not None:
        return hit, shape
    rows = []
    for s in structs:
        factors = _wire_factors(s, onehot, n)
        if local:
            # wire i alone: its factor on i, identity elsewhere
            for i in range(n):
                rows.append([(tuple(k if j == i else 0 for j in range(n)), c) for k, c in factors[i]])
        else:
            rows.append(_strings(factors))
    low = _Lowered(n, rows)
    with _LOWER_LOCK:
        _LOWERED[key] = low
        while len(_LOWERED) > 64:
            _LOWERED.popitem(last=False)
    return low, shape


class _MeasureFn(torch.autograd.Function):
    """psi [N], structure table(s) -> real expectations; one launch of the
    per-term kernel, backward one weighted application
    (dL/dpsi = sum_t 2 (sum_s g_s W[s, t]) P_t psi).  ``low`` is the
    lowered table, computed by the caller from the table's values (inside
    torch.func transforms the Function only sees wrapped tensors); it is
    None only under vmap over the structures, whose rule lowers the
    physical stack."""

    @staticmethod
    def forward(st, structures, low, shape, onehot, local):
        e = low.ham.expect_terms(st.reshape(1, -1))[0]
        vals = low.w_dev(st.device) @ torch.real(e)
        return vals.to(st.real.dtype).reshape(shape)

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, _, low, _, _, _ = inputs
        ctx.low = low
        ctx.save_for_backward(st)

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        w = g.reshape(-1).to(torch.float64) @ ctx.low.w_dev(st.device)
        return engine.PauliTermsVJPFn.apply(st, w, ctx.low.ham), None, None, None, None, None

    @staticmethod
    def vmap(info, in_dims, st, structures, low, shape, onehot, local):
        B = info.batch_size
        n = (st.shape[-1] - 1).bit_length()
        if in_dims[0] is None:
            # only the structures are batched: one launch over the stack
            s = structures.movedim(in_dims[1], 0)
            low_s, shape_s = _lowered(_host_table(s), onehot, local, n)
            return _MeasureFn.apply(st, s, low_s, shape_s, onehot, local), 0
        sts = st.movedim(in_dims[0], 0)
        outs = []
        for b in range(B):
            sb = structures if in_dims[1] is None else structures.select(in_dims[1], b)
            low_b, shape_b = (low, shape) if in_dims[1] is None else _lowered(_host_table(sb), onehot, local, n)
            outs.append(_MeasureFn.apply(sts[b], sb, low_b, shape_b, onehot, local))
        return torch.stack(outs), 0


def _measure(c: Any, structures: Any, onehot: bool, local: bool, reuse: bool) -> torch.Tensor:
    s = structures if isinstance(structures, torch.Tensor) else torch.as_tensor(np.asarray(structures))
    try:
        tab = _host_table(s)
    except RuntimeError:  # a tensor batched by vmap: lowered by the vmap rule
        tab = None
    low, shape = (None, None) if tab is None else _lowered(tab, bool(onehot), local, c._nstate)
    return _MeasureFn.apply(c._state(reuse).reshape(-1), s, low, shape, bool(onehot), local)


def any_measurements(c: Any, structures: Any, onehot: bool = False, reuse: bool = False) -> Any:
    """Re <psi| (x)_i O_i |psi> with O_i = sum_k s[i, k] P_k
    (templates/measurements.py:18-83); a stack of structures returns one
    value per structure, all from one launch."""
    return _measure(c, structures, onehot, False, reuse)


parameterized_measurements = any_measurements


def any_local_measurements(c: Any, structures: Any, onehot: bool = False, reuse: bool = True) -> Any:
    """[Re <psi|O_i|psi> for each wire i], O_i = sum_k s[i, k] P_k acting on
    wire i alone (templates/measurements.py:86-153)."""
    return _measure(c, structures, onehot, True, reuse)


parameterized_local_measurements = any_local_measurements


def pauli_sum_expectation(c: Any, structures: Any, weights: Any = None, reuse: bool = True) -> Any:
    """sum_k w_k <P_k> for integer Pauli codes [nterms, n] as a complex
    scalar, one fused fp64 reduction; the gradient flows through one H|psi>
    application."""
    n = c._nstate
    codes = np.asarray(_host_table(structures)).reshape(-1, n).astype(np.int8)
    w = np.ones(codes.shape[0]) if weights is None else np.asarray(_host_table(weights), dtype=np.float64)
    return c._pauli_expectation(codes, w.reshape(-1), reuse)


def operator_expectation(c: Any, h: Any) -> Any:
    """Real <psi|H|psi> for H a PauliSum (fused Pauli kernel), a sparse
    matrix (CSR kernel) or a dense matrix (templates/measurements.py:156-174)."""
    if isinstance(h, engine.PauliSum):
        return torch.real(engine.PauliExpFn.apply(c._state(), h))
    if isinstance(h, engine.CsrMatrix) or (isinstance(h, torch.Tensor) and h.layout in (torch.sparse_coo,
                                                                                           torch.sparse_csr)):
        return sparse_expectation(c, h)
    w = c._state().reshape(-1, 1)
    hd = torch.as_tensor(h).to(device=w.device, dtype=w.dtype)
    return torch.real((torch.conj(w).T @ (hd @ w))[0, 0])


def sparse_expectation(c: Any, hamiltonian: Any) -> Any:
    """Real <psi|H|psi> for a sparse H (COO or CSR, e.g. from
    quantum.PauliStringSum2COO) by the CSR kernel with fp64 accumulation,
    differentiable in the circuit's parameters (templates/measurements.py:
    177-191)."""
    return torch.real(engine.csr_expectation(hamiltonian, c._state(), c._nstate))
