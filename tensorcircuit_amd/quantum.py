"""Pauli-string encodings (the reference's quantum.py:1025-1068) and the
measurement-record formats (quantum.py:2048-2440): sample_int / sample_bin /
count_vector / count_tuple / count_dict_bin / count_dict_int.  Draws come
from the native sampler (tcm_sample / tcm_sample_prob); the format
conversions are index arithmetic on device tensors."""
from __future__ import annotations

from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch


def ps2xyz(ps: Sequence[int]) -> Dict[str, List[int]]:
    """[1, 2, 2, 0] -> {"x": [0], "y": [1, 2], "z": []} (quantum.py:1025-1044)."""
    xyz: Dict[str, List[int]] = {"x": [], "y": [], "z": []}
    for i, j in enumerate(ps):
        if j == 1:
            xyz["x"].append(i)
        if j == 2:
            xyz["y"].append(i)
        if j == 3:
            xyz["z"].append(i)
    return xyz


def xyz2ps(xyz: Dict[str, List[int]], n: Optional[int] = None) -> List[int]:
    """Inverse of ps2xyz (quantum.py:1047-1068)."""
    if n is None:
        n = max(xyz.get("x", []) + xyz.get("y", []) + xyz.get("z", [])) + 1
    ps = [0 for _ in range(n)]
    for i in range(n):
        if i in xyz.get("x", []):
            ps[i] = 1
        elif i in xyz.get("y", []):
            ps[i] = 2
        elif i in xyz.get("z", []):
            ps[i] = 3
    return ps


# ---------------------------------------------------------------------------
# measurement records
# ---------------------------------------------------------------------------


def _n_of(d: int) -> int:
    return int(np.log(d) / np.log(2) + 1e-8)


def sample_int2bin(sample: Any, n: int) -> torch.Tensor:
    """[trials] ints in [0, 2^n) -> [trials, n] bits, qubit 0 first
    (quantum.py:2104-2119)."""
    s = torch.as_tensor(sample)
    shifts = torch.arange(n - 1, -1, -1, device=s.device)
    return torch.remainder(torch.bitwise_right_shift(s[..., None].to(torch.int64), shifts), 2)


def sample_bin2int(sample: Any, n: int) -> torch.Tensor:
    """[trials, n] bits -> [trials] ints (quantum.py:2122-2134)."""
    s = torch.as_tensor(sample)
    power = torch.as_tensor([2**j for j in reversed(range(n))], device=s.device, dtype=torch.int64)
    return torch.sum(s.to(torch.int64) * power, dim=-1)


def sample2count(sample: Any, n: int, jittable: bool = True) -> Tuple[torch.Tensor, torch.Tensor]:
    """sample_int -> count_tuple (distinct values, counts) (quantum.py:2137-2157);
    ``jittable`` pads to 2^n entries with value -1 as the reference's
    fixed-size unique does."""
    s = torch.as_tensor(sample).reshape(-1)
    vals, cnts = torch.unique(s, return_counts=True)
    if jittable:
        d = 2**n
        pad = d - vals.shape[0]
        vals = torch.cat([vals, torch.full((pad,), -1, dtype=vals.dtype, device=vals.device)])
        cnts = torch.cat([cnts, torch.zeros(pad, dtype=cnts.dtype, device=cnts.device)])
    return vals, cnts


def count_s2d(srepr: Tuple[Any, Any], n: int) -> torch.Tensor:
    """count_tuple -> count_vector of length 2^n (quantum.py:2048-2067)."""
    vals, cnts = (torch.as_tensor(x) for x in srepr)
    keep = vals >= 0
    out = torch.zeros(2**n, dtype=cnts.dtype, device=cnts.device)
    out.index_add_(0, vals[keep].to(torch.int64), cnts[keep])
    return out


def count_d2s(drepr: Any, eps: float = 1e-7) -> Tuple[torch.Tensor, torch.Tensor]:
    """count_vector -> count_tuple over entries above eps (quantum.py:2070-2101)."""
    d = torch.as_tensor(drepr)
    idx = torch.nonzero(torch.abs(d) > eps).reshape(-1)
    return idx, d[idx]


def count_vector2dict(count: Any, n: int, key: str = "bin") -> Dict[Any, Any]:
    """count_vector -> {int or bitstring: count} over all 2^n entries
    (quantum.py:2160-2185)."""
    c = torch.as_tensor(count).detach().cpu().numpy()
    d = {i: c[i].item() for i in range(2**n)}
    if key == "int":
        return d
    return {format(k, "0%db" % n): v for k, v in d.items()}


def count_tuple2dict(count: Tuple[Any, Any], n: int, key: str = "bin") -> Dict[Any, Any]:
    """count_tuple -> {int or bitstring: count} over the drawn values
    (quantum.py:2188-2215)."""
    vals = torch.as_tensor(count[0]).detach().cpu().numpy()
    cnts = torch.as_tensor(count[1]).detach().cpu().numpy()
    d = {int(i): j.item() for i, j in zip(vals, cnts) if i >= 0}
    if key == "int":
        return d
    return {format(k, "0%db" % n): v for k, v in d.items()}


def sample2all(sample: Any, n: int, format: str = "count_vector", jittable: bool = False) -> Any:
    """sample_int or sample_bin -> the requested format (quantum.py:2325-2373)."""
    s = torch.as_tensor(sample)
    if s.dim() == 1:
        sample_int, sample_bin = s, sample_int2bin(s, n)
    elif s.dim() == 2:
        sample_int, sample_bin = sample_bin2int(s, n), s
    else:
        raise ValueError("unrecognized tensor shape for sample")
    if format == "sample_int":
        return sample_int
    if format == "sample_bin":
        return sample_bin
    count_tuple = sample2count(sample_int, n, jittable)
    if format == "count_tuple":
        return count_tuple
    if format == "count_vector":
        return count_s2d(count_tuple, n)
    if format == "count_dict_bin":
        return count_tuple2dict(count_tuple, n, key="bin")
    if format == "count_dict_int":
        return count_tuple2dict(count_tuple, n, key="int")
    raise ValueError("unsupported format %s for finite shots measurement" % format)


def _uniforms(shape: Sequence[int], status: Any = None, random_generator: Any = None) -> torch.Tensor:
    """External randomness in [0, 1): ``status`` when given, else draws from
    ``random_generator`` (a torch.Generator) or torch's default generator."""
    if status is not None:
        return torch.as_tensor(status, dtype=torch.float64).reshape(tuple(shape))
    return torch.rand(tuple(shape), dtype=torch.float64, generator=random_generator)


def measurement_counts(state: Any, counts: Optional[int] = 8192, format: str = "count_vector",
                       is_prob: bool = False, random_generator: Any = None, status: Any = None,
                       jittable: bool = False) -> Any:
    """Measurement record of ``counts`` shots of a state (or probability
    vector with ``is_prob``) in one of the six formats (quantum.py:2218-2322);
    ``counts`` None or <= 0 gives the exact distribution."""
    from . import engine

    st = torch.as_tensor(state)
    if is_prob:
        pi = st.reshape(-1).to(torch.float64) if not st.dtype.is_floating_point else st.reshape(-1)
        pi = pi / torch.sum(pi)
    elif st.dim() == 2 and st.shape[0] == st.shape[1] and st.shape[0] > 1:
        rho = st / torch.trace(st)
        pi = torch.abs(torch.diagonal(rho))
        is_prob = True
    else:
        pi = None
    d = int(st.shape[0]) if is_prob else int(st.reshape(-1).shape[0])
    n = _n_of(d)
    if counts is None or counts <= 0:
        if pi is None:
            v = st.reshape(-1)
            v = v / torch.linalg.norm(v)
            pi = torch.real(torch.conj(v) * v)
        if format == "count_vector":
            return pi
        if format == "count_tuple":
            return count_d2s(pi)
        if format == "count_dict_bin":
            return count_vector2dict(pi, n, key="bin")
        if format == "count_dict_int":
            return count_vector2dict(pi, n, key="int")
        raise ValueError("unsupported format %s for analytical measurement" % format)
    u = _uniforms((1, counts), status, random_generator)
    if pi is not None:
        raw = engine.sample_prob_indices(pi.reshape(1, -1), n, u)[0]
    else:
        raw = engine.sample_indices(st.reshape(1, -1), n, u)[0]
    return sample2all(raw, n, format=format, jittable=jittable)


measurement_results = measurement_counts


def correlation_from_samples(index: Sequence[int], results: Any, n: int) -> torch.Tensor:
    """mean over shots of prod_{i in index} (1 - 2 b_i), from sample_int or
    sample_bin records (quantum.py:2400-2421)."""
    r = torch.as_tensor(results)
    if r.dim() == 1:
        r = sample_int2bin(r, n)
    r = 1 - 2 * r.to(torch.int64)
    prod = r[:, index[0]]
    for i in index[1:]:
        prod = prod * r[:, i]
    return torch.mean(prod.to(torch.float32))


def correlation_from_counts(index: Sequence[int], results: Any) -> torch.Tensor:
    """sum_i p_i prod_{j in index} (1 - 2 b_j(i)) over a probability vector
    (quantum.py:2424-2450)."""
    p = torch.as_tensor(results).reshape(-1)
    if p.is_complex():
        p = torch.real(p)
    p = p / torch.sum(p)
    n = _n_of(p.shape[0])
    i = torch.arange(p.shape[0], device=p.device)
    par = torch.zeros_like(i)
    for j in index:
        par = par ^ ((i >> (n - 1 - j)) & 1)
    sign = 1 - 2 * par
    return torch.sum(p * sign.to(p.dtype))


# ---------------------------------------------------------------------------
# Pauli-string Hamiltonians as matrices (quantum.py:1163-1485)
# ---------------------------------------------------------------------------


def _pauli_of(ls: Any, weight: Any = None) -> Tuple[Any, int]:
    from . import engine

    codes = np.asarray(torch.as_tensor(ls).detach().cpu().numpy() if isinstance(ls, torch.Tensor) else ls)
    codes = codes.astype(np.int8).reshape(len(codes), -1)
    n = codes.shape[1]
    if weight is None:
        w = np.ones(codes.shape[0])
    else:
        w = np.asarray(weight.detach().cpu().numpy() if isinstance(weight, torch.Tensor) else weight)
        w = np.real(w).astype(np.float64).reshape(-1)
    return engine.PauliSum(n, codes, w, jit=0), n


def PauliStringSum2COO(ls: Any, weight: Any = None, numpy: bool = False) -> Any:
    """sum_k w_k P_k as a sparse matrix (quantum.py:1304-1356): a coalesced
    torch sparse COO tensor on the device (entries from tcm_pauli_coo), or a
    scipy COO matrix with ``numpy``.  Rows of ``ls`` are Pauli codes
    0=I 1=X 2=Y 3=Z, qubit 0 most significant."""
    from . import cons, engine

    ham, _ = _pauli_of(ls, weight)
    m = engine.pauli_coo(ham, cons.dtypestr)
    if numpy:
        import scipy.sparse

        idx = m.indices().cpu().numpy()
        return scipy.sparse.coo_matrix((m.values().cpu().numpy(), (idx[0], idx[1])), shape=tuple(m.shape))
    return m


PauliStringSum2COO_numpy = lambda ls, weight=None: PauliStringSum2COO(ls, weight, numpy=True)  # noqa: E731


def PauliString2COO(l: Sequence[int], weight: Optional[float] = None) -> Any:
    """One Pauli string as a sparse matrix (quantum.py:1415-1458)."""
    return PauliStringSum2COO([list(l)], None if weight is None else [weight])


def PauliStringSum2Dense(ls: Any, weight: Any = None, numpy: bool = False) -> Any:
    """Dense 2^n x 2^n matrix of the sum (quantum.py:1257-1283)."""
    m = PauliStringSum2COO(ls, weight).to_dense()
    return m.cpu().numpy() if numpy else m


def heisenberg_hamiltonian(g: Any, hzz: float = 1.0, hxx: float = 1.0, hyy: float = 1.0, hz: float = 0.0,
                           hx: float = 0.0, hy: float = 0.0, sparse: bool = True, numpy: bool = False) -> Any:
    """XXZ-type Hamiltonian on the edges of graph g with fields on its nodes
    (quantum.py:1163-1254): for each edge ZZ, XX, YY terms (weights hzz,
    hxx, hyy), then for each node Z, X, Y terms (hz, hx, hy); zero weights
    are left out.  g needs ``nodes`` and ``edges`` (networkx or alike)."""
    nodes = list(g.nodes)
    n = len(nodes)
    ls, w = [], []
    for e in g.edges:
        for code, h in ((3, hzz), (1, hxx), (2, hyy)):
            if h != 0:
                r = [0] * n
                r[e[0]] = r[e[1]] = code
                ls.append(r)
                w.append(h)
    for node in nodes:
        for code, h in ((3, hz), (1, hx), (2, hy)):
            if h != 0:
                r = [0] * n
                r[node] = code
                ls.append(r)
                w.append(h)
    if sparse:
        return PauliStringSum2COO(ls, w, numpy=numpy)
    return PauliStringSum2Dense(ls, w, numpy=numpy)
