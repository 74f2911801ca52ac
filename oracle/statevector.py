"""complex128 numpy restatement of the reference's plain statevector method.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by tests,
``smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` -- never by the
library.

Conventions (all restated from the reference, tensorcircuit 0.12.1):

* qubit 0 is the most significant bit of the flat amplitude index
  (``circuit.py:803-810``; ``quantum.py:1437-1443``);
* a k-qubit gate tensor has legs ``[out_0..out_{k-1}, in_0..in_{k-1}]`` and
  ``index[0]`` is the most significant row/column bit of its matrix
  (``basecircuit.py:213-215``; ``backends/abstract_backend.py:390-400``);
* gates are applied one at a time by ``tensordot`` onto the ``[2]*n`` state,
  in insertion order, which is what ``plain_contractor`` computes
  (``cons.py:334-370``);
* expectations are unnormalised ``<psi|O|psi>`` and complex
  (``circuit.py:914-990``, ``basecircuit.py:267-319``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# gate matrices (reference: gates.py:32-127 constants, 579-636 rotations,
# 825-865 exp1/rzz, 339-368 adjoint, meta_vgate 949-982 controlled variants)
# ---------------------------------------------------------------------------

_I = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2.0)
_S = np.array([[1, 0], [0, 1j]], dtype=np.complex128)
_T = np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], dtype=np.complex128)
_WROOT = (1 / np.sqrt(2)) * np.array(
    [[1, -(1 + 1j) / np.sqrt(2)], [(1 - 1j) / np.sqrt(2), 1]], dtype=np.complex128
)
PAULI = {0: _I, 1: _X, 2: _Y, 3: _Z}


def _ctrl(u: np.ndarray) -> np.ndarray:
    """|0><0| (x) I + |1><1| (x) u  (reference ``GateF.controlled``, gates.py:291-307)."""
    s = u.shape[0]
    out = np.zeros((2 * s, 2 * s), dtype=np.complex128)
    out[:s, :s] = np.eye(s)
    out[s:, s:] = u
    return out


def _octrl(u: np.ndarray) -> np.ndarray:
    """|0><0| (x) u + |1><1| (x) I  (reference ``GateF.ocontrolled``, gates.py:309-326)."""
    s = u.shape[0]
    out = np.zeros((2 * s, 2 * s), dtype=np.complex128)
    out[:s, :s] = u
    out[s:, s:] = np.eye(s)
    return out


CONST_GATES: Dict[str, np.ndarray] = {
    "i": _I,
    "x": _X,
    "y": _Y,
    "z": _Z,
    "h": _H,
    "s": _S,
    "t": _T,
    "sd": _S.conj().T,
    "td": _T.conj().T,
    "wroot": _WROOT,
    "cnot": _ctrl(_X),
    "cz": _ctrl(_Z),
    "cy": _ctrl(_Y),
    "swap": np.array(
        [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128
    ),
    "ox": _octrl(_X),
    "oy": _octrl(_Y),
    "oz": _octrl(_Z),
    "toffoli": _ctrl(_ctrl(_X)),
    "fredkin": _ctrl(
        np.array(
            [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
            dtype=np.complex128,
        )
    ),
}
CONST_GATES["cx"] = CONST_GATES["cnot"]


def rot(pauli: np.ndarray, theta: float) -> np.ndarray:
    """exp(-i theta/2 P) = cos(theta/2) I - i sin(theta/2) P (gates.py:579-636)."""
    d = pauli.shape[0]
    return np.cos(theta / 2) * np.eye(d) - 1j * np.sin(theta / 2) * pauli


def exp1(u: np.ndarray, theta: float, half: bool = False) -> np.ndarray:
    """cos(theta) I - i sin(theta) U, theta halved when ``half`` (gates.py:825-858)."""
    if half:
        theta = theta / 2.0
    d = u.shape[0]
    return np.cos(theta) * np.eye(d) - 1j * np.sin(theta) * u


def expm_herm(u: np.ndarray, theta: float) -> np.ndarray:
    """exp(-i theta U) for Hermitian U by eigendecomposition (gates.py:792-822)."""
    w, v = np.linalg.eigh(u)
    return (v * np.exp(-1j * theta * w)) @ v.conj().T


def phase(theta: float) -> np.ndarray:
    """diag(1, e^{i theta}) (gates.py:460-478)."""
    return np.array([[1, 0], [0, np.exp(1j * theta)]], dtype=np.complex128)


def u_gate(theta: float, phi: float, lbd: float) -> np.ndarray:
    """OpenQASM3 U gate (gates.py:505-540)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array(
        [
            [c, -np.exp(1j * lbd) * s],
            [np.exp(1j * phi) * s, np.exp(1j * (phi + lbd)) * c],
        ],
        dtype=np.complex128,
    )


def r_gate(theta: float, alpha: float, phi: float) -> np.ndarray:
    """General rotation (gates.py:543-571)."""
    return (
        np.cos(theta) * _I
        - 1j * np.cos(phi) * np.sin(alpha) * np.sin(theta) * _X
        - 1j * np.sin(phi) * np.sin(alpha) * np.sin(theta) * _Y
        - 1j * np.sin(theta) * np.cos(alpha) * _Z
    )


def iswap(theta: float = 1.0) -> np.ndarray:
    """iSWAP(theta) (gates.py:675-705)."""
    c, s = np.cos(theta * np.pi / 2), np.sin(theta * np.pi / 2)
    return np.array(
        [[1, 0, 0, 0], [0, c, 1j * s, 0], [0, 1j * s, c, 0], [0, 0, 0, 1]],
        dtype=np.complex128,
    )


_ZZ = np.kron(_Z, _Z)
_XX = np.kron(_X, _X)
_YY = np.kron(_Y, _Y)


def gate_matrix(name: str, **kw) -> np.ndarray:
    """Matrix of a named gate (the reference's gate factories, gates.py)."""
    name = name.lower()
    if name in CONST_GATES:
        return CONST_GATES[name]
    th = kw.get("theta", 0.0)
    if name == "rx":
        return rot(_X, th)
    if name == "ry":
        return rot(_Y, th)
    if name == "rz":
        return rot(_Z, th)
    if name == "rzz":
        return exp1(_ZZ, th, half=True)
    if name == "rxx":
        return exp1(_XX, th, half=True)
    if name == "ryy":
        return exp1(_YY, th, half=True)
    if name == "phase":
        return phase(th)
    if name == "cphase":
        return _ctrl(phase(th))
    if name == "crx":
        return _ctrl(rot(_X, th))
    if name == "cry":
        return _ctrl(rot(_Y, th))
    if name == "crz":
        return _ctrl(rot(_Z, th))
    if name == "orx":
        return _octrl(rot(_X, th))
    if name == "ory":
        return _octrl(rot(_Y, th))
    if name == "orz":
        return _octrl(rot(_Z, th))
    if name == "iswap":
        return iswap(kw.get("theta", 1.0))
    if name == "u":
        return u_gate(th, kw.get("phi", 0.0), kw.get("lbd", 0.0))
    if name == "cu":
        return _ctrl(u_gate(th, kw.get("phi", 0.0), kw.get("lbd", 0.0)))
    if name == "r":
        return r_gate(th, kw.get("alpha", 0.0), kw.get("phi", 0.0))
    if name == "exp1":
        u = np.asarray(kw["unitary"], dtype=np.complex128)
        d = int(round(np.sqrt(u.size)))
        return exp1(u.reshape(d, d), th, half=kw.get("half", False))
    if name == "exp":
        u = np.asarray(kw["unitary"], dtype=np.complex128)
        d = int(round(np.sqrt(u.size)))
        return expm_herm(u.reshape(d, d), th)
    if name in ("any", "unitary"):
        u = np.asarray(kw["unitary"], dtype=np.complex128)
        d = int(round(np.sqrt(u.size)))
        return u.reshape(d, d)
    raise ValueError("oracle: unknown gate %r" % name)


# ---------------------------------------------------------------------------
# ops and programs
# ---------------------------------------------------------------------------


@dataclass
class Op:
    """One gate: ``name`` on ``qubits``; ``pidx`` names the parameter slot
    that feeds ``theta`` (theta = params[pidx]); other kwargs are fixed."""

    name: str
    qubits: Tuple[int, ...]
    pidx: Optional[int] = None
    kw: Dict = field(default_factory=dict)

    def matrix(self, params: Optional[np.ndarray] = None) -> np.ndarray:
        kw = dict(self.kw)
        if self.pidx is not None:
            kw["theta"] = float(np.asarray(params).reshape(-1)[self.pidx])
        return gate_matrix(self.name, **kw)


def zero_state(n: int) -> np.ndarray:
    s = np.zeros(2**n, dtype=np.complex128)
    s[0] = 1.0
    return s


def apply_gate(state: np.ndarray, u: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    """Contract a k-qubit gate into the state with ``tensordot`` -- one step of
    the reference's ``plain_contractor`` (cons.py:334-370) with the leg wiring
    of ``apply_general_gate`` (basecircuit.py:213-215)."""
    k = len(qubits)
    qubits = [q if q >= 0 else n + q for q in qubits]
    if len(set(qubits)) != k:
        raise ValueError("oracle: repeated qubit in %r" % (qubits,))
    psi = state.reshape([2] * n)
    g = np.asarray(u, dtype=np.complex128).reshape([2] * (2 * k))
    out = np.tensordot(g, psi, axes=(list(range(k, 2 * k)), qubits))
    # out legs: gate outs first, then the untouched state legs in order
    rest = [q for q in range(n) if q not in qubits]
    perm = [0] * n
    for i, q in enumerate(qubits):
        perm[q] = i
    for j, q in enumerate(rest):
        perm[q] = k + j
    return np.transpose(out, perm).reshape(-1)


def apply_gate_bruteforce(state: np.ndarray, u: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    """Independent O(4^n) restatement used only to pin ``apply_gate``:
    (U psi)[y] = sum_x U[y_sub, x_sub] psi[x] over x equal to y off ``qubits``."""
    k = len(qubits)
    N = 2**n
    out = np.zeros(N, dtype=np.complex128)
    bits = [n - 1 - q for q in qubits]
    for y in range(N):
        ysub = 0
        for b in bits:
            ysub = (ysub << 1) | ((y >> b) & 1)
        acc = 0j
        for xsub in range(2**k):
            x = y
            for i, b in enumerate(bits):
                bit = (xsub >> (k - 1 - i)) & 1
                x = (x & ~(1 << b)) | (bit << b)
            acc += u[ysub, xsub] * state[x]
        out[y] = acc
    return out


def run_ops(n: int, ops: Sequence[Op], params: Optional[np.ndarray] = None,
            inputs: Optional[np.ndarray] = None) -> np.ndarray:
    """Final state of a gate list from |0..0> (or ``inputs``)."""
    s = zero_state(n) if inputs is None else np.asarray(inputs, dtype=np.complex128).reshape(-1).copy()
    for op in ops:
        s = apply_gate(s, op.matrix(params), op.qubits, n)
    return s


# ---------------------------------------------------------------------------
# expectations and Pauli sums
# (abstractcircuit.py:1208-1288, basecircuit.py:267-319, quantum.py:1025-1068,
#  quantum.py:1415-1485)
# ---------------------------------------------------------------------------


def expectation(state: np.ndarray, n: int, *ops: Tuple[np.ndarray, Sequence[int]]) -> complex:
    """Unnormalised <psi| O_1 ... O_m |psi>; a repeated index raises ValueError
    like basecircuit.py:306-307."""
    occupied = set()
    phi = state
    for m, idx in ops:
        if isinstance(idx, int):
            idx = [idx]
        idx = [i if i >= 0 else n + i for i in idx]
        for i in idx:
            if i in occupied:
                raise ValueError("Cannot measure two operators in one index")
            occupied.add(i)
        phi = apply_gate(phi, np.asarray(m, dtype=np.complex128).reshape(2 ** len(idx), -1), idx, n)
    return complex(np.vdot(state, phi))


def ps2xyz(ps: Sequence[int]) -> Dict[str, List[int]]:
    """quantum.py:1025-1044."""
    xyz: Dict[str, List[int]] = {"x": [], "y": [], "z": []}
    for i, j in enumerate(ps):
        if j == 1:
            xyz["x"].append(i)
        if j == 2:
            xyz["y"].append(i)
        if j == 3:
            xyz["z"].append(i)
    return xyz


def expectation_ps(state: np.ndarray, n: int, x=None, y=None, z=None, ps=None) -> complex:
    """abstractcircuit.py:1208-1288: ``ps`` (codes 0=I,1=X,2=Y,3=Z) overrides x/y/z."""
    if ps is not None:
        d = ps2xyz(ps)
        x, y, z = d["x"], d["y"], d["z"]
    obs = []
    for i in x or []:
        obs.append((_X, [i]))
    for i in y or []:
        obs.append((_Y, [i]))
    for i in z or []:
        obs.append((_Z, [i]))
    return expectation(state, n, *obs)


def pauli_string_matrix(codes: Sequence[int]) -> np.ndarray:
    """Dense kron of single-qubit Paulis, qubit 0 leftmost (pins bit order like
    tests/test_miscs.py:24-55)."""
    m = np.ones((1, 1), dtype=np.complex128)
    for c in codes:
        m = np.kron(m, PAULI[int(c)])
    return m


def apply_pauli_string(state: np.ndarray, n: int, codes: Sequence[int]) -> np.ndarray:
    """P|psi> by the index rule of ps2coo_core (quantum.py:1461-1485):
    (P psi)[y] = (-i)^{ny} (-1)^{popcount(y & (ymask|zmask))} psi[y ^ flip]."""
    flip = ymask = zmask = 0
    ny = 0
    for q, c in enumerate(codes):
        b = 1 << (n - 1 - q)
        if c == 1:
            flip |= b
        elif c == 2:
            flip |= b
            ymask |= b
            ny += 1
        elif c == 3:
            zmask |= b
    idx = np.arange(2**n, dtype=np.int64)
    par = np.zeros(2**n, dtype=np.int64)
    t = idx & (ymask | zmask)
    for i in range(n):
        par ^= (t >> i) & 1
    sign = 1 - 2 * par
    return ((-1j) ** (ny % 4)) * sign * state[idx ^ flip]


def pauli_sum_apply(state: np.ndarray, n: int, terms: Sequence[Tuple[float, Sequence[int]]]) -> np.ndarray:
    out = np.zeros_like(state)
    for w, codes in terms:
        out += w * apply_pauli_string(state, n, codes)
    return out


def pauli_sum_expectation(state: np.ndarray, n: int, terms) -> complex:
    return complex(np.vdot(state, pauli_sum_apply(state, n, terms)))


# ---------------------------------------------------------------------------
# workloads: HEA + TFIM (the idiom of tests/test_backends.py:426-449 and
# benchmarks/scripts/vqe_tc.py:75-81), brickwork of Haar SU(2) + CZ
# ---------------------------------------------------------------------------


def hea_ops(n: int, nlayers: int) -> List[Op]:
    """Per layer: ry on every qubit, rz on every qubit, CNOT chain i->i+1.
    params laid out [2L, n]: row 2j = ry angles of layer j, row 2j+1 = rz."""
    ops: List[Op] = []
    for j in range(nlayers):
        for i in range(n):
            ops.append(Op("ry", (i,), pidx=(2 * j) * n + i))
        for i in range(n):
            ops.append(Op("rz", (i,), pidx=(2 * j + 1) * n + i))
        for i in range(n - 1):
            ops.append(Op("cnot", (i, i + 1)))
    return ops


def tfim_terms(n: int, h: float = 1.0, j: float = 1.0) -> List[Tuple[float, List[int]]]:
    """Open-chain H = -j sum Z_i Z_{i+1} - h sum X_i as (weight, codes)."""
    terms = []
    for i in range(n - 1):
        c = [0] * n
        c[i] = 3
        c[i + 1] = 3
        terms.append((-j, c))
    for i in range(n):
        c = [0] * n
        c[i] = 1
        terms.append((-h, c))
    return terms


def haar_su2(rng: np.random.Generator) -> np.ndarray:
    """Haar SU(2) by QR of a complex Ginibre matrix, phases fixed by sign(diag R),
    then divided by sqrt(det) (Mezzadri, arXiv:math-ph/0609050)."""
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))
    return q / np.sqrt(np.linalg.det(q))


def brickwork_ops(n: int, depth: int, seed: int) -> List[Op]:
    rng = np.random.default_rng(seed)
    ops: List[Op] = []
    for layer in range(depth):
        for q in range(n):
            ops.append(Op("unitary", (q,), kw={"unitary": haar_su2(rng)}))
        start = 0 if layer % 2 == 0 else 1
        for a in range(start, n - 1, 2):
            ops.append(Op("cz", (a, a + 1)))
    return ops


# ---------------------------------------------------------------------------
# gradients
# ---------------------------------------------------------------------------

_GEN = {
    "rx": (_X, 0.5),
    "ry": (_Y, 0.5),
    "rz": (_Z, 0.5),
    "rzz": (_ZZ, 0.5),
    "rxx": (_XX, 0.5),
    "ryy": (_YY, 0.5),
    "phase": (np.array([[0, 0], [0, -1]], dtype=np.complex128), 1.0),
}


def _dmatrix(op: Op, params: np.ndarray) -> np.ndarray:
    """dU/dtheta.  Rotation-type gates use U = exp(-i s theta G) so
    dU/dtheta = -i s G U; others use a 5-point stencil (error ~h^4)."""
    if op.name in _GEN:
        g, s = _GEN[op.name]
        return -1j * s * g @ op.matrix(params)
    if op.name == "exp1":
        u = np.asarray(op.kw["unitary"], dtype=np.complex128)
        d = int(round(np.sqrt(u.size)))
        s = 0.5 if op.kw.get("half", False) else 1.0
        return -1j * s * u.reshape(d, d) @ op.matrix(params)
    h = 1e-3
    p = np.array(params, dtype=np.float64).reshape(-1)

    def m(dx):
        q = p.copy()
        q[op.pidx] += dx
        return op.matrix(q)

    return (-m(2 * h) + 8 * m(h) - 8 * m(-h) + m(-2 * h)) / (12 * h)


def energy(n: int, ops: Sequence[Op], params: np.ndarray, terms, inputs=None) -> float:
    s = run_ops(n, ops, params, inputs)
    return float(np.real(pauli_sum_expectation(s, n, terms)))


def adjoint_value_and_grad(n: int, ops: Sequence[Op], params: np.ndarray, terms,
                           inputs=None) -> Tuple[float, np.ndarray]:
    """E = <psi|H|psi> and dE/dparams by the reverse sweep
    lambda <- U^dag lambda, psi <- U^dag psi, g += 2 Re <lambda| dU psi_prev>
    (Jones & Gacon, arXiv:2009.02823), all in complex128."""
    params = np.asarray(params, dtype=np.float64).reshape(-1)
    psi = run_ops(n, ops, params, inputs)
    lam = pauli_sum_apply(psi, n, terms)
    e = float(np.real(np.vdot(psi, lam)))
    grad = np.zeros_like(params)
    for op in reversed(ops):
        u = op.matrix(params)
        psi_prev = apply_gate(psi, u.conj().T, op.qubits, n)
        if op.pidx is not None:
            du = _dmatrix(op, params)
            grad[op.pidx] += 2.0 * np.real(np.vdot(lam, apply_gate(psi_prev, du, op.qubits, n)))
        lam = apply_gate(lam, u.conj().T, op.qubits, n)
        psi = psi_prev
    return e, grad


def finite_difference_grad(n, ops, params, terms, eps=1e-3, inputs=None) -> np.ndarray:
    """Central differences in complex128."""
    p = np.asarray(params, dtype=np.float64).reshape(-1)
    g = np.zeros_like(p)
    for i in range(p.size):
        a = p.copy()
        a[i] += eps
        b = p.copy()
        b[i] -= eps
        g[i] = (energy(n, ops, a, terms, inputs) - energy(n, ops, b, terms, inputs)) / (2 * eps)
    return g


def parameter_shift_grad(n, ops, params, terms, inputs=None) -> np.ndarray:
    """dE/dtheta = [E(theta+pi/2) - E(theta-pi/2)] / 2 for exp(-i theta/2 P)
    gates (experimental.py:215-269, shifts (pi/2, 2) at line 219).  Valid when
    every use of a parameter is such a rotation."""
    p = np.asarray(params, dtype=np.float64).reshape(-1)
    g = np.zeros_like(p)
    for i in range(p.size):
        a = p.copy()
        a[i] += np.pi / 2
        b = p.copy()
        b[i] -= np.pi / 2
        g[i] = (energy(n, ops, a, terms, inputs) - energy(n, ops, b, terms, inputs)) / 2
    return g


def tfim_ground_energy(n: int, h: float = 1.0) -> float:
    """Exact diagonalisation of the open-chain TFIM (dense, n <= 12)."""
    dim = 2**n
    hm = np.zeros((dim, dim), dtype=np.complex128)
    for w, codes in tfim_terms(n, h):
        hm += w * pauli_string_matrix(codes)
    return float(np.linalg.eigvalsh(hm)[0])
