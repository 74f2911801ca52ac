"""Gate definitions of the reference (tensorcircuit/gates.py) as host-side
matrices plus the lowering information the HIP engine needs.

Matrices follow the reference exactly: row-major, ``index[0]`` most
significant (gates.py:32-127 constants, 579-636 rotations, 825-865
exp1/rzz/rxx/ryy, 291-326 controlled / ocontrolled variants, 949-982
registration).  Angles are applied by the engine as
``exp(-i a/2 P)`` for rx/ry/rz/rxx/ryy/rzz, ``diag(1, e^{ia})`` for phase and
``cos(a) I - i sin(a) G`` for exp1 (halved when ``half``).
"""
from __future__ import annotations

from typing import Any, Dict, Optional, Sequence

import numpy as np

from . import _lib

_i_matrix = np.array([[1.0, 0.0], [0.0, 1.0]])
_x_matrix = np.array([[0.0, 1.0], [1.0, 0.0]])
_y_matrix = np.array([[0.0, -1j], [1j, 0.0]])
_z_matrix = np.array([[1.0, 0.0], [0.0, -1.0]])
_h_matrix = 1 / np.sqrt(2) * np.array([[1.0, 1.0], [1.0, -1.0]])
_s_matrix = np.array([[1.0, 0.0], [0.0, 1j]])
_t_matrix = np.array([[1.0, 0.0], [0.0, np.exp(np.pi / 4 * 1j)]])
_wroot_matrix = (1 / np.sqrt(2)) * np.array(
    [[1, -1 / np.sqrt(2) * (1 + 1.0j)], [1 / np.sqrt(2) * (1 - 1.0j), 1]])
_xx_matrix = np.kron(_x_matrix, _x_matrix)
_yy_matrix = np.kron(_y_matrix, _y_matrix)
_zz_matrix = np.kron(_z_matrix, _z_matrix)
_cnot_matrix = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0], [0, 0, 1.0, 0]])
_cz_matrix = np.diag([1.0, 1.0, 1.0, -1.0])
_cy_matrix = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]])
_swap_matrix = np.array([[1.0, 0, 0, 0], [0, 0, 1.0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0]])


def _controlled(u: np.ndarray) -> np.ndarray:
    """[[I, 0], [0, u]] (GateF.controlled, gates.py:291-307)."""
    s = u.shape[0]
    out = np.zeros((2 * s, 2 * s), dtype=np.complex128)
    out[:s, :s] = np.eye(s)
    out[s:, s:] = u
    return out


def _ocontrolled(u: np.ndarray) -> np.ndarray:
    """[[u, 0], [0, I]] (GateF.ocontrolled, gates.py:309-326)."""
    s = u.shape[0]
    out = np.zeros((2 * s, 2 * s), dtype=np.complex128)
    out[:s, :s] = u
    out[s:, s:] = np.eye(s)
    return out


_toffoli_matrix = _controlled(_cnot_matrix)
_fredkin_matrix = _controlled(_swap_matrix)

CONST = {
    "i": _i_matrix, "x": _x_matrix, "y": _y_matrix, "z": _z_matrix, "h": _h_matrix,
    "t": _t_matrix, "s": _s_matrix, "td": _t_matrix.conj().T, "sd": _s_matrix.conj().T,
    "wroot": _wroot_matrix, "cnot": _cnot_matrix, "cz": _cz_matrix, "swap": _swap_matrix,
    "cy": _cy_matrix, "ox": _ocontrolled(_x_matrix), "oy": _ocontrolled(_y_matrix),
    "oz": _ocontrolled(_z_matrix), "toffoli": _toffoli_matrix, "fredkin": _fredkin_matrix,
}

# the reference's registration lists (abstractcircuit.py:28-66)
sgates = (["i", "x", "y", "z", "h", "t", "s", "td", "sd", "wroot"]
          + ["cnot", "cz", "swap", "cy", "ox", "oy", "oz"] + ["toffoli", "fredkin"])
vgates = ["r", "cr", "u", "cu", "rx", "ry", "rz", "phase", "rxx", "ryy", "rzz", "cphase",
          "crx", "cry", "crz", "orx", "ory", "orz", "iswap", "any", "exp", "exp1"]
gate_aliases = [["cnot", "cx"], ["fredkin", "cswap"], ["toffoli", "ccnot"], ["toffoli", "ccx"],
                ["any", "unitary"], ["sd", "sdg"], ["td", "tdg"]]

# trainable gates the engine differentiates: name -> (tcm kind, scale)
TRAINABLE = {
    "rx": (_lib.G_RX, 1.0), "ry": (_lib.G_RY, 1.0), "rz": (_lib.G_RZ, 1.0),
    "rxx": (_lib.G_RXX, 1.0), "ryy": (_lib.G_RYY, 1.0), "rzz": (_lib.G_RZZ, 1.0),
    "phase": (_lib.G_PHASE, 1.0),
}


# --- host matrices of the remaining parametrised gates (constant per call) ---

def _rot(p: np.ndarray, th: float) -> np.ndarray:
    return np.cos(th / 2) * np.eye(p.shape[0]) - 1j * np.sin(th / 2) * p


def rx_matrix(theta: float = 0.0) -> np.ndarray:
    return _rot(_x_matrix, theta)


def ry_matrix(theta: float = 0.0) -> np.ndarray:
    return _rot(_y_matrix, theta)


def rz_matrix(theta: float = 0.0) -> np.ndarray:
    return _rot(_z_matrix, theta)


def phase_matrix(theta: float = 0.0) -> np.ndarray:
    return np.array([[1, 0], [0, np.exp(1j * theta)]])


def u_matrix(theta: float = 0.0, phi: float = 0.0, lbd: float = 0.0) -> np.ndarray:
    """OpenQASM3 U (gates.py:505-540)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -np.exp(1j * lbd) * s], [np.exp(1j * phi) * s, np.exp(1j * (phi + lbd)) * c]])


def r_matrix(theta: float = 0.0, alpha: float = 0.0, phi: float = 0.0) -> np.ndarray:
    """gates.py:543-571."""
    return (np.cos(theta) * _i_matrix
            - 1j * np.cos(phi) * np.sin(alpha) * np.sin(theta) * _x_matrix
            - 1j * np.sin(phi) * np.sin(alpha) * np.sin(theta) * _y_matrix
            - 1j * np.sin(theta) * np.cos(alpha) * _z_matrix)


def iswap_matrix(theta: float = 1.0) -> np.ndarray:
    """gates.py:675-705."""
    c, s = np.cos(theta * np.pi / 2), np.sin(theta * np.pi / 2)
    return np.array([[1, 0, 0, 0], [0, c, 1j * s, 0], [0, 1j * s, c, 0], [0, 0, 0, 1]])


def cr_matrix(theta: float = 0.0, alpha: float = 0.0, phi: float = 0.0) -> np.ndarray:
    """gates.py:710-745."""
    return _controlled(r_matrix(theta, alpha, phi))


def exp1_matrix(unitary: Any, theta: float, half: bool = False) -> np.ndarray:
    u = np.asarray(unitary, dtype=np.complex128)
    d = int(round(np.sqrt(u.size)))
    u = u.reshape(d, d)
    if half:
        theta = theta / 2.0
    return np.cos(theta) * np.eye(d) - 1j * np.sin(theta) * u


def exp_matrix(unitary: Any, theta: float) -> np.ndarray:
    """exp(-i theta U) for Hermitian U (gates.py:792-822)."""
    u = np.asarray(unitary, dtype=np.complex128)
    d = int(round(np.sqrt(u.size)))
    u = u.reshape(d, d)
    w, v = np.linalg.eigh(u)
    return (v * np.exp(-1j * theta * w)) @ v.conj().T


HOST_PARAM = {
    "u": (u_matrix, ("theta", "phi", "lbd")),
    "cu": (lambda theta=0.0, phi=0.0, lbd=0.0: _controlled(u_matrix(theta, phi, lbd)), ("theta", "phi", "lbd")),
    "r": (r_matrix, ("theta", "alpha", "phi")),
    "cr": (cr_matrix, ("theta", "alpha", "phi")),
    "iswap": (iswap_matrix, ("theta",)),
    "cphase": (lambda theta=0.0: _controlled(phase_matrix(theta)), ("theta",)),
    "crx": (lambda theta=0.0: _controlled(rx_matrix(theta)), ("theta",)),
    "cry": (lambda theta=0.0: _controlled(ry_matrix(theta)), ("theta",)),
    "crz": (lambda theta=0.0: _controlled(rz_matrix(theta)), ("theta",)),
    "orx": (lambda theta=0.0: _ocontrolled(rx_matrix(theta)), ("theta",)),
    "ory": (lambda theta=0.0: _ocontrolled(ry_matrix(theta)), ("theta",)),
    "orz": (lambda theta=0.0: _ocontrolled(rz_matrix(theta)), ("theta",)),
}


# --- trainable lowering of the other parametrised gates ----------------------
# Every parametrised gate is a product of factors exp(-i scale * theta * G)
# with G Hermitian and G^3 = G, so that dU/dtheta = -i scale G U and the
# adjoint sweep differentiates it like a rotation.

_p0_matrix = np.diag([1.0, 0.0])
_p1_matrix = np.diag([0.0, 1.0])
_p11_matrix = np.diag([0.0, 0.0, 0.0, 1.0])
_hop_matrix = (_xx_matrix + _yy_matrix).real / 2  # |01><10| + |10><01|

# single-parameter gates on their own qubits: name -> (G, scale)
#   c<r>(theta) = |0><0| (x) I + |1><1| (x) exp(-i theta/2 P) = exp(-i theta/2 |1><1| (x) P)
#   o<r>(theta): the same with |0><0| (gates.py:309-326)
#   cphase(theta) = diag(1, 1, 1, e^{i theta}) = exp(+i theta |11><11|)
#   iswap(theta) = exp(+i theta pi/2 (|01><10| + |10><01|))   (gates.py:675-705)
GENERATOR = {
    "crx": (np.kron(_p1_matrix, _x_matrix), 0.5),
    "cry": (np.kron(_p1_matrix, _y_matrix), 0.5),
    "crz": (np.kron(_p1_matrix, _z_matrix), 0.5),
    "orx": (np.kron(_p0_matrix, _x_matrix), 0.5),
    "ory": (np.kron(_p0_matrix, _y_matrix), 0.5),
    "orz": (np.kron(_p0_matrix, _z_matrix), 0.5),
    "cphase": (_p11_matrix, -1.0),
    "iswap": (_hop_matrix, -np.pi / 2),
}

# multi-parameter gates as ordered factor lists (first factor applied first):
# (tcm kind, positions within the gate's qubits, parameter name, scale, G)
#   u(theta, phi, lbd) = phase(phi) ry(theta) phase(lbd)        (gates.py:509-540)
#   cu = the same factors controlled: cphase(phi) cry(theta) cphase(lbd)
#   r(theta, alpha, phi) = exp(-i theta n.sigma), n = (cos phi sin alpha,
#     sin phi sin alpha, cos alpha) = V rz(2 theta) V^dagger with
#     V = rz(phi) ry(alpha)                                     (gates.py:545-575)
#   cr = (I (x) V) crz(2 theta) (I (x) V^dagger)                 (gates.py:720-745)
DECOMPOSE = {
    "u": [(_lib.G_PHASE, (0,), "lbd", 1.0, None), (_lib.G_RY, (0,), "theta", 1.0, None),
          (_lib.G_PHASE, (0,), "phi", 1.0, None)],
    "cu": [(_lib.G_EXPG, (0, 1), "lbd", -1.0, _p11_matrix),
           (_lib.G_EXPG, (0, 1), "theta", 0.5, np.kron(_p1_matrix, _y_matrix)),
           (_lib.G_EXPG, (0, 1), "phi", -1.0, _p11_matrix)],
    "r": [(_lib.G_RZ, (0,), "phi", -1.0, None), (_lib.G_RY, (0,), "alpha", -1.0, None),
          (_lib.G_RZ, (0,), "theta", 2.0, None), (_lib.G_RY, (0,), "alpha", 1.0, None),
          (_lib.G_RZ, (0,), "phi", 1.0, None)],
    "cr": [(_lib.G_RZ, (1,), "phi", -1.0, None), (_lib.G_RY, (1,), "alpha", -1.0, None),
           (_lib.G_EXPG, (0, 1), "theta", 1.0, np.kron(_p1_matrix, _z_matrix)),
           (_lib.G_RY, (1,), "alpha", 1.0, None), (_lib.G_RZ, (1,), "phi", 1.0, None)],
}

# parameter defaults of the reference's gate factories
PARAM_DEFAULT = {"iswap": 1.0}


def factor_matrix(kind: int, g: Optional[np.ndarray], a: float) -> np.ndarray:
    """Host matrix of one lowered factor at angle a = scale * theta (the
    engine's conventions, include/tcm.h)."""
    if kind == _lib.G_RX:
        return _rot(_x_matrix, a)
    if kind == _lib.G_RY:
        return _rot(_y_matrix, a)
    if kind == _lib.G_RZ:
        return _rot(_z_matrix, a)
    if kind == _lib.G_PHASE:
        return phase_matrix(a)
    if kind == _lib.G_EXPG:
        q = g @ g
        return np.eye(g.shape[0]) - q + np.cos(a) * q - 1j * np.sin(a) * g
    raise ValueError("no host matrix for kind %d" % kind)


def is_generator(m: np.ndarray, tol: float = 1e-9) -> bool:
    """Hermitian with m^3 = m (eigenvalues in {-1, 0, 1})."""
    m = np.asarray(m, dtype=np.complex128)
    return bool(np.allclose(m, m.conj().T, atol=tol) and np.allclose(m @ m @ m, m, atol=tol))


def is_involution(m: np.ndarray, tol: float = 1e-9) -> bool:
    return bool(np.allclose(m @ m, np.eye(m.shape[0]), atol=tol))


class Gate:
    """A gate tensor in the reference's leg layout ``[2]*2k`` (what
    ``tc.gates.x()`` returns, gates.py:139-175); ``.tensor`` is a numpy array."""

    def __init__(self, tensor: Any, name: str = "unknowngate"):
        self.tensor = np.asarray(tensor)
        self.name = name

    def matrix(self) -> np.ndarray:
        m = np.asarray(self.tensor, dtype=np.complex128)
        d = int(round(np.sqrt(m.size)))
        return m.reshape(d, d)


def _reshape2(m: np.ndarray) -> np.ndarray:
    d = m.shape[0]
    k = int(round(np.log2(d)))
    return m.reshape([2] * (2 * k))


class _GateFactory:
    def __init__(self, name: str):
        self.n = name

    def __call__(self, **kws: Any) -> Gate:
        name = self.n
        if name in CONST:
            return Gate(_reshape2(np.asarray(CONST[name], dtype=np.complex128)), name)
        th = kws.get("theta", 0.0)
        th = float(np.real(np.asarray(th))) if not hasattr(th, "detach") else float(th.detach().cpu().real)
        if name == "rx":
            m = rx_matrix(th)
        elif name == "ry":
            m = ry_matrix(th)
        elif name == "rz":
            m = rz_matrix(th)
        elif name == "phase":
            m = phase_matrix(th)
        elif name in ("rxx", "ryy", "rzz"):
            m = exp1_matrix({"rxx": _xx_matrix, "ryy": _yy_matrix, "rzz": _zz_matrix}[name], th, half=True)
        elif name == "exp1":
            m = exp1_matrix(kws["unitary"], th, kws.get("half", False))
        elif name == "exp":
            m = exp_matrix(kws["unitary"], th)
        elif name in ("any", "unitary"):
            m = matrix_of(kws["unitary"])
        elif name in HOST_PARAM:
            f, keys = HOST_PARAM[name]
            m = f(**{k: float(np.real(kws.get(k, 1.0 if name == "iswap" else 0.0))) for k in keys})
        else:
            raise ValueError("unknown gate %s" % name)
        return Gate(_reshape2(np.asarray(m, dtype=np.complex128)), name)

    def adjoint(self) -> "_AdjointFactory":
        return _AdjointFactory(self)

    def __repr__(self) -> str:
        return self.n


class _AdjointFactory:
    """The adjoint of a gate factory (GateF.adjoint / GateVF.adjoint,
    gates.py:264-271, 355-368): same parameters, conjugate-transposed
    matrix, name ``<gate>d``."""

    def __init__(self, base: Any):
        self.base = base
        self.n = str(getattr(base, "n", "unknowngate")) + "d"

    def __call__(self, **kws: Any) -> "Gate":
        m = matrix_of(self.base(**kws))
        return Gate(_reshape2(m.conj().T), self.n)

    def adjoint(self) -> Any:
        return self.base

    def __repr__(self) -> str:
        return self.n


def adjoint_factory(gatef: Any) -> Any:
    if hasattr(gatef, "adjoint") and not isinstance(gatef, _GateFactory):
        return gatef.adjoint()
    return _AdjointFactory(gatef)


def matrix_of(u: Any) -> np.ndarray:
    """Square complex128 matrix of a Gate / tensor / array in [2]*2k or 2D
    layout.  Gate matrices are constants of the program: a tensor that
    requires a gradient (or is traced by torch.func) raises instead of being
    silently detached."""
    if isinstance(u, Gate):
        u = u.tensor
    if hasattr(u, "detach"):
        if getattr(u, "requires_grad", False):
            raise NotImplementedError(
                "gradients with respect to gate matrices are not supported: the matrix is a constant of the "
                "compiled program; parametrise the gate by angles (rx/ry/rz/rxx/ryy/rzz/phase/exp1/exp/u/r/"
                "cu/cr/crx/cry/crz/orx/ory/orz/cphase/iswap)")
        try:
            u = u.detach().cpu().numpy()
        except Exception as e:  # functorch-wrapped tensors cannot be read
            raise NotImplementedError("a gate matrix traced by torch.func cannot be used as a constant") from e
    m = np.asarray(u, dtype=np.complex128)
    d = int(round(np.sqrt(m.size)))
    if d * d != m.size:
        raise ValueError("gate matrix of %d entries is not square" % m.size)
    return m.reshape(d, d)


for _name in sgates + vgates:
    globals()[_name] = _GateFactory(_name)
    globals()[_name + "_gate"] = globals()[_name]
    globals()[_name + "gate"] = globals()[_name]
for _a in gate_aliases:
    for _b in _a[1:]:
        globals()[_b] = globals()[_a[0]]

pauli_gates = [globals()["i"](), globals()["x"](), globals()["y"](), globals()["z"]()]


def num_to_tensor(*num: Any, dtype: Optional[str] = None) -> Any:
    """gates.py:180-235: cast to the global complex dtype (torch tensor)."""
    import torch

    from . import cons

    dt = getattr(torch, dtype or cons.dtypestr)
    out = [torch.as_tensor(x).to(dt) if not isinstance(x, torch.Tensor) else x.to(dt) for x in num]
    return out[0] if len(out) == 1 else out


array_to_tensor = num_to_tensor


def pauli_code(m: np.ndarray) -> Optional[int]:
    """0..3 when ``m`` is exactly I/X/Y/Z, else None."""
    for code, p in enumerate((_i_matrix, _x_matrix, _y_matrix, _z_matrix)):
        if m.shape == (2, 2) and np.array_equal(np.asarray(m, dtype=np.complex128), np.asarray(p, dtype=np.complex128)):
            return code
    return None
