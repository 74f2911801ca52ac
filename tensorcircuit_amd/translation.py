"""Circuit wire formats: the JSON gate list (translation.py:587-702 of the
reference: qir2json / json2qir, tensor_to_json / json_to_tensor) and the
qsim text format (abstractcircuit.py:953-1034, after the qsim input format
of quantumlib/qsim).  Host-side conversions; the circuits they build run on
the device like any other."""
from __future__ import annotations

from typing import Any, Dict, List, Tuple

import numpy as np
import torch

from . import gates as G


def tensor_to_json(a: Any) -> List[Any]:
    """[real] or [real, imag] nested lists (translation.py:587-603)."""
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    a = np.asarray(a)
    if np.iscomplexobj(a):
        return [np.real(a).tolist(), np.imag(a).tolist()]
    return [a.tolist()]


def json_to_tensor(a: List[Any]) -> np.ndarray:
    """Inverse of tensor_to_json (translation.py:606-613)."""
    ar = np.array(a[0])
    if len(a) == 1:
        return ar
    assert len(a) == 2
    return ar + 1.0j * np.array(a[1])


def get_u_parameter(m: Any) -> Tuple[float, float, float]:
    """(theta, phi, lbd) of the OpenQASM U gate equal to the 2x2 unitary m up
    to a global phase (gates.py:485-506): with U = m / sqrt(det m) in SU(2),
    |U11| = cos(theta/2), arg U11 = (phi + lbd)/2, arg U10 = (phi - lbd)/2."""
    m = np.asarray(m, dtype=np.complex128).reshape(2, 2)
    u = m * np.linalg.det(m) ** (-0.5)
    theta = 2 * np.arccos(min(1.0, abs(u[1, 1])))
    s = 2 * np.angle(u[1, 1])
    d = -2 * np.angle(u[1, 0])
    return float(theta), float((s - d) / 2), float((s + d) / 2)


_NEEDS_MATRIX = ("any", "unitary", "mpo", "exp", "exp1", "r", "cr")


def _matrix(r: Dict[str, Any]) -> np.ndarray:
    if r.get("gate") is not None:
        return G.matrix_of(r["gate"])
    return G.matrix_of(r["gatef"](**dict(r.get("parameters", {}))))


def qir2json(qir: List[Dict[str, Any]], simplified: bool = False) -> List[Dict[str, Any]]:
    """QIR -> JSON-compatible gate list (translation.py:616-673): name,
    qubits and, unless simplified, the matrix, U parameters (2x2 gates),
    parameters and mpo flag; simplified keeps what is needed to rebuild."""
    out = []
    for r in qir:
        gatef = r.get("gatef")
        if gatef is None and r.get("gate") is None:
            continue  # barriers and other markers
        name = gatef.n if gatef is not None else r.get("name", "any")
        m = _matrix(r)
        mj = tensor_to_json(m)
        uparams = list(get_u_parameter(m)) if m.shape == (2, 2) else []
        params = {k: tensor_to_json(v) for k, v in dict(r.get("parameters", {})).items()}
        d: Dict[str, Any] = {"name": name, "qubits": [int(i) for i in r["index"]]}
        if not simplified:
            d.update({"matrix": mj, "uparams": uparams, "parameters": params, "mpo": bool(r.get("mpo", False))})
        else:
            if name in _NEEDS_MATRIX and uparams:
                d["uparams"] = uparams
            elif name in _NEEDS_MATRIX:
                d["matrix"] = mj
            if params:
                d["parameters"] = params
            if r.get("mpo"):
                d["mpo"] = True
        out.append(d)
    return out


def json2qir(tcqasm: List[Dict[str, Any]]) -> List[Dict[str, Any]]:
    """JSON gate list -> QIR (translation.py:676-702)."""
    qir = []
    for d in tcqasm:
        param = {k: json_to_tensor(v) for k, v in dict(d.get("parameters", {})).items()}
        param = {k: (v.item() if np.ndim(v) == 0 else v) for k, v in param.items()}
        name = d["name"]
        if hasattr(G, name):
            gatef = getattr(G, name)
        elif name.endswith("d") and hasattr(G, name[:-1]):
            gatef = G.adjoint_factory(getattr(G, name[:-1]))  # an inverse()'s adjoint gate
        else:
            raise ValueError("unknown gate %r in JSON" % name)
        if "matrix" in d:
            gatem = G.Gate(G._reshape2(np.asarray(json_to_tensor(d["matrix"]), dtype=np.complex128)), name)
        elif "uparams" in d and name in _NEEDS_MATRIX:
            gatem = G.Gate(G._reshape2(G.u_matrix(*d["uparams"])), name)
        else:
            gatem = gatef(**param)
        qir.append({"gate": gatem, "index": tuple(d["qubits"]), "mpo": d.get("mpo", False), "split": {},
                    "parameters": param, "gatef": gatef, "name": name})
    return qir


def _qsim_token(x: str) -> Any:
    try:
        return int(x)
    except ValueError:
        pass
    try:
        return float(x)
    except ValueError:
        return x.lower()


def apply_qsim(c: Any, lines: List[str]) -> Any:
    """Append the gates of qsim text lines (after the qubit-count line) to
    circuit c (abstractcircuit.py:968-1034): "time name q0 [q1] [params]"."""
    for line in lines:
        if not line.strip():
            continue
        g = tuple(_qsim_token(t) for t in line.strip().split())
        name, q = g[1], g[2:]
        if name in ("h", "x", "y", "z"):
            getattr(c, name)(q[0])
        elif name == "s":
            c.phase(q[0], theta=np.pi / 2)
        elif name == "t":
            c.phase(q[0], theta=np.pi / 4)
        elif name == "x_1_2":
            c.rx(q[0], theta=np.pi / 2)
        elif name == "y_1_2":
            c.ry(q[0], theta=np.pi / 2)
        elif name == "z_1_2":
            c.rz(q[0], theta=np.pi / 2)
        elif name == "w_1_2":
            c.u(q[0], theta=np.pi / 2, phi=-np.pi / 4, lbd=np.pi / 4)
        elif name == "hz_1_2":
            c.wroot(q[0])
        elif name in ("cnot", "cx"):
            c.cnot(q[0], q[1])
        elif name == "cy":
            c.cy(q[0], q[1])
        elif name == "cz":
            c.cz(q[0], q[1])
        elif name in ("is", "iswap"):
            c.iswap(q[0], q[1])
        elif name in ("rx", "ry", "rz"):
            getattr(c, name)(q[0], theta=float(q[1]))
        elif name in ("fs", "fsim"):
            i, j, theta, phi = q
            c.iswap(i, j, theta=-float(theta))
            c.cphase(i, j, theta=-float(phi))
        else:
            raise NotImplementedError("qsim gate %s" % name)
    return c
