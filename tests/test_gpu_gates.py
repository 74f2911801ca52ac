"""Every parametrised gate is trainable on the device: gradients through
crx/cry/crz/orx/ory/orz/cphase/iswap/u/cu/r/cr/phase/exp/exp1 (the
reference's GateVF gates, gates.py:339-368) by the adjoint sweep, against
five-point central differences of the complex128 oracle; the reference's
universal_ad (tests/test_circuit.py:295-316); Circuit.inverse() and append
(abstractcircuit.py:428-507, 1148-1190) with the reference's inverse tests
(tests/test_circuit.py:1314-1343, 1561-1578) and U^dagger U = I on random
circuits."""
import numpy as np
import pytest
import torch

from oracle import statevector as sv
from oracle.statevector import Op

pytestmark = pytest.mark.gpu

N = 4
# (gate, qubits, parameter names)
CASES = [
    ("crx", (0, 2), ("theta",)), ("cry", (3, 1), ("theta",)), ("crz", (1, 2), ("theta",)),
    ("orx", (2, 0), ("theta",)), ("ory", (0, 3), ("theta",)), ("orz", (3, 2), ("theta",)),
    ("cphase", (1, 3), ("theta",)), ("iswap", (0, 1), ("theta",)), ("phase", (2,), ("theta",)),
    ("u", (2,), ("theta", "phi", "lbd")), ("cu", (3, 0), ("theta", "phi", "lbd")),
    ("r", (1,), ("theta", "alpha", "phi")), ("cr", (2, 1), ("theta", "alpha", "phi")),
    ("exp", (1, 2), ("theta",)), ("exp1", (0, 3), ("theta",)),
]
EXP_G = np.kron(np.diag([0.0, 1.0]), sv.CONST_GATES["x"])       # G^3 = G, not an involution
EXP1_U = np.kron(sv.CONST_GATES["x"], sv.CONST_GATES["y"])      # involution
TERMS = [(0.7, [3, 0, 1, 0]), (-1.1, [0, 2, 2, 0]), (0.4, [1, 0, 0, 3]), (0.9, [0, 3, 0, 3]),
         (-0.5, [2, 1, 0, 0])]


def _oracle_gate(name, qs, kw):
    if name == "cr":
        return sv._ctrl(sv.r_gate(kw["theta"], kw["alpha"], kw["phi"]))
    if name == "exp":
        return sv.expm_herm(EXP_G, kw["theta"])
    if name == "exp1":
        return sv.exp1(EXP1_U, kw["theta"])
    return sv.gate_matrix(name, **kw)


def _prefix_ops(rng):
    ops = []
    for i in range(N):
        ops.append(Op("ry", (i,), kw={"theta": float(rng.uniform(0, 2 * np.pi))}))
    for i in range(N - 1):
        ops.append(Op("cnot", (i, i + 1)))
    return ops


def _oracle_energy(name, qs, pnames, x, prefix, suffix):
    kw = dict(zip(pnames, x))
    s = sv.run_ops(N, prefix)
    s = sv.apply_gate(s, _oracle_gate(name, qs, kw), qs, N)
    s = sv.run_ops(N, suffix, inputs=s)
    return float(np.real(sv.pauli_sum_expectation(s, N, TERMS)))


def _five_point(f, x, h=1e-3):
    g = np.zeros_like(x)
    for i in range(x.size):
        def at(d):
            y = x.copy()
            y[i] += d
            return f(y)
        g[i] = (-at(2 * h) + 8 * at(h) - 8 * at(-h) + at(-2 * h)) / (12 * h)
    return g


def _build(tc, name, qs, pnames, p, prefix, suffix):
    c = tc.Circuit(N)
    for o in prefix:
        getattr(c, o.name)(*o.qubits, **o.kw)
    kw = {k: p[j] for j, k in enumerate(pnames)}
    if name == "exp":
        kw["unitary"] = EXP_G
    if name == "exp1":
        kw["unitary"] = EXP1_U
    getattr(c, name)(*qs, **kw)
    for o in suffix:
        getattr(c, o.name)(*o.qubits, **o.kw)
    return c


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_parametrised_gate_gradient(gpu, dtype, case):
    """value_and_grad through the gate's trainable lowering vs central
    differences of the oracle: 1e-5 (complex64), 1e-9 (complex128)."""
    import tensorcircuit_amd as tc

    name, qs, pnames = case
    tc.set_dtype(dtype)
    try:
        rng = np.random.default_rng(sum(map(ord, name)))
        prefix = _prefix_ops(rng)
        suffix = _prefix_ops(rng)[::-1]
        x = rng.uniform(-np.pi, np.pi, len(pnames))
        rdt = torch.float32 if dtype == "complex64" else torch.float64
        if dtype == "complex64":
            x = x.astype(np.float32).astype(np.float64)  # the oracle sees the rounded angles

        def energy(p):
            c = _build(tc, name, qs, pnames, p, prefix, suffix)
            return sum(w * tc.backend.real(c.expectation_ps(ps=codes)) for w, codes in TERMS)

        v, g = tc.backend.value_and_grad(energy)(torch.as_tensor(x, dtype=rdt))
        f = lambda y: _oracle_energy(name, qs, pnames, y, prefix, suffix)  # noqa: E731
        e_ref = f(x)
        g_ref = _five_point(f, x)
        tol = 1e-5 if dtype == "complex64" else 1e-9
        assert abs(float(v) - e_ref) < tol
        np.testing.assert_allclose(g.cpu().numpy(), g_ref, atol=tol)
        # plain autograd (not torch.func) reaches the same gradient
        pt = torch.as_tensor(x, dtype=rdt).requires_grad_(True)
        energy(pt).backward()
        np.testing.assert_allclose(pt.grad.cpu().numpy(), g_ref, atol=tol)
    finally:
        tc.set_dtype("complex64")


def test_universal_ad(gpu):
    """tests/test_circuit.py:295-316: grad and value_and_grad through R with a
    complex theta agree; <Z> = cos^2 t + sin^2 t cos 2a, d/dt = -2 sin 2t sin^2 a."""
    import tensorcircuit_amd as tc

    K = tc.backend

    def forward(theta):
        c = tc.Circuit(2)
        c.R(0, theta=theta, alpha=0.5, phi=0.8)
        return K.real(c.expectation((tc.gates.z(), [0])))

    gg = K.jit(K.grad(forward))
    vg = K.jit(K.value_and_grad(forward))
    theta = tc.gates.num_to_tensor(1.0)
    g1 = gg(theta)
    v2, g2 = vg(theta)
    assert complex(g1) == complex(g2)
    t, a = 1.0, 0.5
    assert abs(float(v2) - (np.cos(t) ** 2 + np.sin(t) ** 2 * np.cos(2 * a))) < 1e-5
    assert abs(complex(g2).real - (-2 * np.sin(2 * t) * np.sin(a) ** 2)) < 1e-5


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_reference_inverse_kats(gpu, dtype):
    """tests/test_circuit.py:1314-1343: c.append(c.inverse()) returns the
    input state."""
    import tensorcircuit_amd as tc

    tc.set_dtype(dtype)
    try:
        tol = 1e-5 if dtype == "complex64" else 1e-12
        rng = np.random.default_rng(0)
        inputs = rng.uniform(size=[8])
        inputs /= np.linalg.norm(inputs)
        c = tc.Circuit(3, inputs=inputs)
        c.H(1)
        c.rx(0, theta=0.5)
        c.cnot(1, 2)
        c.rzz(0, 2, theta=-0.8)
        c.append(c.inverse())
        np.testing.assert_allclose(c.state().cpu().numpy(), inputs, atol=tol)
        c = tc.Circuit(3, inputs=inputs)
        c.iswap(0, 1)
        c.iswap(1, 0, theta=0.6)
        c.rxx(1, 2, theta=-0.2)
        c.cu(0, 1, lbd=2.0, theta=-0.7)
        c.r(2, alpha=0.3)
        c.sd(2)
        c.cx(1, 2)
        c.unitary(0, unitary=tc.gates._x_matrix)
        c.append(c.inverse())
        np.testing.assert_allclose(c.state().cpu().numpy(), inputs, atol=tol)
    finally:
        tc.set_dtype("complex64")


def test_inverse_under_vvag(gpu):
    """tests/test_circuit.py:1561-1578: vvag through an inverse circuit with
    trainable rzz angles; <Z_1> after the inverse and an X is -1."""
    import tensorcircuit_amd as tc

    K = tc.backend

    def simple_ansatz(param):
        c = tc.Circuit(3)
        for i in range(3):
            c.cx(i, (i + 1) % 3)
            c.rzz(i, (i + 1) % 3, theta=param[i])
        c1 = c.inverse()
        c2 = tc.Circuit(3)
        c2.x(1)
        c1.append(c2)
        return K.real(c1.expectation_ps(z=[1]))

    vs, gs = K.jit(K.vvag(simple_ansatz))(torch.ones([2, 3], dtype=torch.float32))
    assert tuple(gs.shape) == (2, 3)
    np.testing.assert_allclose(vs.cpu().numpy(), -np.ones(2), atol=1e-5)
    np.testing.assert_allclose(gs.cpu().numpy(), np.zeros((2, 3)), atol=1e-5)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_inverse_random_circuits(gpu, dtype):
    """U^dagger U = I on random 7-qubit circuits over every gate kind (tensor
    and number parameters); the JSON round trip of the inverse gives the
    same state."""
    import tensorcircuit_amd as tc

    tc.set_dtype(dtype)
    try:
        n = 7
        tol = 2e-5 if dtype == "complex64" else 1e-12
        rng = np.random.default_rng(21)
        names1 = ["h", "x", "y", "z", "s", "t", "sd", "td", "wroot", "rx", "ry", "rz", "phase", "u", "r"]
        names2 = ["cnot", "cz", "swap", "cy", "ox", "oy", "oz", "crx", "cry", "crz", "orx", "ory", "orz",
                  "cphase", "iswap", "rxx", "ryy", "rzz", "cu", "cr"]
        rdt = torch.float32 if dtype == "complex64" else torch.float64
        psi0 = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
        psi0 /= np.linalg.norm(psi0)
        c = tc.Circuit(n, inputs=psi0)
        for k in range(120):
            two = rng.random() < 0.5
            name = names2[rng.integers(len(names2))] if two else names1[rng.integers(len(names1))]
            qs = tuple(int(q) for q in rng.permutation(n)[: 2 if two else 1])
            kw = {}
            pn = {"u": ("theta", "phi", "lbd"), "cu": ("theta", "phi", "lbd"), "r": ("theta", "alpha", "phi"),
                  "cr": ("theta", "alpha", "phi")}.get(name, ("theta",))
            if name not in ("h", "x", "y", "z", "s", "t", "sd", "td", "wroot", "cnot", "cz", "swap", "cy", "ox",
                            "oy", "oz"):
                for pnm in pn:
                    val = float(rng.uniform(-3, 3))
                    kw[pnm] = torch.tensor(val, dtype=rdt) if rng.random() < 0.5 else val
            getattr(c, name)(*qs, **kw)
            if rng.random() < 0.05:
                c.unitary(*qs, unitary=np.linalg.qr(rng.standard_normal((2 ** len(qs),) * 2)
                                                     + 1j * rng.standard_normal((2 ** len(qs),) * 2))[0])
            if rng.random() < 0.05:
                c.exp1(*qs, unitary=np.eye(2 ** len(qs))[::-1].copy(), theta=float(rng.uniform(-2, 2)))
        inv = c.inverse()
        d = tc.Circuit(n, inputs=psi0)
        d.append(c)
        d.append(inv)
        np.testing.assert_allclose(d.state().cpu().numpy(), psi0, atol=tol)
        # inverse applied to the forward state, rebuilt from its JSON
        fwd = c.state()
        e = tc.Circuit.from_json(inv.to_json(), circuit_params={"nqubits": n, "inputs": fwd})
        np.testing.assert_allclose(e.state().cpu().numpy(), psi0, atol=tol)
    finally:
        tc.set_dtype("complex64")
