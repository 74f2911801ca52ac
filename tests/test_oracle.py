"""Pins the CPU oracle before anything is checked against it.

Each check is one the reference's own tests hold (file:line cited), a closed
form, or an independent brute-force restatement.  CPU only.
"""
import numpy as np
import pytest

from oracle import statevector as sv
from oracle.statevector import Op


def run(n, ops, params=None, inputs=None):
    return sv.run_ops(n, ops, params, inputs)


# ------------------------------------------------------------- index order ---

def test_wavefunction_index_order():
    """tests/test_circuit.py:21-43."""
    m16 = np.arange(16).reshape(4, 4)
    s = run(2, [Op("unitary", (0, 1), kw={"unitary": m16})])
    assert np.real(s[2]) == 8
    s = run(2, [Op("unitary", (1, 0), kw={"unitary": m16})])
    assert np.real(s[2]) == 4
    s = run(2, [Op("unitary", (0,), kw={"unitary": np.arange(4).reshape(2, 2)})])
    assert np.real(s[2]) == 2


def test_amplitude_basics():
    """tests/test_circuit.py:46-52: x(0) -> |10>, then CNOT -> |11>."""
    s = run(2, [Op("x", (0,))])
    assert s[0b10] == 1.0
    s = run(2, [Op("x", (0,)), Op("cnot", (0, 1))])
    assert s[0b11] == 1.0


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
def test_tensordot_vs_bruteforce(n):
    rng = np.random.default_rng(n)
    s = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    for k in (1, 2, 3):
        if k > n:
            continue
        u = rng.standard_normal((2**k, 2**k)) + 1j * rng.standard_normal((2**k, 2**k))
        qs = tuple(int(q) for q in rng.permutation(n)[:k])
        np.testing.assert_allclose(sv.apply_gate(s, u, qs, n),
                                   sv.apply_gate_bruteforce(s, u, qs, n), atol=1e-12)


def test_kron_layout():
    """A gate on (0,1,...) of an n-qubit state equals kron(U, I)."""
    rng = np.random.default_rng(3)
    n = 4
    s = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    u = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    np.testing.assert_allclose(sv.apply_gate(s, u, (0, 1), n), np.kron(u, np.eye(4)) @ s, atol=1e-12)
    np.testing.assert_allclose(sv.apply_gate(s, u, (2, 3), n), np.kron(np.eye(4), u) @ s, atol=1e-12)


# ------------------------------------------------------------- gate KATs ---

def test_gates_in_circuit_iswap():
    """tests/test_circuit.py:64-68: inputs=eye(4), iswap -> the iswap matrix."""
    s = sv.apply_gate(np.eye(4).reshape(-1).astype(complex), sv.iswap(1.0), (0, 1), 4)
    np.testing.assert_allclose(s.reshape(4, 4), sv.iswap(1.0), atol=1e-12)


def test_control_vgate():
    """tests/test_circuit.py:71-77."""
    s = run(2, [Op("x", (1,)), Op("crx", (1, 0), kw={"theta": 0.3})])
    np.testing.assert_allclose(sv.expectation(s, 2, (sv._Z, [0])), 0.95533645, atol=1e-7)


def test_adjoint_gate_circuit():
    """tests/test_circuit.py:80-84."""
    s = run(1, [Op("x", (0,)), Op("sd", (0,))])
    np.testing.assert_allclose(s, [0.0, -1.0j])


def test_phase_gate():
    """tests/test_gates.py:18-22."""
    s = run(1, [Op("h", (0,)), Op("phase", (0,), kw={"theta": np.pi / 2})])
    np.testing.assert_allclose(s[1], 0.7071j, atol=1e-4)


def test_cu_gate():
    """tests/test_gates.py:25-31."""
    m = sv.gate_matrix("cu", theta=np.pi / 2, phi=-np.pi / 4, lbd=np.pi / 4)
    np.testing.assert_allclose(m[2:, 2:], sv._WROOT, atol=1e-5)
    np.testing.assert_allclose(m[:2, :2], np.eye(2), atol=1e-5)


def test_fsim_gate():
    """tests/test_gates.py:61-77."""
    ops = [Op("iswap", (0, 1), kw={"theta": -0.2}), Op("cphase", (0, 1), kw={"theta": -0.3})]
    u = np.eye(4, dtype=complex)
    for op in ops:
        u = np.stack([sv.apply_gate(u[:, c], op.matrix(), op.qubits, 2) for c in range(4)], axis=1)
    ans = np.array([
        [1, 0, 0, 0],
        [0, 0.95105654, -0.309017j, 0],
        [0, -0.309017j, 0.95105654, 0],
        [0, 0, 0, 0.9553365 - 0.29552022j]])
    np.testing.assert_allclose(u, ans, atol=1e-5)


def test_exp_gate():
    """tests/test_gates.py:80-90."""
    s = run(2, [Op("exp", (0, 1), kw={"unitary": np.diag([1.0, -1, -1, 1]), "theta": np.pi / 2})])
    np.testing.assert_allclose(s[0], -1j, atol=1e-12)


def test_any_gate():
    """tests/test_gates.py:93-96."""
    s = run(2, [Op("any", (0,), kw={"unitary": np.eye(2)})])
    np.testing.assert_allclose(sv.expectation(s, 2, (sv._Z, [0])), 1.0)


def test_iswap_tensor():
    """tests/test_gates.py:99-104."""
    ans = np.array([[1.0, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1.0]])
    np.testing.assert_allclose(sv.iswap(), ans, atol=1e-12)
    np.testing.assert_allclose(sv.iswap(0), np.eye(4), atol=1e-12)


def test_ocontrolled_any():
    """tests/test_gates.py:111-124: ox(cx) applied on (1, 0, 2) after x(0)."""
    ocx = sv._octrl(sv.CONST_GATES["cnot"])
    s = run(3, [Op("x", (0,)), Op("any", (1, 0, 2), kw={"unitary": ocx})])
    np.testing.assert_allclose(sv.expectation(s, 3, (sv._Z, [2])), -1, atol=1e-12)
    np.testing.assert_allclose(sv.CONST_GATES["toffoli"], sv._ctrl(sv._ctrl(sv._X)))


def test_rxx_is_half_exp1():
    """tests/test_gates.py:147-156."""
    a = run(3, [Op("rxx", (0, 1), kw={"theta": 1.0}), Op("ryy", (0, 2), kw={"theta": 0.5}),
                Op("rzz", (0, 1), kw={"theta": -0.5})])
    b = run(3, [Op("exp1", (0, 1), kw={"theta": 0.5, "unitary": sv._XX}),
                Op("exp1", (0, 2), kw={"theta": 0.25, "unitary": sv._YY}),
                Op("exp1", (0, 1), kw={"theta": -0.25, "unitary": sv._ZZ})])
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_exp1_vs_exp():
    """tests/test_circuit.py:248-272: exp1 and exp agree for an involution."""
    xx = np.array([[0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0], [1, 0, 0, 0]], dtype=complex)
    a = run(2, [Op("exp1", (0, 1), kw={"unitary": xx, "theta": 0.2})])
    b = run(2, [Op("exp", (0, 1), kw={"unitary": xx, "theta": 0.2})])
    np.testing.assert_allclose(a, b, atol=1e-12)


@pytest.mark.parametrize("name", ["rx", "ry", "rz", "rzz", "rxx", "ryy", "phase", "crx", "u", "r"])
def test_param_gates_unitary(name):
    m = sv.gate_matrix(name, theta=0.731, phi=0.2, lbd=-0.4, alpha=1.1)
    np.testing.assert_allclose(m.conj().T @ m, np.eye(m.shape[0]), atol=1e-12)


def test_unitary_kron_pairs():
    """tests/test_circuit.py:404-412: inputs=eye(4), X(0), Y(1) -> kron(X, Y)."""
    s = np.eye(4).reshape(-1).astype(complex)
    s = sv.apply_gate(s, sv._X, (0,), 4)
    s = sv.apply_gate(s, sv._Y, (1,), 4)
    np.testing.assert_allclose(s.reshape(4, 4), np.kron(sv._X, sv._Y), atol=1e-12)


# ------------------------------------------------------------ expectations ---

def test_expectation_h():
    """tests/test_circuit.py:239-245."""
    s = run(2, [Op("h", (0,))])
    np.testing.assert_allclose(sv.expectation(s, 2, (sv._Z, [0])), 0, atol=1e-12)


def test_expectation_ps_kats():
    """tests/test_circuit.py:415-429."""
    s = run(2, [Op("x", (0,))])
    np.testing.assert_allclose(sv.expectation_ps(s, 2, z=[0, 1]), -1, atol=1e-12)
    s = run(2, [Op("h", (0,))])
    np.testing.assert_allclose(sv.expectation_ps(s, 2, z=[1], x=[0]), 1, atol=1e-12)
    np.testing.assert_allclose(sv.expectation_ps(s, 2, ps=[1, 3]), 1, atol=1e-12)
    np.testing.assert_allclose(sv.expectation_ps(s, 2, z=[1, 2], ps=[1, 3]), 1, atol=1e-12)


def test_repeated_index_raises():
    """basecircuit.py:306-307."""
    s = sv.zero_state(2)
    with pytest.raises(ValueError):
        sv.expectation(s, 2, (sv._Z, [0]), (sv._X, [0]))


def test_expectation_y_bug():
    """tests/test_circuit.py:1278-1281."""
    s = np.array([-1, 1j]) / np.sqrt(2)
    np.testing.assert_allclose(sv.expectation_ps(s, 1, y=[0]), -1, atol=1e-12)


def test_minus_index():
    """tests/test_circuit.py:1373-1380."""
    s = run(3, [Op("h", (-2,)), Op("h", (0,))])
    np.testing.assert_allclose(np.real(sv.expectation_ps(s, 3, x=[0])), 1, atol=1e-12)
    np.testing.assert_allclose(np.real(sv.expectation_ps(s, 3, x=[1])), 1, atol=1e-12)
    np.testing.assert_allclose(np.real(sv.expectation_ps(s, 3, x=[-1])), 0, atol=1e-12)
    np.testing.assert_allclose(np.real(sv.expectation_ps(s, 3, z=[-2])), 0, atol=1e-12)


def test_any_inputs_state():
    """tests/test_circuit.py:370-390."""
    for inp, want in [([0, 0, 0, 1.0], 1.0), ([0, 0, 1.0, 0], 1.0), ([1.0, 0, 0, 0], -1.0),
                      ([1 / np.sqrt(2), 0, 1 / np.sqrt(2), 0], 0.0)]:
        s = run(2, [Op("x", (0,))], inputs=np.array(inp))
        np.testing.assert_allclose(sv.expectation(s, 2, (sv._Z, [0])), want, atol=1e-12)


CHECK_PAIRS = [
    ([0, 0], np.eye(4)),
    ([0, 1], np.kron(sv._I, sv._X)),
    ([2, 1], np.kron(sv._Y, sv._X)),
    ([3, 1], np.kron(sv._Z, sv._X)),
    ([3, 2, 2, 0], np.kron(np.kron(np.kron(sv._Z, sv._Y), sv._Y), sv._I)),
    ([0, 1, 1, 1], np.kron(np.kron(np.kron(sv._I, sv._X), sv._X), sv._X)),
]


@pytest.mark.parametrize("codes,mat", CHECK_PAIRS)
def test_pauli_index_rule_vs_dense(codes, mat):
    """tests/test_miscs.py:24-55 pairs: the index rule of ps2coo_core equals the
    dense kron matrix."""
    n = len(codes)
    np.testing.assert_allclose(sv.pauli_string_matrix(codes), mat, atol=1e-12)
    rng = np.random.default_rng(7)
    s = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    np.testing.assert_allclose(sv.apply_pauli_string(s, n, codes), mat @ s, atol=1e-12)


def test_ps2xyz():
    """tests/test_quantum.py:495-503."""
    assert sv.ps2xyz([0, 1, 3]) == {"x": [1], "y": [], "z": [2]}
    assert sv.ps2xyz([0, 1, 3, 0]) == {"x": [1], "y": [], "z": [2]}


def test_heisenberg_ring_ground_energy():
    """tests/test_quantum.py:316-321 (Line1D(6) is a ring): -11.2111."""
    n = 6
    terms = []
    for a in range(n):
        b = (a + 1) % n
        for code in (3, 1, 2):
            c = [0] * n
            c[a] = code
            c[b] = code
            terms.append((1.0, c))
    h = sum(w * sv.pauli_string_matrix(c) for w, c in terms)
    np.testing.assert_allclose(np.linalg.eigvalsh(h)[0], -11.2111, atol=1e-4)


def test_any_measurements():
    """tests/test_templates.py:17-26 and 29-40."""
    s = run(2, [Op("h", (0,)), Op("h", (1,))])
    np.testing.assert_allclose(sv.expectation_ps(s, 2, ps=[1, 1]), 1.0, atol=1e-12)
    np.testing.assert_allclose(sv.expectation_ps(s, 2, ps=[3, 0]), 0.0, atol=1e-12)
    s = run(3, [Op("x", (0,)), Op("cnot", (0, 1)), Op("h", (-1,))])
    r = [sv.expectation_ps(s, 3, ps=[3 if i == q else 0 for i in range(3)]) for q in (0, 1)]
    r.append(sv.expectation_ps(s, 3, x=[2]))
    np.testing.assert_allclose(np.real(r), [-1, -1, 1], atol=1e-12)


def test_pauli_sum_expectation_dense():
    rng = np.random.default_rng(11)
    n = 5
    s = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    terms = [(float(rng.standard_normal()), list(rng.integers(0, 4, n))) for _ in range(9)]
    hm = sum(w * sv.pauli_string_matrix(c) for w, c in terms)
    np.testing.assert_allclose(sv.pauli_sum_expectation(s, n, terms), np.vdot(s, hm @ s), atol=1e-10)


# --------------------------------------------------------------- workloads ---

def test_hea_zero_angles_closed_form():
    """All-zero angles leave |0..0>, so E_TFIM = -(n-1) exactly."""
    n, L = 6, 3
    ops = sv.hea_ops(n, L)
    e = sv.energy(n, ops, np.zeros(2 * L * n), sv.tfim_terms(n))
    assert abs(e + (n - 1)) < 1e-12


def test_hea_pi_ry_basis_state():
    """ry(pi) on every qubit, rz=0: |1..1> -> CNOT chain gives the prefix-parity
    pattern 1,0,1,0,...; one layer."""
    n = 5
    p = np.zeros((2, n))
    p[0, :] = np.pi
    s = sv.run_ops(n, sv.hea_ops(n, 1), p.reshape(-1))
    bits = [(i + 1) % 2 for i in range(n)]  # prefix parity of all ones
    idx = int("".join(map(str, bits)), 2)
    np.testing.assert_allclose(abs(s[idx]), 1.0, atol=1e-12)


def test_tfim_variational_bound():
    n, L = 6, 2
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 2 * np.pi, 2 * L * n)
    e = sv.energy(n, sv.hea_ops(n, L), p, sv.tfim_terms(n))
    assert e >= sv.tfim_ground_energy(n) - 1e-10


def test_haar_su2_unitary():
    rng = np.random.default_rng(1005)
    for _ in range(20):
        u = sv.haar_su2(rng)
        np.testing.assert_allclose(u.conj().T @ u, np.eye(2), atol=1e-12)
        np.testing.assert_allclose(np.linalg.det(u), 1.0, atol=1e-12)


# --------------------------------------------------------------- gradients ---

def test_adjoint_vs_finite_difference_and_shift():
    n, L = 4, 2
    rng = np.random.default_rng(1001)
    p = rng.uniform(0, 2 * np.pi, 2 * L * n)
    ops, terms = sv.hea_ops(n, L), sv.tfim_terms(n)
    e, g = sv.adjoint_value_and_grad(n, ops, p, terms)
    np.testing.assert_allclose(e, sv.energy(n, ops, p, terms), atol=1e-12)
    np.testing.assert_allclose(g, sv.parameter_shift_grad(n, ops, p, terms), atol=1e-10)
    np.testing.assert_allclose(g, sv.finite_difference_grad(n, ops, p, terms), atol=1e-6)


def _example_block_ops(n, nlayers):
    """templates/blocks.py:113-152 with exp1 (half=False) ZZ and rx."""
    ops = [Op("h", (i,)) for i in range(n)]
    for j in range(nlayers):
        for i in range(n - 1):
            ops.append(Op("exp1", (i, i + 1), pidx=(2 * j) * n + i, kw={"unitary": sv._ZZ}))
        for i in range(n):
            ops.append(Op("rx", (i,), pidx=(2 * j + 1) * n + i))
    return ops


def test_interface_gradient_kat_1():
    """tests/test_interfaces.py:30-57: loss=<X_1>^2, param ones [4,4],
    d loss / d param[0,1] = -2.146e-3."""
    n = 4
    ops = _example_block_ops(n, 2)
    p = np.ones(4 * n)
    x1 = [(1.0, [0, 1, 0, 0])]
    e, g = sv.adjoint_value_and_grad(n, ops, p, x1)
    grad = 2 * e * g  # chain rule for loss = e^2
    np.testing.assert_allclose(grad[0 * n + 1], -2.146e-3, atol=1e-5)


def test_interface_gradient_kat_2():
    """tests/test_interfaces.py:59-97: loss=<X_1>-<X_2>, d/dparamzz[0,0] = -0.41609."""
    n = 4
    ops = [Op("h", (i,)) for i in range(n)]
    for j in range(2):
        for i in range(n - 1):
            ops.append(Op("exp1", (i, i + 1), pidx=j * n + i, kw={"unitary": sv._ZZ}))
        for i in range(n):
            ops.append(Op("rx", (i,), pidx=2 * n + j * n + i))
    p = np.ones(4 * n)
    terms = [(1.0, [0, 1, 0, 0]), (-1.0, [0, 0, 1, 0])]
    _, g = sv.adjoint_value_and_grad(n, ops, p, terms)
    np.testing.assert_allclose(g[0], -0.41609, atol=1e-5)


def test_plain_torch_baseline_matches_oracle():
    """The timed CPU baseline (torch tensordot + autograd) computes the same
    HEA/TFIM energy and gradient as the complex128 oracle."""
    torch = pytest.importorskip("torch")
    from oracle import plain_torch

    n, L = 5, 2
    rng = np.random.default_rng(1001)
    p = rng.uniform(0, 2 * np.pi, size=(2 * L, n))
    e_ref, g_ref = sv.adjoint_value_and_grad(n, sv.hea_ops(n, L), p.reshape(-1), sv.tfim_terms(n))
    pt = torch.tensor(p, dtype=torch.float64)
    g, e = torch.func.grad_and_value(
        lambda q: plain_torch.hea_tfim_energy(q, n, L, dtype=torch.complex128))(pt)
    np.testing.assert_allclose(float(e), e_ref, atol=1e-12)
    np.testing.assert_allclose(g.numpy().reshape(-1), g_ref, atol=1e-12)
    # the per-layer rematerialised form (the C4 CPU baseline) is the same function
    pr = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    er = plain_torch.hea_tfim_energy(pr, n, L, dtype=torch.complex128, remat=True)
    er.backward()
    np.testing.assert_allclose(float(er.detach()), e_ref, atol=1e-12)
    np.testing.assert_allclose(pr.grad.numpy().reshape(-1), g_ref, atol=1e-12)


def test_param_gate_closed_forms():
    """Each parametrised matrix equals an independent construction: scipy's
    expm of its generator, or a product of rotations up to the stated
    global phase (a sign or index slip in any factory fails here)."""
    from scipy.linalg import expm

    X, Y, Z = sv._X, sv._Y, sv._Z
    th, phi, lbd, alpha = 0.731, 0.2, -0.4, 1.1
    for name, gen in (("rx", X), ("ry", Y), ("rz", Z), ("rxx", np.kron(X, X)), ("ryy", np.kron(Y, Y)),
                      ("rzz", np.kron(Z, Z))):
        np.testing.assert_allclose(sv.gate_matrix(name, theta=th), expm(-0.5j * th * gen), atol=1e-12)
    rz = lambda a: expm(-0.5j * a * Z)  # noqa: E731
    ry = lambda a: expm(-0.5j * a * Y)  # noqa: E731
    np.testing.assert_allclose(sv.gate_matrix("u", theta=th, phi=phi, lbd=lbd),
                               np.exp(0.5j * (phi + lbd)) * rz(phi) @ ry(th) @ rz(lbd), atol=1e-12)
    np.testing.assert_allclose(sv.phase(th), np.exp(0.5j * th) * rz(th), atol=1e-12)
    nvec = (np.sin(alpha) * np.cos(phi), np.sin(alpha) * np.sin(phi), np.cos(alpha))
    gen = nvec[0] * X + nvec[1] * Y + nvec[2] * Z
    np.testing.assert_allclose(sv.r_gate(th, alpha, phi), expm(-1j * th * gen), atol=1e-12)
    np.testing.assert_allclose(sv.exp1(np.kron(Z, Z), th), expm(-1j * th * np.kron(Z, Z)), atol=1e-12)
    np.testing.assert_allclose(sv.expm_herm(gen, th), expm(-1j * th * gen), atol=1e-12)
