"""Measurement templates on the device: any_measurements /
any_local_measurements (templates/measurements.py:18-153) through the
per-term Pauli kernel, and sparse_expectation with its gradient
(templates/measurements.py:177-191) through the CSR kernels.

KATs: tests/test_templates.py:17-56 of the reference and the
any_measurements docstring example (templates/measurements.py:24-43);
random structures and tables against the oracle's expectation of
(x)_i (sum_k s[i, k] P_k), which the oracle's own tests pin."""
import numpy as np
import pytest
import torch

from oracle import statevector as sv
from oracle.statevector import Op

pytestmark = pytest.mark.gpu

PAULI = [np.eye(2), sv.CONST_GATES["x"], sv.CONST_GATES["y"], sv.CONST_GATES["z"]]


def _oracle_table_expectation(state, n, table):
    ops = [(sum(table[i, k] * PAULI[k] for k in range(4)), [i]) for i in range(n)]
    return float(np.real(sv.expectation(state, n, *ops)))


def _random_circuit(n, seed):
    import tensorcircuit_amd as tc

    rng = np.random.default_rng(seed)
    c = tc.Circuit(n)
    ops = []
    for layer in range(3):
        for i in range(n):
            a, b = rng.uniform(0, 2 * np.pi, 2)
            c.ry(i, theta=a)
            c.rz(i, theta=b)
            ops += [Op("ry", (i,), kw={"theta": a}), Op("rz", (i,), kw={"theta": b})]
        for i in range(n - 1):
            c.cnot(i, i + 1)
            ops.append(Op("cnot", (i, i + 1)))
    return c, sv.run_ops(n, ops)


def test_any_measurements_kats(gpu):
    """tests/test_templates.py:17-40."""
    import tensorcircuit_amd as tc

    M = tc.templates.measurements
    c = tc.Circuit(2)
    c.H(0)
    c.H(1)
    r = M.any_measurements(c, np.array([1, 1]), onehot=True)
    assert r.shape == () and abs(float(r) - 1.0) < 1e-5
    r2 = M.any_measurements(c, np.array([3, 0]), onehot=True)
    assert abs(float(r2)) < 1e-5
    c = tc.Circuit(3)
    c.X(0)
    c.cnot(0, 1)
    c.H(-1)
    basis = tc.backend.convert_to_tensor(np.array([3, 3, 1]))
    r = M.parameterized_local_measurements(c, structures=basis, onehot=True)
    np.testing.assert_allclose(r.cpu().numpy(), [-1, -1, 1], atol=1e-5)


def test_any_measurements_docstring_example(gpu):
    """templates/measurements.py:24-43: <Z0 X2> three ways."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(3)
    c.rx(0, theta=1.0)
    c.cnot(0, 1)
    c.cnot(1, 2)
    c.ry(2, theta=-1.0)
    z0x2 = c.expectation([tc.gates.z(), [0]], [tc.gates.x(), [2]])
    p1 = tc.templates.measurements.parameterized_measurements(c, np.array([3, 0, 1]), onehot=True)
    p2 = tc.templates.measurements.parameterized_measurements(
        c, np.array([[0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]]), onehot=False)
    ref = np.real(sv.expectation(sv.run_ops(3, [Op("rx", (0,), kw={"theta": 1.0}), Op("cnot", (0, 1)),
                                                Op("cnot", (1, 2)), Op("ry", (2,), kw={"theta": -1.0})]), 3,
                                 (sv.CONST_GATES["z"], [0]), (sv.CONST_GATES["x"], [2])))
    for v in (np.real(complex(z0x2)), float(p1), float(p2)):
        assert abs(v - ref) < 1e-5


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_structure_stacks_vs_oracle(gpu, dtype):
    """Stacks of onehot structures and of general coefficient tables: one
    launch, each value against the oracle (1e-5 c64, 1e-12 c128)."""
    import tensorcircuit_amd as tc

    tc.set_dtype(dtype)
    try:
        n = 7
        tol = 1e-5 if dtype == "complex64" else 1e-12
        c, s = _random_circuit(n, 31)
        rng = np.random.default_rng(5)
        codes = rng.integers(0, 4, size=(12, n))
        vals = tc.templates.measurements.any_measurements(c, torch.as_tensor(codes), onehot=True)
        assert tuple(vals.shape) == (12,)
        for k in range(12):
            ref = np.real(sv.expectation_ps(s, n, ps=list(codes[k])))
            assert abs(float(vals[k]) - ref) < tol
        tabs = rng.standard_normal((4, n, 4)) * (rng.random((4, n, 4)) < 0.4)
        tabs[:, :, 0] += 1.0  # every wire non-zero
        vals = tc.templates.measurements.any_measurements(c, tabs)
        for k in range(4):
            ref = _oracle_table_expectation(s, n, tabs[k])
            assert abs(float(vals[k]) - ref) < tol * max(1.0, abs(ref)) * 10
        loc = tc.templates.measurements.any_local_measurements(c, torch.as_tensor(codes[:3]), onehot=True)
        assert tuple(loc.shape) == (3, n)
        for k in range(3):
            for i in range(n):
                cc = [0] * n
                cc[i] = int(codes[k, i])
                assert abs(float(loc[k, i]) - np.real(sv.expectation_ps(s, n, ps=cc))) < tol
    finally:
        tc.set_dtype("complex64")


def test_vmap_over_structures_is_one_call(gpu):
    """backend.vmap over the structures argument (the reference's use of
    any_measurements) equals the stacked call."""
    import tensorcircuit_amd as tc

    K = tc.backend
    n = 6
    c, s = _random_circuit(n, 7)
    codes = torch.as_tensor(np.random.default_rng(2).integers(0, 4, size=(9, n)))
    f = K.vmap(lambda st: tc.templates.measurements.any_measurements(c, st, onehot=True), vectorized_argnums=0)
    a = f(codes)
    b = tc.templates.measurements.any_measurements(c, codes, onehot=True)
    np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), atol=0)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_any_measurements_gradient(gpu, dtype):
    """value_and_grad through any_measurements equals the gradient through
    expectation_ps of the same strings (the adjoint path the oracle pins),
    and central differences of the oracle in complex128."""
    import tensorcircuit_amd as tc

    tc.set_dtype(dtype)
    try:
        K = tc.backend
        n, L = 5, 2
        codes = np.array([[3, 0, 1, 0, 2], [1, 1, 0, 3, 0], [0, 2, 2, 0, 3]])
        wts = np.array([0.7, -1.3, 0.4])

        def build(p):
            c = tc.Circuit(n)
            for j in range(L):
                for i in range(n):
                    c.ry(i, theta=p[2 * j, i])
                    c.rz(i, theta=p[2 * j + 1, i])
                for i in range(n - 1):
                    c.cnot(i, i + 1)
            return c

        w_t = torch.as_tensor(wts)

        def f_meas(p):
            v = tc.templates.measurements.any_measurements(build(p), torch.as_tensor(codes), onehot=True)
            return (v.to(torch.float64) * w_t.to(v.device)).sum()

        def f_ps(p):
            c = build(p)
            return sum(wts[k] * K.real(c.expectation_ps(ps=list(codes[k]))) for k in range(3))

        rdt = torch.float32 if dtype == "complex64" else torch.float64
        p = torch.as_tensor(np.random.default_rng(3).uniform(0, 2 * np.pi, (2 * L, n))).to(rdt)
        v1, g1 = K.value_and_grad(f_meas)(p)
        v2, g2 = K.value_and_grad(f_ps)(p)
        tol = 1e-5 if dtype == "complex64" else 1e-12
        assert abs(float(v1) - float(v2)) < tol
        np.testing.assert_allclose(g1.cpu().numpy(), g2.cpu().numpy(), atol=tol)
        if dtype == "complex128":
            terms = [(float(wts[k]), list(codes[k])) for k in range(3)]
            fd = sv.finite_difference_grad(n, sv.hea_ops(n, L), p.cpu().numpy().reshape(-1), terms)
            np.testing.assert_allclose(g1.cpu().numpy().reshape(-1), fd, atol=1e-6)
    finally:
        tc.set_dtype("complex64")


def test_sparse_expectation_value_and_grad(gpu):
    """tests/test_templates.py:43-56: <X> after rx(theta) then H is
    cos(theta), so the gradient is -sin(theta); also complex128 against the
    adjoint of the equivalent Pauli sum on a 6-qubit HEA."""
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import quantum as qu

    K = tc.backend
    ham = torch.sparse_coo_tensor(torch.tensor([[0, 1], [1, 0]]), torch.ones(2, dtype=torch.complex64), (2, 2))

    def f(param):
        c = tc.Circuit(1)
        c.rx(0, theta=param[0])
        c.H(0)
        return tc.templates.measurements.sparse_expectation(c, ham)

    for th in (0.0, 0.3, -1.1):
        v, g = K.jit(K.value_and_grad(f))(torch.tensor([th]))
        assert abs(float(v) - np.cos(th)) < 1e-6
        assert abs(float(g[0]) + np.sin(th)) < 1e-6

    tc.set_dtype("complex128")
    try:
        n, L = 6, 2
        rng = np.random.default_rng(8)
        ls = rng.integers(0, 4, size=(7, n))
        w = rng.standard_normal(7)
        hs = qu.PauliStringSum2COO(ls, w)

        def energy_sparse(p):
            c = tc.Circuit(n)
            tc.templates.blocks.hea_layers(c, p, L)
            return tc.templates.measurements.sparse_expectation(c, hs)

        p = torch.as_tensor(rng.uniform(0, 2 * np.pi, (2 * L, n)))
        v, g = K.value_and_grad(energy_sparse)(p)
        terms = [(float(w[k]), list(ls[k])) for k in range(7)]
        e_ref, g_ref = sv.adjoint_value_and_grad(n, sv.hea_ops(n, L), p.numpy().reshape(-1), terms)
        assert abs(float(v) - e_ref) < 1e-12
        np.testing.assert_allclose(g.cpu().numpy().reshape(-1), g_ref, atol=1e-11)
    finally:
        tc.set_dtype("complex64")
