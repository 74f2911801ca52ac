"""Sparse Hamiltonians: tcm_pauli_coo (PauliStringSum2COO / ps2coo_core,
quantum.py:1304-1485) and tcm_csr_expect (sparse_expectation,
templates/measurements.py:177-191) against dense Kronecker products of the
oracle; the 6-site Heisenberg ground energy KAT (tests/test_quantum.py:
316-321)."""
import numpy as np
import pytest
import torch

from oracle import statevector as sv

pytestmark = pytest.mark.gpu


def _random_terms(rng, n, nterms):
    ls = rng.integers(0, 4, size=(nterms, n))
    ls[0] = 0  # an identity term
    w = rng.standard_normal(nterms)
    return ls, w


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_pauli_coo_matches_kron(gpu, dtype):
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import quantum as qu

    tc.set_dtype(dtype)
    try:
        rng = np.random.default_rng(4)
        n = 5
        ls, w = _random_terms(rng, n, 9)
        ls[3] = ls[2]  # a repeated string: coordinates add
        dense = sum(wk * sv.pauli_string_matrix(l) for wk, l in zip(w, ls))
        m = qu.PauliStringSum2COO(ls, w)
        assert m.is_sparse and m.device.type == "cuda"
        tol = 1e-6 if dtype == "complex64" else 1e-13
        np.testing.assert_allclose(m.to_dense().cpu().numpy(), dense, atol=tol)
        np.testing.assert_allclose(qu.PauliStringSum2Dense(ls, w, numpy=True), dense, atol=tol)
        sp = qu.PauliStringSum2COO(ls, w, numpy=True)
        np.testing.assert_allclose(sp.toarray(), dense, atol=tol)
        one = qu.PauliString2COO([1, 2, 0, 3, 2], 0.7)
        np.testing.assert_allclose(one.to_dense().cpu().numpy(), 0.7 * sv.pauli_string_matrix([1, 2, 0, 3, 2]),
                                   atol=tol)
    finally:
        tc.set_dtype("complex64")


def test_heisenberg_ground_energy_kat(gpu):
    import networkx as nx

    from tensorcircuit_amd import quantum as qu

    # templates/graphs.py Line1D(6) is periodic: the 6-site ring
    g = nx.cycle_graph(6)
    h = qu.heisenberg_hamiltonian(g, sparse=False)
    e = np.linalg.eigvalsh(h.cpu().numpy().astype(np.complex128))
    assert abs(e[0] - (-11.2111025)) < 1e-5
    hs = qu.heisenberg_hamiltonian(g, hz=0.3, hx=-0.2, sparse=True)
    ref = sum(wk * sv.pauli_string_matrix(l) for wk, l in zip(
        [1.0] * 18 + [0.3, -0.2] * 6,
        [r for e_ in g.edges for r in ([3 if i in e_ else 0 for i in range(6)], [1 if i in e_ else 0 for i in range(6)],
                                        [2 if i in e_ else 0 for i in range(6)])]
        + [r for v in g.nodes for r in ([3 if i == v else 0 for i in range(6)], [1 if i == v else 0 for i in range(6)])]))
    np.testing.assert_allclose(hs.to_dense().cpu().numpy(), ref, atol=1e-6)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_sparse_expectation_vs_oracle(gpu, dtype):
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine, quantum as qu
    from tensorcircuit_amd.templates import measurements as ms

    tc.set_dtype(dtype)
    try:
        rng = np.random.default_rng(12)
        n = 9
        ls, w = _random_terms(rng, n, 12)
        psi = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
        psi /= np.linalg.norm(psi)
        c = tc.Circuit(n, inputs=torch.as_tensor(psi.astype(np.complex64 if dtype == "complex64" else np.complex128)))
        want = sv.pauli_sum_expectation(psi, n, list(zip(w, ls))).real
        h = qu.PauliStringSum2COO(ls, w)
        tol = 1e-5 if dtype == "complex64" else 1e-12
        assert abs(float(ms.sparse_expectation(c, h)) - want) < tol
        assert abs(float(ms.operator_expectation(c, h.to_sparse_csr())) - want) < tol
        assert abs(float(ms.operator_expectation(c, engine.PauliSum(n, ls, w))) - want) < tol
        assert abs(float(ms.operator_expectation(c, h.to_dense())) - want) < tol
    finally:
        tc.set_dtype("complex64")
