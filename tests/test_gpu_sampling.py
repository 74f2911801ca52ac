"""Measurement on the device (tcm_probability, tcm_sample, tcm_sample_prob,
tcm_measure) against the oracle (oracle/sampling.py) on the same seeded
states and uniforms, and the Circuit / quantum APIs built on them.

Exact index parity for draws whose r lies farther than 1e-9 (relative) from
a cumulative-sum boundary (the device sums in fp64 in a different order);
probabilities of records to 1e-6 (complex64) / 1e-12 (complex128).
"""
import numpy as np
import pytest
import torch

from oracle import sampling as osm

pytestmark = pytest.mark.gpu


def _state(rng, n, batch, dtype, sparse=False):
    psi = rng.standard_normal((batch, 2**n)) + 1j * rng.standard_normal((batch, 2**n))
    if sparse:
        psi *= rng.random((batch, 2**n)) > 0.7
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    t = torch.as_tensor(psi.astype(np.complex64 if dtype == "complex64" else np.complex128))
    return t, t.numpy().astype(np.complex128)


@pytest.mark.parametrize("n", [1, 3, 7, 12, 17])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_sample_indices_vs_oracle(gpu, n, dtype):
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(100 + n)
    batch, shots = 3, 4000
    t, psi = _state(rng, n, batch, dtype, sparse=n >= 7)
    u = rng.random((batch, shots))
    u[:, :3] = [0.0, 1.0, 0.5]
    got = engine.sample_indices(t.to(gpu), n, torch.as_tensor(u)).cpu().numpy()
    for b in range(batch):
        p = np.abs(psi[b]) ** 2
        want = osm.probability_sample(p, u[b])
        m = osm.boundary_margin(p, u[b]) > 1e-9
        assert m.mean() > 0.99
        np.testing.assert_array_equal(got[b][m], want[m])
        assert np.all(got[b] >= 0) and np.all(got[b] < 2**n)
        # away from boundaries a zero-probability index is never drawn (u = 1
        # gives r = 0, whose left search returns index 0 whatever p_0 is)
        assert np.all(p[got[b][m]] > 0)


def test_sample_prob_vs_oracle_and_frequencies(gpu):
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(7)
    p = np.array([0.05] * 8 + [0.2, 0.4] + [0.0] * 6, dtype=np.float32)
    u = rng.random(20000)
    got = engine.sample_prob_indices(torch.as_tensor(p).to(gpu), 4, torch.as_tensor(u)).cpu().numpy()[0]
    want = osm.probability_sample(p, u)
    m = osm.boundary_margin(p, u) > 1e-9
    np.testing.assert_array_equal(got[m], want[m])
    counts = np.bincount(got, minlength=16)
    np.testing.assert_allclose(counts - p * 20000, 0, atol=400)  # tests/test_backends.py:287-291
    assert counts[10:].sum() == 0


def test_sample_large_state_properties(gpu):
    """24 qubits: a basis state always samples to itself; a product state's
    per-qubit marginals match over 2^18 draws (4 sigma)."""
    from tensorcircuit_amd import engine

    n = 24
    e = torch.zeros(1, 2**n, dtype=torch.complex64, device=gpu)
    e[0, 123457] = 1.0
    u = torch.rand(1, 1000, dtype=torch.float64)
    assert (engine.sample_indices(e, n, u) == 123457).all()
    # product state: qubit q has P(1) = sin^2(a_q / 2)
    import tensorcircuit_amd as tc

    a = np.linspace(0.3, 2.8, n)
    c = tc.Circuit(n)
    for q in range(n):
        c.ry(q, theta=float(a[q]))
    s = c.state().reshape(1, -1)
    shots = 1 << 18
    idx = engine.sample_indices(s, n, torch.rand(1, shots, dtype=torch.float64)).cpu().numpy()[0]
    for q in range(n):
        f = ((idx >> (n - 1 - q)) & 1).mean()
        pq = np.sin(a[q] / 2) ** 2
        assert abs(f - pq) < 4 * np.sqrt(pq * (1 - pq) / shots) + 1e-6


def test_probability_native(gpu):
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(1)
    for dtype in ("complex64", "complex128"):
        t, psi = _state(rng, 11, 2, dtype)
        p = engine.probability(t.to(gpu), 11).cpu().numpy()
        np.testing.assert_allclose(p, np.abs(psi) ** 2, rtol=1e-6 if dtype == "complex64" else 1e-14, atol=1e-12)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_measure_vs_oracle(gpu, dtype):
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(11)
    n, batch = 9, 4
    t, psi = _state(rng, n, batch, dtype)
    tol = 1e-6 if dtype == "complex64" else 1e-12
    for order in ([0, 1, 2, 3, 4, 5, 6, 7, 8], [4, 0, 8, 2], [8], [3, 5, 1, 0, 7, 6, 2, 4, 8]):
        st = rng.random((batch, len(order)))
        outc, prob = engine.measure(t.to(gpu), n, order, torch.as_tensor(st))
        outc, prob = outc.cpu().numpy(), prob.cpu().numpy()
        for b in range(batch):
            want, pw = osm.measure_jit(psi[b], order, st[b])
            np.testing.assert_array_equal(outc[b], want)
            assert abs(prob[b] - pw) < tol


def test_measure_rejects_bad_qubits(gpu):
    from tensorcircuit_amd import engine

    s = torch.zeros(1, 8, dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError):
        engine.measure(s, 3, [0, 0], torch.zeros(1, 2))
    with pytest.raises(ValueError):
        engine.measure(s, 3, [3], torch.zeros(1, 1))


def test_circuit_measure_and_perfect_sampling(gpu):
    import tensorcircuit_amd as tc

    c = tc.Circuit(3)
    c.h(0)
    c.cnot(0, 1)
    c.rx(2, theta=0.7)
    psi = c.state().cpu().numpy().astype(np.complex128)
    for st in ([0.1, 0.9, 0.3], [0.7, 0.2, 0.95], [0.5, 0.5, 0.5]):
        r, p = c.measure(0, 2, with_prob=True, status=torch.tensor(st))
        want, pw = osm.measure_jit(psi, [0, 2], st)
        np.testing.assert_array_equal(r.cpu().numpy(), want)
        assert abs(float(p) - pw) < 1e-6
        bits, p = c.perfect_sampling(torch.tensor(st))
        want, pw = osm.measure_jit(psi, [0, 1, 2], st)
        np.testing.assert_array_equal(bits.cpu().numpy(), want)
        assert bits[0] == bits[1]  # Bell pair on qubits 0, 1


def test_sample_formats(gpu):
    """tests/test_circuit.py:1409-1440: every format, with and without
    allow_state, batch None / 1 / 3."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(2)
    c.H(0)
    c.cnot(0, 1)
    g = torch.Generator().manual_seed(42)
    for allow_state in (False, True):
        for batch in (None, 1, 3):
            for fmt in (None, "sample_int", "sample_bin", "count_vector", "count_tuple", "count_dict_bin",
                        "count_dict_int"):
                r = c.sample(batch=batch, allow_state=allow_state, format=fmt, random_generator=g)
                if fmt is None:
                    items = [r] if batch is None else r
                    for bits, p in items:
                        assert int(bits[0]) == int(bits[1])
                        assert abs(float(p) - 0.5) < 1e-6
                elif fmt == "count_vector":
                    v = r.cpu().numpy()
                    assert v[1] == 0 and v[2] == 0 and v.sum() == (1 if batch is None else batch)
                elif fmt == "count_dict_bin":
                    assert set(r) <= {"00", "11"} or all(r[k] == 0 for k in ("01", "10") if k in r)
    r = c.sample(batch=8, allow_state=True, status=np.linspace(0.01, 0.99, 8), format="sample_bin")
    assert r.shape == (8, 2)
    # status maps through the inverse CDF: u > 0.5 -> r < half -> "00"
    np.testing.assert_array_equal(r.cpu().numpy()[:, 0], [1, 1, 1, 1, 0, 0, 0, 0])


def test_sample_expectation_ps(gpu):
    """tests/test_circuit.py:1381-1405."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(1, inputs=torch.tensor(np.array([1.0, 1.0j]) / np.sqrt(2), dtype=torch.complex64))
    assert abs(float(c.sample_expectation_ps(y=[0])) - 1.0) < 1e-5
    c = tc.Circuit(4)
    c.H(0)
    c.cnot(0, 1)
    c.rx(1, theta=0.3)
    c.rz(2, theta=-1.2)
    c.ccnot(2, 3, 1)
    c.rzz(0, 3, theta=0.5)
    c.ry(3, theta=2.2)
    c.s(1)
    c.td(2)
    y = c.sample_expectation_ps(x=[1], y=[0], z=[2, 3])
    ye = c.expectation_ps(x=[1], y=[0], z=[2, 3])
    assert abs(float(y) - float(torch.real(ye))) < 1e-5
    torch.manual_seed(0)
    y2 = c.sample_expectation_ps(x=[1], y=[0], z=[2, 3], shots=81920)
    assert abs(float(y2) - float(y)) < 0.01


def test_sample_expectation_readout_vs_oracle(gpu):
    import tensorcircuit_amd as tc

    c = tc.Circuit(3)
    c.h(0)
    c.cnot(0, 1)
    c.ry(2, theta=0.4)
    ro = [[0.9, 0.75], [0.4, 0.7], [0.8, 0.95]]
    v = c.sample_expectation_ps(z=[0, 1, 2], readout_error=ro)
    p = np.abs(c.state().cpu().numpy().astype(np.complex128)) ** 2
    want = osm.correlation_from_counts([0, 1, 2], osm.readout(p, ro))
    assert abs(float(v) - want) < 1e-6


def test_probability_gradient(gpu):
    """d/dtheta of sum_i w_i p_i through the native probability kernel vs
    the parameter-shift rule."""
    import tensorcircuit_amd as tc

    w = torch.linspace(-1, 1, 8, device=gpu)

    def f(th):
        c = tc.Circuit(3)
        c.h(0)
        c.rx(1, theta=th[0])
        c.cnot(1, 2)
        c.ry(2, theta=th[1])
        return torch.sum(w * c.probability())

    th = torch.tensor([0.4, -1.1], device=gpu, requires_grad=True)
    f(th).backward()
    g = th.grad.cpu().numpy()
    for i in range(2):
        e = np.zeros(2)
        e[i] = np.pi / 2
        gp = float(f(torch.tensor(th.detach().cpu().numpy() + e, device=gpu)))
        gm = float(f(torch.tensor(th.detach().cpu().numpy() - e, device=gpu)))
        assert abs(g[i] - (gp - gm) / 2) < 1e-5


def test_backend_probability_sample(gpu):
    """tests/test_backends.py:282-291: a length-10 p, 10000 draws."""
    from tensorcircuit_amd import backend as K

    p = torch.tensor([0.05] * 8 + [0.2, 0.4], dtype=torch.float32)
    st = np.random.default_rng(4).uniform(size=[10000])
    r = K.probability_sample(10000, p, status=st).cpu().numpy()
    _, cnt = np.unique(r, return_counts=True)
    np.testing.assert_allclose(cnt - p.numpy() * 10000.0, np.zeros([10]), atol=200, rtol=1)
    np.testing.assert_array_equal(r[osm.boundary_margin(p.numpy(), st) > 1e-9],
                                  osm.probability_sample(p.numpy(), st)[osm.boundary_margin(p.numpy(), st) > 1e-9])
    g = K.get_random_state(42)
    a = K.probability_sample(100, p, g=g)
    assert a.shape == (100,) and int(a.max()) < 10
