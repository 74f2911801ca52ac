"""GPU parity of the HIP path against the complex128 CPU oracle.

Tolerances (from the target): complex64 results within 1e-5 (relative to
the state norm / energy scale), complex128 within 1e-12 (1e-10 for values
accumulated over hundreds of gates), gradients max|d| <= 1e-5 (c64).
"""
import numpy as np
import pytest
import torch

from oracle import statevector as sv

pytestmark = pytest.mark.gpu

C64_TOL = 1e-5
C128_TOL = 1e-10


@pytest.fixture(autouse=True)
def _reset_dtype():
    import tensorcircuit_amd as tc

    tc.set_dtype("complex64")
    yield
    tc.set_dtype("complex64")


def np_of(t):
    return t.detach().cpu().numpy()


# ------------------------------------------------------------------ KATs ---

def test_wavefunction_index_order(gpu):
    """tests/test_circuit.py:21-43."""
    import tensorcircuit_amd as tc

    m = np.arange(16).reshape(2, 2, 2, 2).astype(np.complex64)
    qc = tc.Circuit(2)
    qc.unitary(0, 1, unitary=tc.gates.Gate(m))
    assert np.real(np_of(qc.wavefunction())[2]) == 8
    qc = tc.Circuit(2)
    qc.unitary(1, 0, unitary=tc.gates.Gate(m))
    assert np.real(np_of(qc.wavefunction())[2]) == 4
    qc = tc.Circuit(2)
    qc.unitary(0, unitary=tc.gates.Gate(np.arange(4).reshape(2, 2).astype(np.complex64)))
    assert np.real(np_of(qc.wavefunction())[2]) == 2


def test_basics_amplitude(gpu):
    """tests/test_circuit.py:46-52."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(2)
    c.x(0)
    np.testing.assert_allclose(np_of(c.amplitude("10")), 1.0)
    c.CNOT(0, 1)
    np.testing.assert_allclose(np_of(c.amplitude("11")), 1.0)


def test_gates_in_circuit_and_controls(gpu):
    """tests/test_circuit.py:64-84."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(2, inputs=np.eye(2**2))
    c.iswap(0, 1)
    np.testing.assert_allclose(np_of(c.state()).reshape(4, 4), sv.iswap(), atol=1e-5)
    c = tc.Circuit(2)
    c.x(1)
    c.crx(1, 0, theta=0.3)
    np.testing.assert_allclose(np_of(c.expectation([tc.gates._z_matrix, 0])), 0.95533645, atol=1e-5)
    c = tc.Circuit(1)
    c.X(0)
    c.SD(0)
    np.testing.assert_allclose(np_of(c.state()), np.array([0.0, -1.0j]), atol=1e-7)


def test_expectation_kats(gpu):
    """tests/test_circuit.py:239-245, 370-390, 415-429, 1278-1281, 1373-1380."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(2)
    c.H(0)
    np.testing.assert_allclose(np_of(c.expectation((tc.gates.z(), [0]))), 0, atol=1e-7)
    c = tc.Circuit(2)
    c.X(0)
    np.testing.assert_allclose(np_of(c.expectation_ps(z=[0, 1])), -1, atol=1e-5)
    c = tc.Circuit(2)
    c.H(0)
    np.testing.assert_allclose(np_of(c.expectation_ps(z=[1], x=[0])), 1, atol=1e-5)
    np.testing.assert_allclose(np_of(c.expectation_ps(ps=[1, 3])), 1, atol=1e-5)
    np.testing.assert_allclose(np_of(c.expectation_ps(z=[1, 2], ps=[1, 3])), 1, atol=1e-5)
    c = tc.Circuit(1, inputs=1 / np.sqrt(2) * np.array([-1, 1.0j]))
    np.testing.assert_allclose(np_of(c.expectation_ps(y=[0])), -1, atol=1e-5)
    c = tc.Circuit(3)
    c.H(-2)
    c.H(0)
    np.testing.assert_allclose(np_of(c.expectation_ps(x=[0])).real, 1, atol=1e-5)
    np.testing.assert_allclose(np_of(c.expectation_ps(x=[1])).real, 1, atol=1e-5)
    np.testing.assert_allclose(np_of(c.expectation_ps(x=[-1])).real, 0, atol=1e-5)
    np.testing.assert_allclose(np_of(c.expectation_ps(z=[-2])).real, 0, atol=1e-5)
    for inp, want in [([0, 0, 0, 1.0], 1.0), ([0, 0, 1.0, 0], 1.0), ([1.0, 0, 0, 0], -1.0)]:
        c = tc.Circuit(2, inputs=tc.array_to_tensor(np.array(inp)))
        c.X(0)
        assert np_of(c.expectation((tc.gates.z(), [0]))) == want


def test_gate_kats(gpu):
    """tests/test_gates.py:18-22, 80-90, 93-96, 111-124, 147-156."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(1)
    c.h(0)
    c.phase(0, theta=np.pi / 2)
    np.testing.assert_allclose(np_of(c.state())[1], 0.7071j, atol=1e-4)
    c = tc.Circuit(2)
    c.exp(0, 1, unitary=np.array([[1.0, 0, 0, 0], [0, -1.0, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]]),
          theta=np.pi / 2)
    np.testing.assert_allclose(np_of(c.wavefunction())[0], -1j, atol=1e-6)
    c = tc.Circuit(2)
    c.any(0, unitary=np.eye(2))
    np.testing.assert_allclose(np_of(c.expectation((tc.gates.z(), [0]))), 1.0)
    ocx = sv._octrl(sv.CONST_GATES["cnot"])
    c = tc.Circuit(3)
    c.x(0)
    c.any(1, 0, 2, unitary=ocx)
    np.testing.assert_allclose(np_of(c.expectation([tc.gates.z(), [2]])), -1, atol=1e-5)
    c1 = tc.Circuit(3)
    c1.rxx(0, 1, theta=1.0)
    c1.ryy(0, 2, theta=0.5)
    c1.rzz(0, 1, theta=-0.5)
    c2 = tc.Circuit(3)
    c2.exp1(0, 1, theta=1.0 / 2, unitary=tc.gates._xx_matrix)
    c2.exp1(0, 2, theta=0.5 / 2, unitary=tc.gates._yy_matrix)
    c2.exp1(0, 1, theta=-0.5 / 2, unitary=tc.gates._zz_matrix)
    np.testing.assert_allclose(np_of(c1.state()), np_of(c2.state()), atol=1e-5)


def test_unitary_kron_inputs(gpu):
    """tests/test_circuit.py:404-412."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(2, inputs=np.eye(4))
    c.X(0)
    c.Y(1)
    np.testing.assert_allclose(np_of(c.wavefunction()).reshape(4, 4), np.kron(sv._X, sv._Y), atol=1e-5)


# ------------------------------------------------------- random circuits ---

NAMES_1Q = ["h", "x", "y", "z", "s", "t", "sd", "td", "rx", "ry", "rz", "phase", "wroot"]
NAMES_2Q = ["cnot", "cz", "swap", "cy", "rzz", "rxx", "ryy", "iswap", "crx", "unitary2"]


def _random_circuit(rng, n, ngates, with_3q=True):
    ops = []
    for _ in range(ngates):
        r = rng.random()
        if with_3q and n >= 3 and r < 0.04:
            qs = tuple(int(q) for q in rng.permutation(n)[:3])
            ops.append(("toffoli", qs, {}))
        elif r < 0.55:
            ops.append((str(rng.choice(NAMES_1Q)), (int(rng.integers(n)),), {"theta": float(rng.uniform(0, 2 * np.pi))}))
        else:
            qs = tuple(int(q) for q in rng.permutation(n)[:2])
            name = str(rng.choice(NAMES_2Q))
            kw = {"theta": float(rng.uniform(0, 2 * np.pi))}
            if name == "unitary2":
                u = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
                kw = {"unitary": u}
            ops.append((name, qs, kw))
    return ops


def _build(tc, n, ops):
    c = tc.Circuit(n)
    for name, qs, kw in ops:
        if name in ("h", "x", "y", "z", "s", "t", "sd", "td", "wroot", "cnot", "cz", "swap", "cy", "toffoli"):
            getattr(c, name)(*qs)
        elif name == "unitary2":
            c.unitary(*qs, unitary=kw["unitary"])
        else:
            getattr(c, name)(*qs, theta=kw["theta"])
    return c


def _oracle_ops(ops):
    out = []
    for name, qs, kw in ops:
        if name == "unitary2":
            out.append(sv.Op("unitary", qs, kw={"unitary": kw["unitary"]}))
        elif name in ("h", "x", "y", "z", "s", "t", "sd", "td", "wroot", "cnot", "cz", "swap", "cy", "toffoli"):
            out.append(sv.Op(name, qs))
        else:
            out.append(sv.Op(name, qs, kw={"theta": kw["theta"]}))
    return out


@pytest.mark.parametrize("n", [3, 6, 10, 12, 14])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("fuse", [0, 1])
def test_random_circuit_state(gpu, n, dtype, fuse):
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine

    tc.set_dtype(dtype)
    rng = np.random.default_rng(1000 + n)
    ops = _random_circuit(rng, n, 6 * n)
    c = _build(tc, n, ops)
    specs, thetas = c._lower()
    prog = engine.get_program(n, specs, len(thetas), fuse=fuse)
    params = c._params(thetas, gpu)
    st, _ = prog.forward(params.reshape(1, -1), 1, dtype)
    ref = sv.run_ops(n, _oracle_ops(ops))
    tol = C64_TOL if dtype == "complex64" else C128_TOL
    np.testing.assert_allclose(np_of(st).reshape(-1), ref, atol=tol * 4)
    # the drop-in path (fused default) agrees too
    np.testing.assert_allclose(np_of(c.wavefunction()), ref, atol=tol * 4)


# --------------------------------------------------------------- HEA/TFIM ---

def _hea_params(seed, L, n, batch=None):
    rng = np.random.default_rng(seed)
    shape = (2 * L, n) if batch is None else (batch, 2 * L, n)
    return rng.uniform(0, 2 * np.pi, size=shape).astype(np.float32)


def _tfim_codes(n, h=1.0):
    terms = sv.tfim_terms(n, h)
    return np.array([c for _, c in terms], dtype=np.int8), np.array([w for w, _ in terms])


@pytest.mark.parametrize("n,L", [(12, 4), (16, 3)])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_hea_energy_and_grad(gpu, n, L, dtype):
    """C1-shaped workload: fused adjoint energy + gradient vs the c128 oracle."""
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    tc.set_dtype(dtype)
    p = _hea_params(1001, L, n)
    c = tc.Circuit(n)
    hea_layers(c, p, L)
    specs, thetas = c._lower()
    codes, coeffs = _tfim_codes(n)
    ham = engine.get_pauli(n, codes, coeffs)
    ops = sv.hea_ops(n, L)
    e_ref, g_ref = sv.adjoint_value_and_grad(n, ops, p.astype(np.float64).reshape(-1), sv.tfim_terms(n))
    tol = 1e-5 if dtype == "complex64" else C128_TOL
    for fuse, ck in ((0, 0), (1, 0), (1, 2)):
        prog = engine.get_program(n, specs, len(thetas), fuse=fuse, ckpt_every=ck)
        e, g = prog.energy_grad(torch.as_tensor(p.reshape(1, -1)), ham, 1, dtype)
        assert abs(np_of(e)[0] - e_ref) <= tol * max(1.0, abs(e_ref)), (fuse, ck, np_of(e)[0], e_ref)
        assert np.max(np.abs(np_of(g).reshape(-1) - g_ref)) <= tol, (fuse, ck)


def test_hea_batch_matches_single(gpu):
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L, B = 12, 3, 5
    ps = _hea_params(7, L, n, batch=B)
    c = Circuit(n)
    hea_layers(c, ps[0], L)
    specs, thetas = c._lower()
    codes, coeffs = _tfim_codes(n)
    ham = engine.get_pauli(n, codes, coeffs)
    prog = engine.get_program(n, specs, len(thetas), fuse=1, ckpt_every=1)
    e, g = prog.energy_grad(torch.as_tensor(ps.reshape(B, -1)), ham, B)
    for b in range(B):
        e1, g1 = prog.energy_grad(torch.as_tensor(ps[b].reshape(1, -1)), ham, 1)
        np.testing.assert_allclose(np_of(e)[b], np_of(e1)[0], rtol=0, atol=1e-12)
        np.testing.assert_allclose(np_of(g)[b], np_of(g1)[0], rtol=0, atol=1e-12)
        e_ref, g_ref = sv.adjoint_value_and_grad(n, sv.hea_ops(n, L), ps[b].astype(np.float64).reshape(-1),
                                                 sv.tfim_terms(n))
        assert abs(np_of(e)[b] - e_ref) < 1e-5
        assert np.max(np.abs(np_of(g)[b] - g_ref)) < 1e-5


def test_zero_angles_closed_form_24q(gpu):
    """All-zero angles keep |0..0>: E = -(n-1) exactly; gradient of the
    rz angles vanishes (|0..0> is an rz eigenstate)."""
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L = 24, 2
    p = np.zeros((2 * L, n), dtype=np.float32)
    c = Circuit(n)
    hea_layers(c, p, L)
    specs, thetas = c._lower()
    codes, coeffs = _tfim_codes(n)
    ham = engine.get_pauli(n, codes, coeffs)
    prog = engine.get_program(n, specs, len(thetas), fuse=1, ckpt_every=1)
    e, g = prog.energy_grad(torch.as_tensor(p.reshape(1, -1)), ham, 1)
    assert abs(np_of(e)[0] + (n - 1)) < 1e-6
    g = np_of(g).reshape(2 * L, n)
    assert np.max(np.abs(g[1::2])) < 1e-6
    st, _ = prog.forward(torch.as_tensor(p.reshape(1, -1)), 1, "complex64")
    s = np_of(st).reshape(-1)
    assert abs(s[0] - 1) < 1e-6 and np.max(np.abs(s[1:])) < 1e-7


def test_norm_preserved_large(gpu):
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L = 22, 3
    p = _hea_params(1002, L, n)
    c = Circuit(n)
    hea_layers(c, p, L)
    specs, thetas = c._lower()
    prog = engine.get_program(n, specs, len(thetas), fuse=1)
    st, _ = prog.forward(torch.as_tensor(p.reshape(1, -1)), 1, "complex64")
    nrm = engine.inner(st, st, n).real
    assert abs(float(nrm[0]) - 1.0) < 1e-5


# ------------------------------------------------------- drop-in autograd ---

def test_interface_gradient_kats(gpu):
    """tests/test_interfaces.py:30-97 through backend.grad on the Circuit API."""
    import tensorcircuit_amd as tc

    K = tc.backend
    n = 4

    def f(param):
        c = tc.Circuit(n)
        c = tc.templates.blocks.example_block(c, param)
        return K.real(c.expectation([tc.gates.x(), [1]])) ** 2

    g = K.grad(f)(torch.ones([4, n]))
    np.testing.assert_allclose(np_of(g)[0, 1], -2.146e-3, atol=1e-5)

    def f2(paramzz, paramx):
        c = tc.Circuit(n)
        for i in range(n):
            c.H(i)
        for j in range(2):
            for i in range(n - 1):
                c.exp1(i, i + 1, unitary=tc.gates._zz_matrix, theta=paramzz[j, i])
            for i in range(n):
                c.rx(i, theta=paramx[j, i])
        return K.real(c.expectation([tc.gates.x(), [1]])) - K.real(c.expectation([tc.gates.x(), [2]]))

    g1, g2 = K.grad(f2, argnums=(0, 1))(torch.ones([2, n]), torch.ones([2, n]))
    np.testing.assert_allclose(np_of(g1)[0, 0], -0.41609, atol=1e-5)


def _vqe_energy(inputs, param, n, nlayers):
    """tests/test_backends.py:426-449."""
    import tensorcircuit_amd as tc

    c = tc.Circuit(n, inputs=inputs)
    paramc = tc.backend.cast(param, "complex64")
    for i in range(n):
        c.H(i)
    for j in range(nlayers):
        for i in range(n - 1):
            c.ryy(i, i + 1, theta=paramc[2 * j, i])
        for i in range(n):
            c.rx(i, theta=paramc[2 * j + 1, i])
    e = 0.0
    for i in range(n):
        e += c.expectation((tc.gates.x(), [i]))
    for i in range(n - 1):
        e += c.expectation((tc.gates.z(), [i]), (tc.gates.z(), [(i + 1) % n]))
    return tc.backend.real(e)


def test_vvag_consistency(gpu):
    """tests/test_backends.py:452-475: vvag == value_and_grad per sample."""
    from functools import partial

    import tensorcircuit_amd as tc

    K = tc.backend
    n, nlayers = 4, 3
    inp = K.ones([2**n]) / 2 ** (n / 2)
    param = K.ones([2 * nlayers, n])
    f = partial(_vqe_energy, n=n, nlayers=nlayers)
    v0, (g00, g01) = K.value_and_grad(f, argnums=(0, 1))(inp, param)
    batch = 8
    inps = K.cast(K.ones([batch, 2**n]) / 2 ** (n / 2), "complex64")
    v1, (g10, g11) = K.vvag(f, argnums=(0, 1))(inps, param)
    np.testing.assert_allclose(np_of(v1)[0], np_of(v0), atol=1e-4)
    np.testing.assert_allclose(np_of(g10)[0], np_of(g00), atol=1e-4)
    np.testing.assert_allclose(np_of(g11) / batch, np_of(g01), atol=1e-4)
    # against the oracle: same energy function in complex128
    ops = [sv.Op("h", (i,)) for i in range(n)]
    for j in range(nlayers):
        for i in range(n - 1):
            ops.append(sv.Op("ryy", (i, i + 1), pidx=(2 * j) * n + i))
        for i in range(n):
            ops.append(sv.Op("rx", (i,), pidx=(2 * j + 1) * n + i))
    terms = [(1.0, [1 if q == i else 0 for q in range(n)]) for i in range(n)]
    terms += [(1.0, [3 if q in (i, i + 1) else 0 for q in range(n)]) for i in range(n - 1)]
    e_ref, g_ref = sv.adjoint_value_and_grad(n, ops, np.ones(2 * nlayers * n), terms,
                                             inputs=np.ones(2**n) / 2 ** (n / 2))
    np.testing.assert_allclose(np_of(v0), e_ref, atol=1e-5)
    np.testing.assert_allclose(np_of(g01).reshape(-1), g_ref, atol=1e-5)


def test_grad_vs_finite_difference_c128(gpu, highp):
    """Drop-in value_and_grad in complex128 vs central differences of the oracle."""
    import tensorcircuit_amd as tc

    K = tc.backend
    n, L = 6, 2
    p = torch.as_tensor(_hea_params(3, L, n).astype(np.float64))

    def energy(param):
        c = tc.Circuit(n)
        tc.templates.blocks.hea_layers(c, param, L)
        e = 0.0
        for w, codes in sv.tfim_terms(n):
            e = e + w * c.expectation_ps(ps=codes)
        return K.real(e)

    v, g = K.value_and_grad(energy)(p)
    ops = sv.hea_ops(n, L)
    fd = sv.finite_difference_grad(n, ops, p.numpy().reshape(-1), sv.tfim_terms(n), eps=1e-3)
    ps = sv.parameter_shift_grad(n, ops, p.numpy().reshape(-1), sv.tfim_terms(n))
    np.testing.assert_allclose(np_of(g).reshape(-1), ps, atol=1e-10)
    np.testing.assert_allclose(np_of(g).reshape(-1), fd, atol=1e-5)
    np.testing.assert_allclose(np_of(v), sv.energy(n, ops, p.numpy().reshape(-1), sv.tfim_terms(n)), atol=1e-12)


def test_parameter_shift_grad_matches_adjoint(gpu, highp):
    """experimental.parameter_shift_grad (experimental.py:215-269): all
    shifted parameter sets in one vmapped batch equal the adjoint gradient
    and the oracle's shift rule (complex128)."""
    import tensorcircuit_amd as tc

    K = tc.backend
    n, L = 5, 2
    p = torch.as_tensor(_hea_params(7, L, n).astype(np.float64)).to("cuda")

    def energy(param):
        c = tc.Circuit(n)
        tc.templates.blocks.hea_layers(c, param, L)
        e = 0.0
        for w, codes in sv.tfim_terms(n):
            e = e + w * c.expectation_ps(ps=codes)
        return K.real(e)

    g_ps = tc.experimental.parameter_shift_grad(energy)(p)
    _, g_ad = K.value_and_grad(energy)(p)
    ref = sv.parameter_shift_grad(n, sv.hea_ops(n, L), p.cpu().numpy().reshape(-1), sv.tfim_terms(n))
    assert tuple(g_ps.shape) == (2 * L, n)
    np.testing.assert_allclose(np_of(g_ps).reshape(-1), ref, atol=1e-10)
    np.testing.assert_allclose(np_of(g_ps), np_of(g_ad), atol=1e-10)


# ----------------------------------------------------------------- Pauli ---

@pytest.mark.parametrize("jit", [0, 1])
def test_pauli_sums_vs_oracle(gpu, jit):
    """Generic (jit=0) and compiled (jit=1) tiled Pauli passes."""
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(99)
    for n, T in ((5, 9), (11, 20), (13, 7)):
        s = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
        s /= np.linalg.norm(s)
        codes = rng.integers(0, 4, size=(T, n)).astype(np.int8)
        w = rng.standard_normal(T)
        ham = engine.PauliSum(n, codes, w, jit=jit)
        terms = [(float(w[k]), list(codes[k])) for k in range(T)]
        for dt in (torch.complex64, torch.complex128):
            st = torch.as_tensor(s).to(dt).cuda()
            e = np_of(ham.expect(st))[0]
            ref = sv.pauli_sum_expectation(s.astype(np.complex64 if dt == torch.complex64 else np.complex128), n,
                                           terms)
            tol = 1e-5 if dt == torch.complex64 else 1e-12
            assert abs(e - ref) < tol * np.abs(w).sum(), (n, dt, e, ref)
            hs = np_of(ham.apply(st))[0]
            np.testing.assert_allclose(hs, sv.pauli_sum_apply(s, n, terms), atol=tol * 10 * np.abs(w).sum())


@pytest.mark.parametrize("jit", [0, 1])
def test_pauli_tfim_21q_wide_tiles(gpu, jit):
    """A 21-qubit TFIM-like sum (ZZ chain with random weights, X on every
    qubit): complex64 takes 14-bit tiles here (the 4 low bits plus 10 flip
    bits per pass: 2 passes instead of 3), the second one of scattered high
    bits; energies and H psi against the oracle, with a batch of 2."""
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(21)
    n = 21
    terms = [(float(rng.uniform(-1, 1)), [3 if q in (i, i + 1) else 0 for q in range(n)]) for i in range(n - 1)]
    terms += [(float(rng.uniform(-1, 1)), [1 if q == i else 0 for q in range(n)]) for i in range(n)]
    ham = engine.PauliSum(n, [t[1] for t in terms], [t[0] for t in terms], jit=jit)
    wsum = sum(abs(t[0]) for t in terms)
    s = rng.standard_normal((2, 2**n)) + 1j * rng.standard_normal((2, 2**n))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    for dt, tol in ((torch.complex64, 1e-5), (torch.complex128, 1e-12)):
        st = torch.as_tensor(s).to(dt).cuda()
        e = np_of(ham.expect(st))
        hs = np_of(ham.apply(st))
        for b in range(2):
            ref = sv.pauli_sum_expectation(s[b], n, terms)
            assert abs(e[b] - ref) < tol * wsum, (dt, b, e[b], ref)
            np.testing.assert_allclose(hs[b], sv.pauli_sum_apply(s[b], n, terms), atol=tol * 10 * wsum)


@pytest.mark.parametrize("jit", [0, 1])
def test_pauli_wide_terms_and_tiles(gpu, jit):
    """Terms wider than an LDS tile (X on every qubit) run beside tiled ones."""
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(5)
    n = 14
    s = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    s /= np.linalg.norm(s)
    codes = [[1] * n, [2] * n, [3] * n, [2, 1] + [0] * (n - 2), [0] * (n - 1) + [1]]
    codes += [list(rng.integers(0, 4, n)) for _ in range(6)]
    w = rng.standard_normal(len(codes))
    terms = [(float(w[k]), list(codes[k])) for k in range(len(codes))]
    ham = engine.PauliSum(n, codes, w, jit=jit)
    for dt, tol in ((torch.complex64, 1e-5), (torch.complex128, 1e-12)):
        st = torch.as_tensor(s).to(dt).cuda()
        e = np_of(ham.expect(st))[0]
        ref = sv.pauli_sum_expectation(s, n, terms)
        assert abs(e - ref) < tol * np.abs(w).sum() * 4, (dt, e, ref)
        alpha = torch.tensor([0.5], dtype=torch.float64, device="cuda")
        hs = np_of(ham.apply(st, alpha))[0]
        np.testing.assert_allclose(hs, 0.5 * sv.pauli_sum_apply(s, n, terms), atol=tol * 10 * np.abs(w).sum())


def test_pauli_compiled_batch(gpu):
    """Compiled Pauli passes on a batch of states: energies and alpha-scaled
    H psi per item against the oracle (complex128, 16 qubits, Y terms)."""
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(12)
    n, B, T = 16, 3, 12
    codes = rng.integers(0, 4, size=(T, n)).astype(np.int8)
    w = rng.standard_normal(T)
    terms = [(float(w[k]), list(codes[k])) for k in range(T)]
    ham = engine.PauliSum(n, codes, w, jit=1)
    s = rng.standard_normal((B, 2**n)) + 1j * rng.standard_normal((B, 2**n))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    st = torch.as_tensor(s).cuda()
    e = np_of(ham.expect(st))
    alpha = torch.tensor([0.5, -1.0, 2.0], dtype=torch.float64, device="cuda")
    hs = np_of(ham.apply(st, alpha))
    for b in range(B):
        assert abs(e[b] - sv.pauli_sum_expectation(s[b], n, terms)) < 1e-12 * np.abs(w).sum()
        np.testing.assert_allclose(hs[b], float(alpha[b]) * sv.pauli_sum_apply(s[b], n, terms), atol=1e-11)


def test_heisenberg_ring_energy(gpu):
    """tests/test_quantum.py:316-321: the 6-site ring ground energy -11.2111
    from <psi|H|psi> on the exact ground state."""
    from tensorcircuit_amd import engine

    n = 6
    codes, w = [], []
    for a in range(n):
        for code in (3, 1, 2):
            c = [0] * n
            c[a] = code
            c[(a + 1) % n] = code
            codes.append(c)
            w.append(1.0)
    hm = sum(wi * sv.pauli_string_matrix(ci) for wi, ci in zip(w, codes))
    evals, evecs = np.linalg.eigh(hm)
    ham = engine.PauliSum(n, codes, w)
    e = np_of(ham.expect(torch.as_tensor(evecs[:, 0]).to(torch.complex128).cuda()))[0]
    np.testing.assert_allclose(e.real, -11.2111, atol=1e-4)


# ----------------------------------------------------- MFMA 5-qubit blocks ---

@pytest.mark.parametrize("n", [10, 12, 15])
def test_mfma_dense5_vs_oracle(gpu, n):
    """Dense 5-qubit gates on FP32 matrix cores (every placement class: gate
    bits among the low bits, the high bits and mixed) vs the c128 oracle, and
    the same gates undone through the adjoint (U^dagger on the matrix cores)."""
    from tensorcircuit_amd import _lib, engine
    from tensorcircuit_amd.engine import GateSpec

    rng = np.random.default_rng(n)
    specs, ops = [], []
    placements = [tuple(range(5)), tuple(range(n - 5, n)), (0, n - 1, 2, n // 2, n - 3)]
    for k in range(6):
        qs = placements[k % 3] if k < 3 else tuple(int(q) for q in rng.permutation(n)[:5])
        u = np.linalg.qr(rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32)))[0]
        specs.append(GateSpec(_lib.G_MAT, qs, -1, 0.0, 1.0, u))
        ops.append(sv.Op("unitary", qs, kw={"unitary": u}))
    s0 = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    s0 /= np.linalg.norm(s0)
    ref = sv.run_ops(n, ops, inputs=s0)
    for mfma in (2, 1, 0):
        prog = engine.Program(n, specs, 0, fuse=1, mfma=mfma)
        st, _ = prog.forward(None, 2, "complex64", inputs=torch.as_tensor(np.stack([s0, s0])).to(torch.complex64))
        np.testing.assert_allclose(np_of(st)[1], ref, atol=2e-6)
        # the vector-Jacobian product with grad = final state maps back to U^dag psi = s0
        _, gi = prog.vjp(None, None, st, st, 2, "complex64", True)
        np.testing.assert_allclose(np_of(gi)[0], s0, atol=2e-6)


def test_brickwork_fused_mfma_vs_c128(gpu):
    """C5-shaped check at 16 qubits: fused 5-qubit blocks on FP32 MFMA vs the
    complex128 VALU run of the same blocks and vs the c128 oracle at 14 qubits."""
    from tensorcircuit_amd import engine, fusion, workloads

    for n, depth in ((14, 8), (16, 10)):
        gates = workloads.brickwork(n, depth, 1005)
        blocks = fusion.fuse(gates, n)
        p32 = engine.Program(n, blocks, 0, fuse=1, mfma=1)
        p64 = engine.Program(n, blocks, 0, fuse=1, mfma=0)
        s32, _ = p32.forward(None, 1, "complex64")
        s64, _ = p64.forward(None, 1, "complex128")
        fid = abs(complex(engine.inner(s64, s32.to(torch.complex128), n)[0])) ** 2
        assert abs(fid - 1.0) < 1e-5, fid
        # the same blocks grouped into LDS tile passes
        pt = engine.Program(n, blocks, 0, fuse=1, mfma=2)
        assert pt.npasses < p32.npasses
        st_, _ = pt.forward(None, 1, "complex64")
        np.testing.assert_allclose(np_of(st_), np_of(s32), atol=2e-6)
        if n == 14:
            ref = sv.zero_state(n)
            for g in gates:
                ref = sv.apply_gate(ref, g.mat, g.qubits, n)
            np.testing.assert_allclose(np_of(s64).reshape(-1), ref, atol=1e-12)
            np.testing.assert_allclose(np_of(s32).reshape(-1), ref, atol=2e-6)


def test_mfma_tile_passes_every_width(gpu):
    """Matrix-core tile passes at every tile width (10..13 bits) on a 20-qubit
    batch of 3 states: more tiles than resident blocks, so every block runs
    several tiles and loads the next one during the matrix-core work.  Twelve
    blocks on 5-subsets of 9 qubits share one 13-bit tile and are split into
    passes of 8 and 4; two more sit on the low qubits.  The result equals the
    complex128 VALU run of the same gates (the c128 path is pinned against the
    oracle above)."""
    import itertools

    from tensorcircuit_amd import _lib, engine
    from tensorcircuit_amd.engine import GateSpec

    n, batch = 20, 3
    rng = np.random.default_rng(20)
    subsets = list(itertools.combinations(range(9), 5))
    qsets = [tuple(int(q) for q in rng.permutation(subsets[i])) for i in rng.choice(len(subsets), 12, replace=False)]
    qsets += [tuple(range(n - 5, n)), (n - 1, 3, n - 7, 12, n - 2)]
    specs = []
    for qs in qsets:
        u = np.linalg.qr(rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32)))[0]
        specs.append(GateSpec(_lib.G_MAT, qs, -1, 0.0, 1.0, u))
    s0 = rng.standard_normal((batch, 2**n)) + 1j * rng.standard_normal((batch, 2**n))
    s0 /= np.linalg.norm(s0, axis=1, keepdims=True)
    p64 = engine.Program(n, specs, 0, fuse=1, mfma=0)
    ref = np_of(p64.forward(None, batch, "complex128", inputs=torch.as_tensor(s0))[0])
    for tb in (13, 12, 11, 10):
        pt = engine.Program(n, specs, 0, fuse=1, mfma=2, tile_bits=tb)
        passes = pt.describe("complex64", 0)["sched"]["passes"]
        assert all(p["mfma"] == 1 and len(p["tile"]) == tb for p in passes)
        assert max(len(p["gates"]) for p in passes) <= 8
        assert sorted(g for p in passes for g in p["gates"]) == list(range(len(specs)))
        if tb == 13:
            assert [len(p["gates"]) for p in passes][:2] == [8, 4]
        st, _ = pt.forward(None, batch, "complex64", inputs=torch.as_tensor(s0).to(torch.complex64))
        np.testing.assert_allclose(np_of(st), ref, atol=3e-6, err_msg="tile bits %d" % tb)


# ------------------------------------------------------ batched uneven (C3) ---

def test_forward_multi_uneven(gpu):
    """Circuits of 10..17 qubits (and a batch of one program) in pass-synchronous
    multi-program launches equal their single-program runs and the oracle."""
    from tensorcircuit_amd import Circuit, engine, workloads
    from tensorcircuit_amd.templates.blocks import hea_layers

    items, singles, refs = [], [], []
    for i, n in enumerate(range(10, 18)):
        L = 3
        c = Circuit(n)
        hea_layers(c, np.zeros((2 * L, n)), L, barrier=False)
        specs, thetas = c._lower()
        prog = engine.get_program(n, specs, len(thetas), fuse=1)
        b = 1 + i % 3
        p = torch.as_tensor(workloads.hea_params(2000 + n, L, n, batch=b).reshape(b, -1)).cuda()
        st = torch.empty(b, 1 << n, dtype=torch.complex64, device="cuda")
        items.append((prog, p, st))
        s1, _ = prog.forward(p, b, "complex64")
        singles.append(s1)
        if n <= 14:
            refs.append((n, sv.run_ops(n, sv.hea_ops(n, L), p[0].double().cpu().numpy().reshape(-1))))
    engine.forward_multi(items)
    for (prog, p, st), s1 in zip(items, singles):
        np.testing.assert_allclose(np_of(st), np_of(s1), atol=1e-6)
    for n, ref in refs:
        st = [it[2] for it in items if it[0].n == n][0]
        np.testing.assert_allclose(np_of(st)[0], ref, atol=2e-6)


# ------------------------------------------------------- torch interface ---

def test_torchlayer_quantumnet(gpu):
    """tests/test_torchnn.py:20-48: a TorchLayer over a 6-qubit classifier,
    batch of 3 inputs -> [3, 6]; values against the oracle and weight
    gradients against central differences of the oracle (complex128)."""
    import tensorcircuit_amd as tc

    K = tc.backend
    n, nlayers = 6, 2

    def qpred(x, weights):
        c = tc.Circuit(n)
        for i in range(n):
            c.rx(i, theta=x[i])
        for j in range(nlayers):
            for i in range(n - 1):
                c.cnot(i, i + 1)
            for i in range(n):
                c.rx(i, theta=weights[2 * j, i])
                c.ry(i, theta=weights[2 * j + 1, i])
        return K.real(K.stack([c.expectation_ps(x=[i]) for i in range(n)]))

    torch.manual_seed(3)
    ql = tc.TorchLayer(qpred, weights_shape=[2 * nlayers, n])
    x = torch.rand(3, n) * 2
    yp = ql(x)
    assert tuple(yp.shape) == (3, n)

    def oracle(xv, w):
        ops = [sv.Op("rx", (i,), kw={"theta": float(xv[i])}) for i in range(n)]
        for j in range(nlayers):
            ops += [sv.Op("cnot", (i, i + 1)) for i in range(n - 1)]
            for i in range(n):
                ops.append(sv.Op("rx", (i,), kw={"theta": float(w[2 * j, i])}))
                ops.append(sv.Op("ry", (i,), kw={"theta": float(w[2 * j + 1, i])}))
        s = sv.run_ops(n, ops)
        return np.array([np.real(sv.expectation_ps(s, n, x=[i])) for i in range(n)])

    w = ql.q_weights[0].detach().cpu().numpy().astype(np.float64)
    xs = x.numpy().astype(np.float64)
    ref = np.stack([oracle(xs[b], w) for b in range(3)])
    np.testing.assert_allclose(np_of(yp), ref, atol=2e-5)
    # gradient of sum(yp) w.r.t. the weights
    yp.sum().backward()
    g = ql.q_weights[0].grad.detach().cpu().numpy()
    eps = 1e-4
    fd = np.zeros_like(w)
    for idx in np.ndindex(*w.shape):
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd[idx] = sum((oracle(xs[b], wp) - oracle(xs[b], wm)).sum() for b in range(3)) / (2 * eps)
    np.testing.assert_allclose(g, fd, atol=2e-4)


def test_torch_interface_backward(gpu):
    """interfaces/torch.py docstring example: <Z> after rx(a0) ry(a1) is
    cos(a0) cos(a1); d(<Z>^2)/da is checked against the closed form."""
    import tensorcircuit_amd as tc

    K = tc.backend

    def f(params):
        c = tc.Circuit(1)
        c.rx(0, theta=params[0])
        c.ry(0, theta=params[1])
        return K.real(c.expectation((tc.gates.z(), [0])))

    f_torch = tc.interfaces.torch_interface(f, jit=True)
    a = torch.tensor([0.3, 0.7], requires_grad=True)
    b = f_torch(a)
    (b ** 2).backward()
    z = np.cos(0.3) * np.cos(0.7)
    np.testing.assert_allclose(float(b.detach()), z, atol=1e-6)
    np.testing.assert_allclose(a.grad.numpy(), [2 * z * -np.sin(0.3) * np.cos(0.7),
                                                2 * z * -np.cos(0.3) * np.sin(0.7)], atol=1e-5)
