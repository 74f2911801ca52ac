"""Compiled pass kernels (``Program(..., jit=1)``): every tile pass runs as
straight-line generated code.  Checked against the complex128 oracle and
against the generic fused kernels on the same inputs.

Tolerances: complex64 states 4e-5 absolute (norm-1 states); energies and
gradients 1e-5 absolute at complex64 (the target's max|dgrad| <= 1e-5) and
1e-9 at complex128 (sums over up to a few hundred gates).  The 24-qubit
workload the benchmark times (20 layers, checkpoints every 5, compiled
passes, 5 register slots, searched schedule) is compared with its own
complex128 run: max|dgrad| <= 1e-5 and |dE| <= 1e-5 |E|.
"""
import numpy as np
import pytest
import torch

from oracle import statevector as sv

pytestmark = pytest.mark.gpu

ONE_Q = ["h", "x", "y", "s", "t", "rx", "ry", "rz", "phase", "exp1_h"]
TWO_Q = ["cnot", "cz", "swap", "rzz", "rxx", "ryy", "exp1_zz", "exp1_xy", "unitary2", "iswap"]
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
_XY = np.kron(np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]))
_ZZ = np.diag([1.0, -1.0, -1.0, 1.0]).astype(np.complex128)


def np_of(t):
    return t.detach().cpu().numpy()


def _random_ops(rng, n, ngates):
    ops = []
    for _ in range(ngates):
        if rng.random() < 0.5:
            ops.append((str(rng.choice(ONE_Q)), (int(rng.integers(n)),), float(rng.uniform(0, 2 * np.pi))))
        else:
            qs = tuple(int(q) for q in rng.permutation(n)[:2])
            name = str(rng.choice(TWO_Q))
            th = float(rng.uniform(0, 2 * np.pi))
            if name == "unitary2":
                th = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
            ops.append((name, qs, th))
    return ops


def _circuit(tc, n, ops):
    """The circuit and the oracle's op list (trainable gates take parameter
    slots in program order, as Circuit lowering assigns them)."""
    c = tc.Circuit(n)
    o = []
    pidx = 0
    for name, qs, th in ops:
        if name in ("h", "x", "y", "s", "t", "cnot", "cz", "swap"):
            getattr(c, name)(*qs)
            o.append(sv.Op(name, qs))
        elif name == "iswap":
            c.iswap(*qs, theta=th)
            o.append(sv.Op("iswap", qs, kw={"theta": th}))
        elif name == "unitary2":
            c.unitary(*qs, unitary=th)
            o.append(sv.Op("unitary", qs, kw={"unitary": th}))
        elif name.startswith("exp1_"):
            g = {"exp1_h": _H, "exp1_zz": _ZZ, "exp1_xy": _XY}[name]
            c.exp1(*qs, unitary=g, theta=th)
            o.append(sv.Op("exp1", qs, pidx=pidx, kw={"unitary": g}))
            pidx += 1
        else:
            getattr(c, name)(*qs, theta=th)
            o.append(sv.Op(name, qs, pidx=pidx))
            pidx += 1
    return c, o


def _terms(rng, n, nterms):
    out = []
    for _ in range(nterms):
        codes = [0] * n
        for q in rng.permutation(n)[: int(rng.integers(1, 4))]:
            codes[int(q)] = int(rng.integers(1, 4))
        out.append((float(rng.uniform(-1, 1)), codes))
    return out


@pytest.mark.parametrize("n", [10, 13])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("R", [5, 4, 3])
def test_jit_energy_grad_random_vs_oracle(gpu, n, dtype, R):
    """All gate kinds and generator kinds of the fused passes, compiled."""
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(2000 + 10 * n + R)
    ops = _random_ops(rng, n, 8 * n)
    c, oops = _circuit(tc, n, ops)
    specs, thetas = c._lower()
    params = np.array([float(np.real(t)) for t in thetas])
    terms = _terms(rng, n, 6)
    ham = engine.PauliSum(n, np.array([t[1] for t in terms], dtype=np.int8), np.array([t[0] for t in terms]))
    e_ref, g_ref = sv.adjoint_value_and_grad(n, oops, params, terms)
    prog = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=0, jit=1, reg_slots=R)
    rdt = torch.float32 if dtype == "complex64" else torch.float64
    p = torch.as_tensor(params, dtype=rdt).reshape(1, -1).to(gpu)
    e, g = prog.energy_grad(p, ham, 1, dtype)
    tol = 1e-5 if dtype == "complex64" else 1e-9
    de = abs(float(np_of(e)[0]) - e_ref)
    dg = float(np.max(np.abs(np_of(g).reshape(-1) - g_ref)))
    print("n=%d %s R=%d |dE|=%.2e max|dgrad|=%.2e" % (n, dtype, R, de, dg))
    assert de < tol
    assert dg < tol
    st, _ = prog.forward(p, 1, dtype)
    ref = sv.run_ops(n, oops, params)
    np.testing.assert_allclose(np_of(st).reshape(-1), ref, atol=4e-5 if dtype == "complex64" else 1e-11)


@pytest.mark.parametrize("R", [5, 4, 3])
def test_jit_matches_generic_hea_batch(gpu, R):
    """20-qubit HEA, batch 4, checkpoints every 2 layers: compiled passes
    against the generic fused kernels and the oracle."""
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L, B = 20, 6, 4
    rng = np.random.default_rng(77)
    ps = rng.uniform(0, 2 * np.pi, size=(B, 2 * L, n)).astype(np.float32)
    c = Circuit(n)
    hea_layers(c, ps[0], L)
    specs, thetas = c._lower()
    terms = sv.tfim_terms(n)
    ham = engine.PauliSum(n, np.array([t[1] for t in terms], dtype=np.int8), np.array([t[0] for t in terms]))
    pt = torch.as_tensor(ps.reshape(B, -1)).to(gpu)
    gen = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=2)
    jit = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=2, jit=1, reg_slots=R)
    e0, g0 = gen.energy_grad(pt, ham, B)
    e1, g1 = jit.energy_grad(pt, ham, B)
    np.testing.assert_allclose(np_of(e1), np_of(e0), atol=2e-5)
    np.testing.assert_allclose(np_of(g1), np_of(g0), atol=2e-5)
    s0, _ = gen.forward(pt, B, "complex64")
    s1, _ = jit.forward(pt, B, "complex64")
    np.testing.assert_allclose(np_of(s1), np_of(s0), atol=2e-5)
    e_ref, g_ref = sv.adjoint_value_and_grad(n, sv.hea_ops(n, L), ps[2].astype(np.float64).reshape(-1), terms)
    de = abs(float(np_of(e1)[2]) - e_ref)
    dg = float(np.max(np.abs(np_of(g1)[2] - g_ref)))
    print("20q HEA R=%d |dE|=%.2e max|dgrad|=%.2e" % (R, de, dg))
    assert de < 1e-5
    assert dg < 1e-5


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_jit_first_pass_generates_zero_state(gpu, dtype):
    """A compiled program run from |0..0> (its first pass generates the
    state instead of reading it) equals, bit for bit, the same program fed
    |0..0> as an input state (every pass reads); for forward states,
    checkpoints, energies and gradients, on a 16-qubit HEA batch of 3."""
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L, B = 16, 4, 3
    rng = np.random.default_rng(16)
    ps = rng.uniform(0, 2 * np.pi, size=(B, 2 * L, n)).astype(np.float32)
    c = Circuit(n)
    hea_layers(c, ps[0], L)
    specs, thetas = c._lower()
    terms = sv.tfim_terms(n)
    ham = engine.PauliSum(n, np.array([t[1] for t in terms], dtype=np.int8), np.array([t[0] for t in terms]))
    prog = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=2, jit=1)
    pt = torch.as_tensor(ps.reshape(B, -1)).to(gpu)
    zero = torch.zeros(B, 1 << n, dtype=getattr(torch, dtype), device=gpu)
    zero[:, 0] = 1
    s0, c0 = prog.forward(pt, B, dtype, want_ckpts=True)
    s1, c1 = prog.forward(pt, B, dtype, inputs=zero, want_ckpts=True)
    assert torch.equal(s0, s1) and torch.equal(c0, c1)
    e0, g0 = prog.energy_grad(pt, ham, B, dtype)
    e1, g1 = prog.energy_grad(pt, ham, B, dtype, inputs=zero)
    assert torch.equal(e0, e1) and torch.equal(g0, g1)
    ref = sv.run_ops(n, sv.hea_ops(n, L), ps[1].astype(np.float64).reshape(-1))
    np.testing.assert_allclose(np_of(s0)[1], ref, atol=1e-5 if dtype == "complex64" else 1e-12)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_jit_ry_shear_angle_edges(gpu, dtype):
    """ry runs in the compiled passes as three shears of the angle reduced
    into [-pi/2, pi/2], with the dropped sign restored at the end of each
    pass.  Angles at and around the reduction's edges (ry angle +-pi and
    +-2 pi, i.e. a half-angle of +-pi/2 and +-pi), negative and large ones:
    states, energies and gradients of a 12-qubit HEA against the oracle."""
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L, B = 12, 4, 3
    pi = np.pi
    edges = [0.0, pi, -pi, 2 * pi, -2 * pi, 3 * pi, -3 * pi, pi + 1e-6, pi - 1e-6, 2 * pi - 1e-6, 4 * pi + 0.1,
             -7.5, 20.0, -20.0, 1e-7, -1e-7, 0.5 * pi, 1.5 * pi, 2.5 * pi, -0.5 * pi]
    rng = np.random.default_rng(31)
    ps = np.empty((B, 2 * L, n))
    ps[0] = np.resize(np.array(edges), (2 * L, n))
    ps[1] = rng.uniform(-20, 20, size=(2 * L, n))
    ps[2] = rng.choice(np.array(edges), size=(2 * L, n)) + rng.uniform(-1e-3, 1e-3, size=(2 * L, n))
    rdt = np.float32 if dtype == "complex64" else np.float64
    ps = ps.astype(rdt)
    c = Circuit(n)
    hea_layers(c, np.zeros((2 * L, n)), L)
    specs, thetas = c._lower()
    terms = sv.tfim_terms(n)
    ham = engine.PauliSum(n, np.array([t[1] for t in terms], dtype=np.int8), np.array([t[0] for t in terms]))
    prog = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=2, jit=1, reg_slots=5)
    pt = torch.as_tensor(ps.reshape(B, -1)).to(gpu)
    e, g = prog.energy_grad(pt, ham, B, dtype)
    st, _ = prog.forward(pt, B, dtype)
    tol = 1e-5 if dtype == "complex64" else 1e-9
    for b in range(B):
        p64 = ps[b].astype(np.float64).reshape(-1)
        e_ref, g_ref = sv.adjoint_value_and_grad(n, sv.hea_ops(n, L), p64, terms)
        de = abs(float(np_of(e)[b]) - e_ref)
        dg = float(np.max(np.abs(np_of(g)[b] - g_ref)))
        print("set %d %s |dE|=%.2e max|dgrad|=%.2e" % (b, dtype, de, dg))
        assert de < tol and dg < tol
        ref = sv.run_ops(n, sv.hea_ops(n, L), p64)
        np.testing.assert_allclose(np_of(st)[b], ref, atol=4e-6 if dtype == "complex64" else 1e-12)


def test_jit_vjp_matches_generic(gpu):
    """Vector-Jacobian product of the forward map (the autograd backward)."""
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L, B = 14, 4, 3
    rng = np.random.default_rng(5)
    ps = rng.uniform(0, 2 * np.pi, size=(B, 2 * L, n)).astype(np.float32)
    c = Circuit(n)
    hea_layers(c, ps[0], L)
    specs, thetas = c._lower()
    pt = torch.as_tensor(ps.reshape(B, -1)).to(gpu)
    gen = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=1)
    jit = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=1, jit=1)
    gs = torch.as_tensor(rng.standard_normal((B, 2**n)) + 1j * rng.standard_normal((B, 2**n)),
                         dtype=torch.complex64).to(gpu)
    out = []
    for prog in (gen, jit):
        st, ck = prog.forward(pt, B, "complex64", want_ckpts=True)
        gp, gi = prog.vjp(pt, ck, st, gs, B, "complex64", True)
        out.append((np_of(gp), np_of(gi)))
    np.testing.assert_allclose(out[1][0], out[0][0], atol=1e-3, rtol=1e-5)
    np.testing.assert_allclose(out[1][1], out[0][1], atol=1e-5)


def test_jit_compile_reports_and_caches(gpu, tmp_path, monkeypatch):
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    monkeypatch.setenv("TCM_CACHE_DIR", str(tmp_path))
    n, L = 12, 2
    c = Circuit(n)
    hea_layers(c, np.zeros((2 * L, n), np.float32), L)
    specs, thetas = c._lower()
    a = engine.Program(n, specs, len(thetas), fuse=1, jit=1)
    t1 = a.compile("complex64", 0)
    b = engine.Program(n, specs, len(thetas), fuse=1, jit=1)
    t2 = b.compile("complex64", 0)
    assert t1 > 0.0 and t2 == 0.0  # second build reads the cached code objects
    assert list((tmp_path / "jit").glob("*.co"))


def test_c4_workload_vs_complex128(gpu):
    """The benchmark's workload exactly (24 qubits, 20 HEA layers, TFIM
    h = 1, checkpoints every 5 layers, compiled passes, default register
    slots and searched schedule, parameter sets from default_rng(1004)):
    the complex64 energies and gradients of two parameter sets against the
    complex128 run of the same program (itself oracle-pinned at n <= 16 and
    by the closed forms at 24 qubits)."""
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L, B, P = 24, 20, 32, 960
    c = Circuit(n)
    hea_layers(c, np.zeros((2 * L, n)), L)
    specs, thetas = c._lower()
    assert len(thetas) == P
    terms = sv.tfim_terms(n)
    ham = engine.PauliSum(n, np.array([t[1] for t in terms], dtype=np.int8), np.array([t[0] for t in terms]))
    rng = np.random.default_rng(1004)
    ps = rng.uniform(0, 2 * np.pi, size=(B, P)).astype(np.float32)[:2]
    p64 = torch.as_tensor(ps).to(gpu)
    prog = engine.Program(n, specs, P, fuse=1, ckpt_every=5, jit=1)
    e, g = prog.energy_grad(p64, ham, 2, "complex64")
    prog2 = engine.Program(n, specs, P, fuse=1, ckpt_every=5, jit=1)
    e2, g2 = prog2.energy_grad(p64.double(), ham, 2, "complex128")
    de = np.abs(np_of(e) - np_of(e2)) / np.abs(np_of(e2))
    dg = np.max(np.abs(np_of(g) - np_of(g2)))
    print("C4 24q L=20: rel|dE| %s max|dgrad| %.2e (|E| %s)" % (de, dg, np.abs(np_of(e2))))
    assert np.all(de <= 1e-5)
    assert dg <= 1e-5
