"""Batched uneven circuits (the reference's vmap over circuit structures,
backends/pytorch_backend.py:605-617): compiled multi-program launches
(engine.MultiPlan, tcm_forward_multi_plan) against single-program runs and
the complex128 oracle at n <= 16; the program-level autotuner never picks a
configuration slower than the default by its own measurement; the
engine.UnevenBatch step (autotuned programs, step-level choice among the
launch strategies, one HIP graph) against single-program runs and the
oracle."""
import numpy as np
import pytest
import torch

from oracle import statevector as sv

pytestmark = pytest.mark.gpu


def np_of(t):
    return t.detach().cpu().numpy()


def _hea_program(n, L, **opts):
    from tensorcircuit_amd import Circuit, engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    c = Circuit(n)
    hea_layers(c, np.zeros((2 * L, n)), L, barrier=False)
    specs, thetas = c._lower()
    return engine.Program(n, specs, len(thetas), fuse=1, **opts), specs, len(thetas)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_multi_plan_matches_single_programs_and_oracle(gpu, dtype):
    """Programs of 6..16 qubits (batches 1..5, register slots 3/4/5, one
    batch empty, one 6-qubit program small enough to run plain passes) in
    one plan: every state equals its single-program compiled run and the
    oracle (4e-6 complex64, 1e-12 complex128)."""
    from tensorcircuit_amd import engine, workloads

    cdt = torch.complex64 if dtype == "complex64" else torch.complex128
    rdt = torch.float32 if dtype == "complex64" else torch.float64
    items, progs, refs = [], [], []
    for i, n in enumerate([6, 10, 11, 12, 13, 14, 16]):
        L = 2 + i % 3
        R = (3, 4, 5)[i % 3] if n >= 10 else 0
        prog, specs, P = _hea_program(n, L, jit=1, reg_slots=R)
        b = 0 if n == 13 else 1 + i % 5
        p = torch.as_tensor(workloads.hea_params(3000 + n, L, n, batch=max(b, 1)).reshape(max(b, 1), -1))
        p = p[:b].to(rdt).to(gpu)
        st = torch.full((b, 1 << n), 7.0, dtype=cdt, device=gpu)
        items.append((prog, p, st))
        progs.append(prog)
        refs.append((n, L, p))
    plan = engine.MultiPlan(progs, dtype)
    plan.forward(items)
    torch.cuda.synchronize()
    tol = 4e-6 if dtype == "complex64" else 1e-12
    for (prog, p, st), (n, L, pp) in zip(items, refs):
        if st.shape[0] == 0:
            continue
        s1, _ = prog.forward(p, p.shape[0], dtype)
        np.testing.assert_allclose(np_of(st), np_of(s1), atol=tol, err_msg="n=%d" % n)
        for b in range(p.shape[0]):
            ref = sv.run_ops(n, sv.hea_ops(n, L), pp[b].double().cpu().numpy().reshape(-1))
            np.testing.assert_allclose(np_of(st)[b], ref, atol=tol, err_msg="n=%d b=%d" % (n, b))
    # a second forward with new parameters (plan reuse, init_zero each time)
    for _, p, _ in items:
        p.add_(0.37)
    plan.forward(items)
    prog, p, st = items[-1]
    s1, _ = prog.forward(p, p.shape[0], dtype)
    np.testing.assert_allclose(np_of(st), np_of(s1), atol=tol)


def test_multi_plan_refuses_foreign_descriptors(gpu):
    from tensorcircuit_amd import engine

    a, _, _ = _hea_program(10, 2, jit=1)
    b, _, _ = _hea_program(11, 2, jit=1)
    plan = engine.MultiPlan([a, b])
    st = torch.zeros(1, 1 << 10, dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError):
        plan.forward([(a, torch.zeros(1, a.nparams, device=gpu), st)])  # one descriptor for two programs
    with pytest.raises(ValueError):
        plan.forward([(b, torch.zeros(1, b.nparams, device=gpu), st), (a, torch.zeros(1, a.nparams, device=gpu), st)])


def test_autotune_program_never_slower_than_default(gpu, tmp_path, monkeypatch):
    """engine.autotune_program times the real program on the real batch for
    (tile bits, register slots) candidates and keeps the default unless a
    candidate is faster by the margin; the chosen program computes the same
    states."""
    from tensorcircuit_amd import engine, workloads

    monkeypatch.setenv("TCM_CACHE_DIR", str(tmp_path))
    n, L, B = 14, 4, 64
    prog, specs, P = _hea_program(n, L)
    opts, times = engine.autotune_program(n, specs, P, B, "complex64", jit=1)
    assert times["default"] > 0
    assert times[engine.tune_key(opts)] <= times["default"] * (1 + 1e-9)
    assert set(opts) <= {"tile_bits", "reg_slots"}
    tuned = engine.Program(n, specs, P, fuse=1, jit=1, **opts)
    p = torch.as_tensor(workloads.hea_params(77, L, n, batch=B).reshape(B, -1)).cuda()
    a, _ = tuned.forward(p, B, "complex64")
    b, _ = engine.Program(n, specs, P, fuse=1, jit=1).forward(p, B, "complex64")
    np.testing.assert_allclose(np_of(a), np_of(b), atol=4e-6)
    # cached: the second call returns the same choice without timing
    opts2, times2 = engine.autotune_program(n, specs, P, B, "complex64", jit=1)
    assert opts2 == opts and times2 == {}


def test_circuit_api_routes_constant_circuits_to_matrix_cores(gpu):
    """A constant 20-qubit brickwork built through the Circuit API (Haar
    SU(2) layers and CZ pairs, as the 30-qubit matrix-core workload) runs as
    fused dense 5-qubit blocks on the FP32 matrix cores; its state matches
    the complex128 oracle within 1e-5 per amplitude.  The same circuit with
    a tensor angle keeps the trainable VALU path."""
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import cons

    n, depth = 20, 6
    ops = sv.brickwork_ops(n, depth, 1005)
    c = tc.Circuit(n)
    for o in ops:
        if o.name == "cz":
            c.cz(*o.qubits)
        else:
            c.unitary(*o.qubits, unitary=o.matrix())
    assert cons.mfma_for(n)
    prog = c._mfma_program()
    assert prog is not None
    d = prog.describe("complex64", 0)
    assert d["sched"]["passes"] and all(p["mfma"] for p in d["sched"]["passes"])
    st = np_of(c.wavefunction())
    ref = sv.run_ops(n, ops)
    err = np.max(np.abs(st - ref))
    fid = abs(np.vdot(ref, st.astype(np.complex128))) ** 2
    print("20q brickwork via Circuit: max|dpsi| %.2e fidelity %.9f" % (err, fid))
    assert err < 1e-5 and abs(fid - 1) < 1e-5
    c2 = tc.Circuit(n)
    c2.rx(0, theta=torch.tensor(0.3))
    for o in ops[:10]:
        c2.unitary(*o.qubits, unitary=o.matrix()) if o.name != "cz" else c2.cz(*o.qubits)
    assert c2._mfma_program() is None


def test_circuit_api_matrix_cores_random_long_range(gpu):
    """A constant 20-qubit circuit of random two-qubit unitaries on random
    (mostly distant) qubit pairs and random single-qubit unitaries, built
    through the Circuit API: the fused blocks (qubit sets chosen by the gates
    they absorb, not by locality) on the FP32 matrix cores match the
    complex128 oracle of the unfused gates within 1e-5 per amplitude."""
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import cons

    n = 20
    rng = np.random.default_rng(2024)

    def ru(d):
        q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
        return q * (np.diag(r) / np.abs(np.diag(r)))

    c = tc.Circuit(n)
    ref = sv.zero_state(n)
    for _ in range(240):
        if rng.random() < 0.6:
            qs = [int(q) for q in rng.choice(n, 2, replace=False)]
            u = ru(4)
        else:
            qs = [int(rng.integers(n))]
            u = ru(2)
        c.unitary(*qs, unitary=u)
        ref = sv.apply_gate(ref, u, qs, n)
    assert cons.mfma_for(n)
    prog = c._mfma_program()
    assert prog is not None
    st = np_of(c.wavefunction())
    err = np.max(np.abs(st - ref))
    print("20q random long-range via Circuit: %d passes, max|dpsi| %.2e" % (prog.npasses, err))
    assert err < 1e-5


def test_uneven_batch_step(gpu, tmp_path, monkeypatch):
    """engine.UnevenBatch on four HEA structures (12 qubits x 96 parameter
    sets holding most amplitudes, and 8, 9, 10 qubits x 2..5) with TFIM
    energies: the step chosen among sequence / multi-program / hybrid /
    two-stream hybrid candidates is not slower than the default by its own
    measurement; replays of its HIP graph give every state of its program's
    single run (4e-6) and of the oracle (4e-6 at complex64), the energies of
    PauliSum.expect on those states (1e-9); parameters updated in place show
    in the next replay; the eager step gives the same states."""
    from tensorcircuit_amd import Circuit, engine, workloads
    from tensorcircuit_amd.templates.blocks import hea_layers

    monkeypatch.setenv("TCM_CACHE_DIR", str(tmp_path))
    rng = np.random.default_rng(41)
    sizes = [(12, 3, 96), (8, 2, 5), (9, 2, 3), (10, 3, 2)]
    structs, hams = [], []
    for n, L, B in sizes:
        c = Circuit(n)
        hea_layers(c, np.zeros((2 * L, n)), L, barrier=False)
        specs, thetas = c._lower()
        structs.append((n, specs, len(thetas)))
        hams.append(engine.PauliSum(n, *workloads.tfim_terms(n)))
    ub = engine.UnevenBatch(structs, [B for _, _, B in sizes], hams=hams)
    assert ub.large == [True, False, False, False]
    assert {"sequence_default", "multi_default", "hybrid_default", "hybrid_streams_default"} <= set(ub.candidate_ms)
    assert ub.candidate_ms[ub.choice] <= ub.candidate_ms["sequence_default"] * (1 + 1e-9)

    def check(es):
        torch.cuda.synchronize()
        for (n, L, B), prog, pa, st, h, e in zip(sizes, ub.programs, ub.params, ub.states, hams, es):
            s1, _ = prog.forward(pa, B, "complex64")
            np.testing.assert_allclose(np_of(st), np_of(s1), atol=4e-6, err_msg="n=%d" % n)
            np.testing.assert_allclose(np_of(e), np_of(h.expect(s1)), atol=1e-9)
            ref = sv.run_ops(n, sv.hea_ops(n, L), pa[B - 1].double().cpu().numpy())
            np.testing.assert_allclose(np_of(st)[B - 1], ref, atol=4e-6, err_msg="oracle n=%d" % n)

    for pa in ub.params:
        pa.copy_(torch.as_tensor(rng.uniform(0, 2 * np.pi, size=tuple(pa.shape)).astype(np.float32)))
    check(ub.run())
    for pa in ub.params:
        pa.add_(0.37)
    check(ub.run())
    replayed = [np_of(st).copy() for st in ub.states]
    ub.run_eager()
    torch.cuda.synchronize()
    for st, r in zip(ub.states, replayed):
        np.testing.assert_allclose(np_of(st), r, atol=1e-7)
