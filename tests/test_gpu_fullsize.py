"""The benchmark configurations at their full sizes on the GPU, checked by
what the full size allows: the complex64 result against the complex128 run
of the same program (which the oracle pins at n <= 16 and the closed forms
pin at 24 qubits), norms, fidelities, and the batched uneven step against
single-program runs and the oracle for its smallest circuits.

  26-qubit 10-layer HEA forward + TFIM energy (seed 1002): |dE| <= 1e-5 |E|,
      fidelity with complex128 within 1e-5 of 1;
  30-qubit depth-20 brickwork fused into 5-qubit blocks on the FP32 matrix
      cores (seed 1005): fidelity with the complex128 VALU run within 1e-5;
  1024 x 16-qubit 8-layer HEA + 256 circuits of 10..20 qubits (seed 1003)
      through a compiled multi-program plan: every state equals its program's
      single run (4e-6) and the 10-qubit states the oracle (4e-6)."""
import numpy as np
import pytest
import torch

from oracle import statevector as sv

pytestmark = pytest.mark.gpu


def _hea(n, L):
    from tensorcircuit_amd import Circuit
    from tensorcircuit_amd.templates.blocks import hea_layers

    c = Circuit(n)
    hea_layers(c, np.zeros((2 * L, n)), L)
    return c._lower()


def test_c2_26q_forward_energy_vs_complex128(gpu):
    from tensorcircuit_amd import engine, workloads

    n, L = 26, 10
    specs, thetas = _hea(n, L)
    codes, w = workloads.tfim_terms(n)
    ham = engine.PauliSum(n, codes, w)
    p = torch.as_tensor(workloads.hea_params(1002, L, n).reshape(1, -1)).cuda()
    prog = engine.Program(n, specs, len(thetas), fuse=1, jit=1)
    st, _ = prog.forward(p, 1, "complex64")
    st64, _ = prog.forward(p.double(), 1, "complex128")
    e32 = float(ham.expect(st)[0].real)
    e64 = float(ham.expect(st64)[0].real)
    fid = abs(complex(engine.inner(st64, st.to(torch.complex128), n)[0])) ** 2
    norm = float(engine.inner(st, st, n)[0].real)
    print("C2 26q: E %.9f (c128 %.9f) |dE| %.2e fidelity %.9f norm %.9f" % (e32, e64, abs(e32 - e64), fid, norm))
    assert abs(e32 - e64) <= 1e-5 * abs(e64)
    assert abs(fid - 1.0) < 1e-5 and abs(norm - 1.0) < 1e-5


def test_c5_30q_brickwork_mfma_vs_complex128(gpu):
    from tensorcircuit_amd import engine, fusion, workloads

    n, depth = 30, 20
    gates = workloads.brickwork(n, depth, 1005)
    blocks = fusion.fuse(gates, n)
    p32 = engine.Program(n, blocks, 0, fuse=1, mfma=2)
    st32, _ = p32.forward(None, 1, "complex64")
    p64 = engine.Program(n, blocks, 0, fuse=1, mfma=0)
    st64, _ = p64.forward(None, 1, "complex128")
    fid = abs(complex(engine.inner(st64, st32.to(torch.complex128), n)[0])) ** 2
    del st64
    norm = float(engine.inner(st32, st32, n)[0].real)
    print("C5 30q: %d blocks in %d passes, fidelity %.9f norm %.9f" % (len(blocks), p32.npasses, fid, norm))
    assert abs(fid - 1.0) < 1e-5 and abs(norm - 1.0) < 1e-5


def test_c3_batched_uneven_plan(gpu):
    from tensorcircuit_amd import engine

    rng = np.random.default_rng(1003)
    ns = rng.integers(10, 21, 256)
    groups = {}
    for i, n in enumerate(ns):
        groups.setdefault(int(n), []).append(i)
    sizes = [(16, 8, 1024)] + [(n, 6, len(groups[n])) for n in sorted(groups)]
    progs, items, ops = [], [], []
    for n, L, B in sizes:
        P = 2 * L * n
        pa = torch.as_tensor(rng.uniform(0, 2 * np.pi, size=(B, P)).astype(np.float32)).cuda()
        st = torch.empty(B, 1 << n, dtype=torch.complex64, device="cuda")
        specs, thetas = _hea(n, L)
        prog = engine.Program(n, specs, len(thetas), fuse=1, jit=1)
        progs.append(prog)
        items.append((prog, pa, st))
        ops.append(sv.hea_ops(n, L))
    plan = engine.MultiPlan(progs, "complex64", batches=[B for _, _, B in sizes])
    plan.forward(items)
    torch.cuda.synchronize()
    for (prog, pa, st), (n, L, B) in zip(items, sizes):
        s1, _ = prog.forward(pa[:3], 3, "complex64")
        np.testing.assert_allclose(st[:3].cpu().numpy(), s1.cpu().numpy(), atol=4e-6, err_msg="n=%d" % n)
        if n == 10:
            for b in range(B):
                ref = sv.run_ops(n, sv.hea_ops(n, L), pa[b].double().cpu().numpy())
                np.testing.assert_allclose(st[b].cpu().numpy(), ref, atol=4e-6)
