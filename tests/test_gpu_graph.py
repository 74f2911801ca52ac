"""HIP graph capture of engine steps (engine.CapturedStep): every replay,
with eager work and synchronisation between replays and parameters changed
in place, equals the eager result."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _hea(tc, n, L):
    from tensorcircuit_amd.templates.blocks import hea_layers

    c = tc.Circuit(n)
    hea_layers(c, np.zeros((2 * L, n)), L, barrier=False)
    return c._lower()


def test_captured_multi_forward_and_energies(gpu):
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine, workloads

    rng = np.random.default_rng(3)
    items, hams = [], []
    for n, L, B in ((10, 3, 16), (13, 2, 4), (7, 4, 9)):
        specs, thetas = _hea(tc, n, L)
        prog = engine.Program(n, specs, len(thetas), fuse=1)
        p = torch.as_tensor(rng.uniform(0, 2 * np.pi, (B, len(thetas))).astype(np.float32)).to(gpu)
        items.append((prog, p, torch.empty(B, 1 << n, dtype=torch.complex64, device=gpu)))
        hams.append(engine.PauliSum(n, *workloads.tfim_terms(n)))

    def step():
        engine.forward_multi(items)
        return [h.expect(st) for (_, _, st), h in zip(items, hams)]

    g = engine.CapturedStep(step)
    for rep in range(4):
        for _, p, _ in items:
            p.copy_(torch.as_tensor(rng.uniform(0, 2 * np.pi, tuple(p.shape)).astype(np.float32)))
        out = [e.clone() for e in g.replay()]
        torch.cuda.synchronize()
        states = [st.clone() for _, _, st in items]
        # eager work between replays, then the eager reference
        ref = step()
        torch.cuda.synchronize()
        for a, b in zip(out, ref):
            np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=0, atol=1e-12)
        for s, (_, _, st) in zip(states, items):
            np.testing.assert_allclose(s.cpu().numpy(), st.cpu().numpy(), atol=0)


def test_captured_energy_grad_compiled(gpu):
    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine, workloads

    n, L, B = 14, 3, 4
    specs, thetas = _hea(tc, n, L)
    prog = engine.Program(n, specs, len(thetas), fuse=1, ckpt_every=1, jit=1)
    ham = engine.PauliSum(n, *workloads.tfim_terms(n), jit=1)
    rng = np.random.default_rng(9)
    p = torch.as_tensor(rng.uniform(0, 2 * np.pi, (B, len(thetas))).astype(np.float32)).to(gpu)
    g = engine.CapturedStep(lambda: prog.energy_grad(p, ham, B))
    for rep in range(3):
        p.copy_(torch.as_tensor(rng.uniform(0, 2 * np.pi, (B, len(thetas))).astype(np.float32)))
        e1, g1 = (t.clone() for t in g.replay())
        torch.cuda.synchronize()
        e0, g0 = prog.energy_grad(p, ham, B)
        torch.cuda.synchronize()
        np.testing.assert_allclose(e1.cpu().numpy(), e0.cpu().numpy(), atol=1e-12)
        np.testing.assert_allclose(g1.cpu().numpy(), g0.cpu().numpy(), atol=1e-12)


def test_captured_step_outlives_cache_eviction(gpu):
    """A step captured through the Circuit API (cached programs and Pauli
    sums) keeps replaying correctly after the engine's caches are emptied:
    the step holds every object whose device tables the graph reads."""
    import gc

    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine

    n, L = 9, 2
    p = torch.as_tensor(np.random.default_rng(4).uniform(0, 2 * np.pi, (2 * L, n)).astype(np.float32)).to(gpu)

    def step():
        c = tc.Circuit(n)
        tc.templates.blocks.hea_layers(c, p, L)
        return tc.backend.real(c.expectation_ps(z=[0, 1]) + c.expectation_ps(x=[3]))

    g = engine.CapturedStep(step)
    held = {type(o).__name__ for o in g.held}
    assert "Program" in held and "PauliSum" in held
    engine._PROG_CACHE.clear()
    engine._PAULI_CACHE.clear()
    gc.collect()
    for rep in range(3):
        p.add_(0.1)
        out = g.replay().clone()
        torch.cuda.synchronize()
        ref = step()
        torch.cuda.synchronize()
        np.testing.assert_allclose(out.cpu().numpy(), ref.cpu().numpy(), atol=1e-6)
