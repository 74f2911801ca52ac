"""tcm_expect_pauli_sum_noalloc (partial sums in a static device array, no
allocation) against tcm_expect_pauli_sum on the same inputs: identical
kernels and reduction order, so the results are equal bit for bit; two calls
in a row, the second with a smaller batch; a batch beyond the static
scratch and an unprepared Pauli sum are refused with an error."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import statevector as sv

def _noalloc(ham, state, code):
    from tensorcircuit_amd import _lib, engine

    s = state.reshape(-1, 1 << ham.n).contiguous()
    out = torch.empty(s.shape[0], 2, dtype=torch.float64, device=s.device)
    _lib.check(_lib.lib.tcm_expect_pauli_sum_noalloc(ham.handle, code, s.data_ptr(), s.shape[0], out.data_ptr(),
                                                      engine._stream()), "tcm_expect_pauli_sum_noalloc")
    return torch.view_as_complex(out)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("n,jit", [(12, 1), (18, 0), (18, 1)])
def test_noalloc_matches_workspace_entry_point(gpu, dtype, n, jit):
    from tensorcircuit_amd import _lib, engine, workloads

    code = _lib.TCM_C64 if dtype == "complex64" else _lib.TCM_C128
    cdt = torch.complex64 if dtype == "complex64" else torch.complex128
    rng = np.random.default_rng(n + 100 * jit)
    codes = rng.integers(0, 4, size=(9, n))
    w = rng.standard_normal(9)
    ham = engine.PauliSum(n, codes, w, jit=jit)
    st = torch.as_tensor(rng.standard_normal((6, 1 << n)) + 1j * rng.standard_normal((6, 1 << n))).to(cdt).to(gpu)
    ref = ham.expect(st)  # prepares the Pauli sum on this device (tables, compiled passes)
    torch.cuda.synchronize()
    a = _noalloc(ham, st, code)
    torch.cuda.synchronize()
    assert torch.equal(a, ref)
    b = _noalloc(ham, st[:2], code)  # second call, smaller batch
    ref_b = ham.expect(st[:2])
    torch.cuda.synchronize()
    assert torch.equal(b, ref_b)
    # and against the oracle at the stated tolerance
    tol = 1e-5 if dtype == "complex64" else 1e-12
    terms = [(float(w[k]), list(codes[k])) for k in range(9)]
    s1 = st[1].cpu().numpy().astype(np.complex128)
    e_ref = sv.pauli_sum_expectation(s1, n, terms)
    assert abs(complex(b[1]) - e_ref) <= tol * max(1.0, abs(e_ref)) * np.abs(s1).dot(np.abs(s1))


@pytest.mark.gpu
def test_noalloc_refuses_oversize_and_unprepared(gpu):
    from tensorcircuit_amd import _lib, engine, workloads

    n = 20
    ham = engine.PauliSum(n, *workloads.tfim_terms(n), jit=0)
    st = torch.zeros(1, 1 << n, dtype=torch.complex64, device=gpu)
    st[0, 0] = 1
    with pytest.raises(ValueError, match="not prepared"):
        _noalloc(ham, st, _lib.TCM_C64)
    ham.expect(st)
    assert complex(_noalloc(ham, st, _lib.TCM_C64)[0]).real == pytest.approx(-(n - 1))
    nb = C.c_int64()
    big = 1
    while True:
        _lib.lib.tcm_pauli_workspace_bytes(ham.handle, _lib.TCM_C64, big, C.byref(nb))
        if nb.value > (8 << 20):
            break
        big *= 2
    # the size check comes before any launch: a zero-filled pointer range is never read
    sbig = torch.zeros(big, 1 << n, dtype=torch.complex64, device=gpu)
    with pytest.raises(RuntimeError, match="static scratch"):
        _noalloc(ham, sbig, _lib.TCM_C64)


def test_noalloc_size_check_is_host_side():
    """CPU: a batch whose partial sums exceed the static scratch is refused
    before the device is touched (no GPU needed)."""
    from tensorcircuit_amd import _lib, engine, workloads

    n = 20
    ham = engine.PauliSum(n, *workloads.tfim_terms(n), jit=0)
    out = (C.c_double * 2)()
    rc = _lib.lib.tcm_expect_pauli_sum_noalloc(ham.handle, _lib.TCM_C64, C.cast(out, C.c_void_p), 1 << 12,
                                               C.cast(out, C.c_void_p), None)
    assert rc == _lib.E_WORKSPACE and "static scratch" in _lib.last_error()
    assert _lib.lib.tcm_expect_pauli_sum_noalloc(None, _lib.TCM_C64, None, 1, None, None) == _lib.E_ARG
