"""CPU oracle for the dense statevector path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The library in ``tensorcircuit_amd`` never imports it and has no
CPU fallback: its compute calls go to the HIP library or raise.

Contents
--------
``statevector``  complex128 numpy restatement of the reference's plain method:
                 gate matrices, gate-by-gate ``tensordot`` application,
                 ``expectation_ps`` / Pauli sums, the HEA/TFIM workloads,
                 an adjoint gradient, central finite differences and the
                 parameter-shift rule.
``sampling``     float64 restatements of the reference's measurement
                 algorithms: probability_sample, measure_jit, correlations,
                 readout-error matrices.
``plain_torch``  the reference's pytorch-backend plain method on the CPU
                 (gate-by-gate tensordot + torch autograd); used only as the
                 timed CPU baseline.

Pinning (what checks each function; details in ``tests/test_oracle.py``)
--------------------------------------------------------------------------
* gate matrices: the reference's gate KATs (tests/test_gates.py:18-156,
  tests/test_circuit.py:64-84) and unitarity / closed forms.
* ``apply_gate`` / ``run_ops``: brute-force dense kron construction at n <= 8
  (independent code path), index-order KATs (tests/test_circuit.py:21-43).
* ``expectation_ps`` / ``pauli_sum_*``: reference KATs
  (tests/test_circuit.py:239-245, 370-429, 1278-1281, 1373-1380), the dense
  Pauli-matrix pairs of tests/test_miscs.py:24-55, and dense ``<psi|H|psi>``.
* ``adjoint_value_and_grad``: central finite differences, the parameter-shift
  rule (exact for rotation gates) and the reference's gradient KATs
  (tests/test_interfaces.py:30-97).
* ``tfim`` energies: exact diagonalisation lower bound at n <= 10, all-zero
  angles give E = -(n-1) exactly.
* ``haar_su2`` / ``brickwork_ops``: unitarity and det = 1 of every draw.
* ``sampling``: the reference's KATs (quantum.py:2432-2436,
  tests/test_backends.py:278-281, tests/test_circuit.py:1384-1388),
  brute-force prefix sums, P(full record) = |psi_idx|^2, basis states,
  Kronecker readout matrices, draw frequencies (tests/test_sampling.py).
* ``plain_torch``: energy and autograd gradient equal ``statevector``'s
  adjoint at complex128 (5 qubits, 2 layers) to 1e-12.
"""
