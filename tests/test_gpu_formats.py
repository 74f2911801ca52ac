"""Circuits rebuilt from JSON and qsim text produce the same states
(tests/test_circuit.py:1443-1458 of the reference; qsim input format)."""
import numpy as np
import pytest
import torch

from oracle import statevector as sv

pytestmark = pytest.mark.gpu


def test_circuit_to_json_state(gpu):
    import tensorcircuit_amd as tc

    c = tc.Circuit(3)
    c.h(0)
    c.CNOT(1, 2)
    c.rxx(0, 2, theta=0.3)
    c.crx(0, 1, theta=-0.8)
    c.r(1, theta=1.0, alpha=0.2)
    c.toffoli(0, 2, 1)
    c.ccnot(0, 1, 2)
    for simplified in (False, True):
        c2 = tc.Circuit.from_json(c.to_json(simplified=simplified))
        np.testing.assert_allclose(c.state().cpu().numpy(), c2.state().cpu().numpy(), atol=1e-5)


def test_qsim_state_vs_oracle(gpu, tmp_path):
    import tensorcircuit_amd as tc

    q = tmp_path / "c.qsim"
    q.write_text("4\n0 h 0\n0 h 1\n0 x_1_2 2\n0 y_1_2 3\n1 cz 0 1\n1 fs 2 3 0.7 0.2\n"
                 "2 rz 0 0.3\n2 cnot 1 2\n3 hz_1_2 3\n3 t 0\n4 is 0 3\n")
    c = tc.Circuit.from_qsim_file(str(q))
    ops = [sv.Op("h", (0,)), sv.Op("h", (1,)), sv.Op("rx", (2,), kw={"theta": np.pi / 2}),
           sv.Op("ry", (3,), kw={"theta": np.pi / 2}), sv.Op("cz", (0, 1)),
           sv.Op("iswap", (2, 3), kw={"theta": -0.7}), sv.Op("cphase", (2, 3), kw={"theta": -0.2}),
           sv.Op("rz", (0,), kw={"theta": 0.3}), sv.Op("cnot", (1, 2)), sv.Op("wroot", (3,)),
           sv.Op("phase", (0,), kw={"theta": np.pi / 4}), sv.Op("iswap", (0, 3), kw={"theta": 1.0})]
    ref = sv.run_ops(4, ops, np.zeros(0))
    np.testing.assert_allclose(c.state().cpu().numpy(), ref, atol=1e-5)
