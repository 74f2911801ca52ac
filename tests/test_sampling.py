"""Measurement oracle pinned by the reference's KATs and by brute force, and
the host-side measurement-record conversions (quantum.py:2048-2450) on CPU
tensors.  No GPU."""
import numpy as np
import pytest
import torch

from oracle import sampling as osm
from oracle import statevector as sv
from tensorcircuit_amd import quantum as qu


def test_correlation_from_counts_kat():
    # quantum.py:2432-2436
    p = np.array([0.6, 0.4, 0, 0])
    assert abs(osm.correlation_from_counts([0, 1], p) - 0.2) < 1e-12
    assert abs(osm.correlation_from_counts([1], p) - 0.2) < 1e-12
    assert abs(float(qu.correlation_from_counts([0, 1], torch.tensor(p))) - 0.2) < 1e-12
    assert abs(float(qu.correlation_from_counts([1], torch.tensor(p))) - 0.2) < 1e-12


def test_searchsorted_kat_and_probability_sample_brute_force():
    # tests/test_backends.py:278-281
    edges = np.array([-1, 3.3, 9.1, 10.0])
    vals = np.array([0.0, 4.1, 12.0])
    np.testing.assert_array_equal(osm.searchsorted_left(edges, vals), [1, 2, 4])
    np.testing.assert_array_equal(np.searchsorted(edges, vals, side="left"), [1, 2, 4])
    rng = np.random.default_rng(3)
    p = rng.random(16) * (rng.random(16) > 0.3)
    u = rng.random(200)
    got = osm.probability_sample(p, u)
    q = p / p.sum()
    cum = [sum(q[: i + 1]) for i in range(16)]
    r = cum[-1] * (1 - u)
    want = osm.searchsorted_left(np.array(cum), r)
    m = osm.boundary_margin(p, u) > 1e-12
    np.testing.assert_array_equal(got[m], want[m])
    # zero-probability entries are never drawn away from boundaries
    assert np.all(p[got[m]] > 0)


def test_measure_jit_oracle_closed_forms():
    rng = np.random.default_rng(5)
    n = 5
    psi = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    psi /= np.linalg.norm(psi)
    for _ in range(20):
        st = rng.random(n)
        bits, p = osm.measure_jit(psi, list(range(n)), st)
        idx = int("".join(str(int(b)) for b in bits), 2)
        assert abs(p - abs(psi[idx]) ** 2) < 1e-12
    # a basis state measures to itself with probability 1
    for idx in (0, 7, 2**n - 1):
        e = np.zeros(2**n, complex)
        e[idx] = 1
        bits, p = osm.measure_jit(e, [3, 0, 4, 1, 2], rng.random(n))
        want = [(idx >> (n - 1 - q)) & 1 for q in [3, 0, 4, 1, 2]]
        np.testing.assert_array_equal(bits, want)
        assert abs(p - 1.0) < 1e-12


def test_sample_frequencies_match_p():
    rng = np.random.default_rng(8)
    p = np.array([0.05] * 8 + [0.2, 0.4] + [0.0] * 6)
    s = osm.probability_sample(p, rng.random(20000))
    counts = np.bincount(s, minlength=16)
    # the reference's statistical check (tests/test_backends.py:287-291)
    np.testing.assert_allclose(counts - p * 20000, 0, atol=400)
    assert counts[10:].sum() == 0


def test_sexpps_closed_form_oracle():
    # tests/test_circuit.py:1384-1388: (1, i)/sqrt2 has <Y> = 1
    psi = np.array([1.0, 1.0j]) / np.sqrt(2)
    rx = sv.gate_matrix("rx", theta=np.pi / 2)
    p = np.abs(rx @ psi) ** 2
    assert abs(osm.correlation_from_counts([0], p) - 1.0) < 1e-12


def test_readout_oracle_matches_kron():
    rng = np.random.default_rng(2)
    p = rng.random(8)
    ro = [[0.9, 0.75], [0.4, 0.7], [0.8, 0.95]]
    m = np.array([[1.0]])
    for e0, e1 in ro:
        m = np.kron(m, np.array([[e0, 1 - e1], [1 - e0, e1]]))
    np.testing.assert_allclose(osm.readout(p, ro), m @ p, atol=1e-14)


def test_record_formats_roundtrip():
    n = 3
    s = torch.tensor([5, 1, 5, 7, 0, 5])
    b = qu.sample_int2bin(s, n)
    assert b.tolist()[0] == [1, 0, 1]
    assert qu.sample_bin2int(b, n).tolist() == s.tolist()
    vals, cnts = qu.sample2count(s, n, jittable=False)
    assert dict(zip(vals.tolist(), cnts.tolist())) == {0: 1, 1: 1, 5: 3, 7: 1}
    cv = qu.sample2all(s, n, format="count_vector")
    assert cv.tolist() == [1, 1, 0, 0, 0, 3, 0, 1]
    assert qu.sample2all(b, n, format="count_dict_bin") == {"000": 1, "001": 1, "101": 3, "111": 1}
    assert qu.sample2all(s, n, format="count_dict_int") == {0: 1, 1: 1, 5: 3, 7: 1}
    vt, ct = qu.sample2all(s, n, format="count_tuple", jittable=True)
    assert vt.shape[0] == 8 and int(ct.sum()) == 6
    idx, val = qu.count_d2s(cv)
    assert idx.tolist() == [0, 1, 5, 7] and val.tolist() == [1, 1, 3, 1]
    assert qu.correlation_from_samples([0], s, n) == pytest.approx(np.mean([1 - 2 * ((x >> 2) & 1) for x in s.tolist()]))
    assert abs(float(qu.correlation_from_samples([0, 2], s, n))
               - osm.correlation_from_samples([0, 2], s.numpy(), n)) < 1e-6
