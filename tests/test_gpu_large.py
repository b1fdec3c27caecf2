"""Registers of 2^31 amplitudes and more: one-qubit kernels at their largest
single launch (31 qubits), and beyond it launches stay below 2^32 threads
and the kernels loop over their virtual blocks (32 qubits, 68.7 GB
complex128).  The checks use closed forms instead of the oracle."""
import numpy as np
import pytest

from doki_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_32_qubits_gates_in_place(capi):
    n = 32
    s = capi.new(n)
    # H on the top qubit (pair kernel, highest stride), X on qubit 0
    # controlled by it (lane kernel), a dense 4-qubit unitary on 5..8
    capi.apply(s, capi.Gate(W.gate_h()), [n - 1], in_place=True)
    capi.apply(s, capi.Gate(W.gate_x()), [0], controls=[n - 1], in_place=True)
    u = W.haar_unitary(W.rng_for(32), 4)
    capi.apply(s, capi.Gate(u), [5, 6, 7, 8], in_place=True)
    # RY on qubit 12: a mid target, looping pair kernel at this size
    capi.apply(s, capi.Gate(W.gate_ry(0.7)), [12], in_place=True)
    p = capi.prob_all(s)
    want = np.zeros(n)
    want[0] = want[n - 1] = 0.5
    want[12] = np.sin(0.35) ** 2
    col = np.abs(u[:, 0]) ** 2
    for b in range(4):
        want[5 + b] = sum(col[j] for j in range(16) if (j >> b) & 1)
    assert np.max(np.abs(p - want)) <= 1e-12
    raw, nrm = s.raw(0, 1)
    assert abs(raw[0] / nrm - u[0, 0] * np.cos(0.35) / np.sqrt(2)) <= 1e-12
    hi = (1 << (n - 1)) | 1 | (3 << 5)
    raw, nrm = s.raw(hi, 1)
    assert abs(raw[0] / nrm - u[3, 0] * np.cos(0.35) / np.sqrt(2)) <= 1e-12


def test_31_qubits_largest_single_launch(capi):
    """2^31 amplitudes (34.4 GB): the one-qubit kernels at their largest
    single launch (the half-wave form for targets 6..17 uses exactly 2^31
    threads), against closed forms."""
    n = 31
    s = capi.new(n)
    angles = {3: 0.3, 9: 0.9, 16: 1.3, 18: 0.5, 30: 2.1}   # lane, half-wave, pair
    for q, a in angles.items():
        capi.apply(s, capi.Gate(W.gate_ry(a)), [q], in_place=True)
    # X on qubit 7 controlled by qubit 9 (half-wave kernel, control outside
    # the pair)
    capi.apply(s, capi.Gate(W.gate_x()), [7], controls=[9], in_place=True)
    p = capi.prob_all(s)
    want = np.zeros(n)
    for q, a in angles.items():
        want[q] = np.sin(a / 2) ** 2
    want[7] = want[9]
    assert np.max(np.abs(p - want)) <= 1e-12
    assert abs(s.norm - 1.0) <= 1e-12
    # one amplitude: qubits 3, 9, 7, 16, 18, 30 all set
    idx = (1 << 3) | (1 << 9) | (1 << 7) | (1 << 16) | (1 << 18) | (1 << 30)
    raw, nrm = s.raw(idx, 1)
    want_amp = np.prod([np.sin(a / 2) for a in angles.values()])
    assert abs(raw[0] / nrm - want_amp) <= 1e-12


def test_join_to_32_qubits(capi):
    a = W.random_normalised_state(1, 16)
    b = W.random_normalised_state(2, 16)
    j = capi.join(capi.from_host(a), capi.from_host(b))
    rng = W.rng_for(3)
    for _ in range(8):
        i, k = int(rng.integers(1 << 16)), int(rng.integers(1 << 16))
        raw, nrm = j.raw((i << 16) | k, 1)
        assert abs(raw[0] / nrm - a[i] * b[k]) <= 1e-12
    pa = [float(np.sum(np.abs(a[(np.arange(1 << 16) >> q) & 1 == 1]) ** 2)) for q in range(16)]
    pb = [float(np.sum(np.abs(b[(np.arange(1 << 16) >> q) & 1 == 1]) ** 2)) for q in range(16)]
    assert np.max(np.abs(capi.prob_all(j) - np.array(pb + pa))) <= 1e-12
