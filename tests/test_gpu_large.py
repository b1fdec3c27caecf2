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


def _brickwork_closed_form(steps, n):
    """Amplitudes of one brickwork layer applied to |0...0>: the U3 layer
    makes the product state prod_q U3_q[b_q, 0]; each CZ on (q, q+1) then
    flips the sign where both bits are set.  Returns (amp(idx), P(q = 1))."""
    col = np.zeros((n, 2), dtype=np.complex128)
    pairs = []
    for m, t, c, a in steps:
        if c:
            pairs.append((min(c), t[0]))
        else:
            col[t[0]] = m[:, 0]

    def amp(idx):
        idx = np.asarray(idx, dtype=np.uint64)
        v = np.ones(idx.shape, dtype=np.complex128)
        for q in range(n):
            v *= col[q][((idx >> np.uint64(q)) & np.uint64(1)).astype(np.int64)]
        for lo, hi in pairs:
            both = ((idx >> np.uint64(lo)) & (idx >> np.uint64(hi)) & np.uint64(1)) == 1
            v = np.where(both, -v, v)
        return v

    return amp, np.abs(col[:, 1]) ** 2


@pytest.mark.parametrize("mode", ["eager", "fuse"])
def test_33_qubit_brickwork_layer_closed_form(doki, capi, mode):
    """The 33-qubit register (137.4 GB complex128) of the brickwork workload:
    one layer from |0...0> through the drop-in module, gate by gate out of
    place ("eager": old and new state together, 274.9 GB) or as
    cache-blocked passes in place ("fuse"), then in place through the C ABI;
    sampled amplitudes and every marginal against the closed form."""
    n = 33
    steps = W.brickwork_layer(W.rng_for(3333), n, 0)
    assert len(steps) == 33 + 16
    amp, p1 = _brickwork_closed_form(steps, n)
    rng = W.rng_for(33)
    idx = np.concatenate([rng.integers(0, 1 << n, size=1500, dtype=np.uint64),
                          np.array([0, (1 << n) - 1, 1 << (n - 1), 0x155555555],
                                   dtype=np.uint64)])
    capi.check(capi.lib().dk_memory_trim())
    gates = [(doki.gate_new(1, m.tolist(), False), t, set(c)) for m, t, c, a in steps]
    prev = doki.execution_mode(mode, False)
    try:
        reg = doki.registry_new(n, False)
        for g, t, c in gates:
            reg = doki.registry_apply(reg, g, t, c, set(), -1, False)
        doki.registry_sync(reg, False)
    finally:
        doki.execution_mode(prev, False)
    got = np.array([doki.registry_get(reg, int(i), False, False) for i in idx])
    assert np.max(np.abs(got - amp(idx))) <= 1e-12
    probs = np.array([doki.registry_prob(reg, q, -1, False) for q in range(n)])
    assert np.max(np.abs(probs - p1)) <= 1e-12
    doki.registry_del(reg, False)
    del reg
    capi.synchronize()

    s = capi.new(n)
    for m, t, c, a in steps:
        capi.apply(s, capi.Gate(m), t, c, a, in_place=True)
    assert np.max(np.abs(capi.prob_all(s) - p1)) <= 1e-12
    assert abs(s.norm - 1.0) <= 1e-12
    for first in (0, (1 << 20) + 7, (1 << 32) - 4096, (1 << n) - 65536):
        raw, nrm = s.raw(first, 65536)
        want = amp(np.arange(first, first + 65536, dtype=np.uint64))
        assert np.max(np.abs(raw / nrm - want)) <= 1e-12, first
    del s
    capi.synchronize()
