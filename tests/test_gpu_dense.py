"""Dense k-qubit gates on the matrix cores, the complex64 mode and fused
circuit runs, on a gfx950 GPU, through the C ABI.

FP64 results are held to 1e-12 max-abs against the oracle.  The complex64
mode stores amplitudes in FP32 and accumulates products in FP32 (norms and
probabilities in FP64); its error against the FP64 oracle is bounded by
5e-6 max-abs per amplitude for these circuit sizes (FP32 unit roundoff is
6e-8, and each amplitude passes through at most a few hundred FP32 FMAs).
"""
import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, oracle_run

pytestmark = pytest.mark.gpu

C64_TOL = 5e-6


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


@pytest.mark.parametrize("form", [0, 2])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("k", [4, 5, 6])
def test_dense_forms_vs_oracle(capi, oracle, form, dtype, k):
    """Both dense forms (0: LDS tiles, 2: streaming from HBM into registers)
    in both precisions, whatever the default dispatch picks, against the
    oracle: in place and out of place, targets anywhere (lowest qubits in
    reverse order, spread out, the highest ones)."""
    n = 15
    dt = capi.C128 if dtype == "c128" else capi.C64
    tol = ATOL if dt == capi.C128 else C64_TOL
    rng = W.rng_for(9300 + 7 * k + form)
    psi = W.random_normalised_state(93 + k, n)
    o = oracle.from_data(n, psi)
    d = capi.convert(capi.from_host(psi), dt)
    capi.tune("dense_stream", form)
    try:
        for targets in (list(range(k))[::-1],
                        [int(x) for x in rng.choice(n, size=k, replace=False)],
                        list(range(n - k, n))):
            u = W.haar_unitary(rng, k)
            g = capi.Gate(u)
            d = capi.apply(d, g, targets)
            d = capi.apply(d, g, targets, in_place=True)
            o = oracle.apply(oracle.apply(o, u, targets), u, targets)
            assert maxabs(d.values(), o.values()) <= tol, targets
            assert abs(d.norm - 1.0) <= tol
    finally:
        capi.tune("dense_stream", 1)


@pytest.mark.parametrize("k", [4, 5, 6])
@pytest.mark.parametrize("n", [12, 13, 16])
def test_dense_mfma_vs_oracle(capi, oracle, k, n):
    rng = W.rng_for(9100 + 10 * n + k)
    psi = W.random_normalised_state(91 + n + k, n) * 1.3
    d = capi.from_host(psi)
    o = oracle.from_data(n, psi)
    for rep in range(3):
        targets = [int(x) for x in rng.choice(n, size=k, replace=False)]
        if rep == 0:
            targets = list(range(k))[::-1]       # lowest qubits, reversed order
        u = W.haar_unitary(rng, k)
        g = capi.Gate(u)
        d = capi.apply(d, g, targets)
        o = oracle.apply(o, u, targets)
        assert maxabs(d.values(), o.values()) <= ATOL, targets
        assert abs(d.norm - o.norm) <= ATOL


def test_dense_mfma_equals_vector_path(capi):
    n = 14
    psi = W.random_normalised_state(5, n)
    rng = W.rng_for(6)
    for k in (4, 5, 6):
        u = W.haar_unitary(rng, k)
        targets = [int(x) for x in rng.choice(n, size=k, replace=False)]
        g = capi.Gate(u)
        a = capi.apply(capi.from_host(psi), g, targets)
        capi.tune("dense_mfma", 0)
        try:
            b = capi.apply(capi.from_host(psi), g, targets)
        finally:
            capi.tune("dense_mfma", 1)
        assert maxabs(a.values(), b.values()) <= 1e-13


@pytest.mark.parametrize("n", [9, 14])
def test_c64_gates_and_probabilities(capi, oracle, n):
    rng = W.rng_for(300 + n)
    psi = W.random_normalised_state(30 + n, n)
    d = capi.convert(capi.from_host(psi), capi.C64)
    assert capi.dtype_of(d) == capi.C64
    o = oracle.from_data(n, psi)
    steps = []
    for t in (0, 3, 6, n - 1):
        steps.append((W.gate_u3(*(rng.random(3) * 6)), [t], set(), set()))
    steps.append((W.gate_x(), [1], {2}, {4}))
    for m, t, c, a in steps:
        d = capi.apply(d, capi.Gate(m), t, c, a)
        o = oracle.apply(o, m, t, sorted(c), sorted(a))
    if n >= 12:
        for k in (4, 6):
            u = W.haar_unitary(rng, k)
            tg = [int(x) for x in rng.choice(n, size=k, replace=False)]
            d = capi.apply(d, capi.Gate(u), tg)
            o = oracle.apply(o, u, tg)
    assert maxabs(d.values(), o.values()) <= C64_TOL
    pa = capi.prob_all(d)
    assert maxabs(pa, [oracle.prob(o, q) for q in range(n)]) <= 2 * C64_TOL
    back = capi.convert(d, capi.C128)
    assert maxabs(back.values(), o.values()) <= C64_TOL


def test_c64_measure_outcomes(capi, oracle):
    n = 10
    steps, _ = W.random_circuit(77, n, 40)
    st = oracle_run(oracle, oracle.new(n), steps)
    d = capi.convert(capi.from_host(st.values()), capi.C64)
    rng = W.rng_for(78)
    compared = 0
    for rep in range(6):
        rolls = rng.random(n)
        left, lst, p1 = capi.measure(d, (1 << n) - 1, rolls)
        rc, oleft, olst = oracle.measure(st, (1 << n) - 1, rolls)
        # list slot i is decided on roll i: outcomes agree up to the first
        # slot whose roll lies within FP32 error of its P1 (after a flip the
        # two collapse chains differ)
        close = [abs(p1[i] - rolls[i]) < 1e-5 for i in range(n)]
        upto = close.index(True) if any(close) else n
        assert lst[:upto] == olst[:upto], rep
        compared += upto == n
    assert compared >= 4     # the check is not vacuous


@pytest.mark.parametrize("k", [0, 3, 4, 5, 6])
def test_circuit_run_fused_vs_oracle(capi, oracle, k):
    n = 14
    steps, _ = W.random_circuit(4400 + k, n, 150)
    circ = capi.Circuit(n, steps)
    nb = circ.fuse(k) if k >= 4 else circ.fuse(0)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    out = circ.run(capi.new(n))
    assert maxabs(out.values(), want) <= ATOL
    if k >= 4:
        assert nb < len(steps)
    # in place on a fresh state gives the same
    s2 = capi.new(n)
    circ.run(s2, in_place=True)
    assert maxabs(s2.values(), want) <= ATOL


def test_circuit_c64_fused(capi, oracle):
    n = 16
    steps = W.brickwork_circuit(8, n, 6)
    circ = capi.Circuit(n, steps)
    circ.fuse(5)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    out = circ.run(capi.new_typed(n, capi.C64))
    assert maxabs(out.values(), want) <= C64_TOL


@pytest.mark.parametrize("cw", [1, 2, 4])
def test_dense_k4_column_sets_vs_oracle(capi, oracle, cw):
    """Every block shape of the 4-qubit dense kernel (64, 128 or 256 groups
    per block) against the oracle, in both amplitude types."""
    n = 13
    rng = W.rng_for(7700 + cw)
    psi = W.random_normalised_state(77 + cw, n)
    capi.tune("dense_k4_cw", cw)
    try:
        for dtype, tol in ((capi.C128, ATOL), (capi.C64, C64_TOL)):
            d = capi.convert(capi.from_host(psi), dtype)
            o = oracle.from_data(n, psi)
            for rep in range(3):
                targets = [int(x) for x in rng.choice(n, size=4, replace=False)]
                u = W.haar_unitary(rng, 4)
                d = capi.apply(d, capi.Gate(u), targets)
                o = oracle.apply(o, u, targets)
                assert maxabs(d.values(), o.values()) <= tol, (dtype, targets)
    finally:
        capi.tune("dense_k4_cw", 0)


@pytest.mark.parametrize("form", [0, 2])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("k", [4, 5, 6])
def test_dense_zero_blocks_of_fresh_states(capi, oracle, form, dtype, k):
    """Dense gates on a new |0...0> state skip the tiles (column blocks)
    outside the state's support, out of place (zeros written) and in place
    (zeros left where they are): every result equals the one on an input the
    library cannot know to be sparse, and the oracle."""
    n = 16
    dt = capi.C128 if dtype == "c128" else capi.C64
    tol = ATOL if dt == capi.C128 else C64_TOL
    rng = W.rng_for(9500 + 7 * k + form)
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1.0
    fresh = capi.new_typed(n, dt)
    dense = capi.convert(capi.from_host(v), dt)
    o = oracle.new(n)
    capi.tune("dense_stream", form)
    try:
        for i, targets in enumerate((list(range(n - k, n)), list(range(3, 3 + k)),
                                     [int(x) for x in rng.choice(n, size=k, replace=False)])):
            u = W.haar_unitary(rng, k)
            g = capi.Gate(u)
            in_place = i % 2 == 1
            fresh = capi.apply(fresh, g, targets, in_place=in_place)
            dense = capi.apply(dense, g, targets, in_place=in_place)
            o = oracle.apply(o, u, targets)
            a = fresh.values()
            assert np.array_equal(a, dense.values()), targets
            assert maxabs(a, o.values()) <= tol, targets
    finally:
        capi.tune("dense_stream", 1)
