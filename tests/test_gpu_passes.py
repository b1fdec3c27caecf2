"""Cache-blocked multi-gate passes (k_pass) on a gfx950 GPU, through the C ABI.

complex128 results are held to 1e-12 max-abs against the oracle's
gate-by-gate run; complex64 to 5e-6 per amplitude (FP32 storage and
arithmetic; the same bound as the other complex64 paths).  At 30 qubits the
pass plan is compared with the GPU's own per-gate path on a slice of the
state and through the norm (the oracle would take minutes there).
"""
import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, oracle_run
from test_passes import mixed_steps

pytestmark = pytest.mark.gpu

C64_TOL = 5e-6


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def run_passes(capi, n, steps, dtype, in_place):
    circ = capi.Circuit(n, steps)
    npass = circ.passes()
    st = capi.new_typed(n, dtype)
    out = circ.run(st, in_place=in_place)
    return out, npass, st


@pytest.mark.parametrize("n", [12, 13, 14, 17, 20])
def test_random_circuit_passes_vs_oracle(capi, oracle, n):
    print(f"[passes] random circuit n={n}", flush=True)
    steps, _ = W.random_circuit(1200 + n, n, 150)
    want = oracle_run(oracle, oracle.new(n), steps)
    out, npass, st = run_passes(capi, n, steps, capi.C128, in_place=False)
    assert npass > 0
    assert maxabs(out.values(), want.values()) <= ATOL
    assert abs(out.norm - want.norm) <= 1e-12
    # out of place: the input is untouched
    z = st.values()
    assert z[0] == 1 and not np.any(z[1:])


@pytest.mark.parametrize("n", [12, 15, 18])
def test_mixed_circuit_passes_in_place(capi, oracle, n):
    """Diagonal gates, controls and anticontrols everywhere, and lone
    multi-qubit gates between passes, in place."""
    print(f"[passes] mixed circuit n={n}", flush=True)
    steps = mixed_steps(50 + n, n, 220)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    out, npass, st = run_passes(capi, n, steps, capi.C128, in_place=True)
    assert out is st and npass > 0
    assert maxabs(out.values(), want) <= ATOL


def test_passes_on_nonzero_input_norm(capi, oracle):
    """The first pass reads the input through its norm constant."""
    n = 14
    psi = W.random_normalised_state(3, n) * 2.5   # unnormalised input data
    steps, _ = W.random_circuit(77, n, 80)
    d = capi.from_host(psi)
    circ = capi.Circuit(n, steps)
    assert circ.passes() > 0
    d = circ.run(d)
    want = oracle_run(oracle, oracle.from_data(n, psi), steps).values()
    assert maxabs(d.values(), want) <= ATOL


def test_brickwork_passes(capi, oracle):
    n = 16
    print("[passes] brickwork n=16", flush=True)
    steps = W.brickwork_circuit(44, n, 6)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    out, npass, _ = run_passes(capi, n, steps, capi.C128, in_place=False)
    assert npass * 6 < len(steps)
    assert maxabs(out.values(), want) <= ATOL


@pytest.mark.parametrize("n", [12, 16])
def test_c64_passes(capi, oracle, n):
    print(f"[passes] complex64 n={n}", flush=True)
    steps = mixed_steps(600 + n, n, 160)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    out, npass, _ = run_passes(capi, n, steps, capi.C64, in_place=True)
    assert npass > 0
    assert maxabs(out.values(), want) <= C64_TOL


def test_passes_then_measure_and_prob(capi, oracle):
    """States written by passes feed the other operations as usual."""
    n = 13
    steps, rolls = W.random_circuit(4242, n, 120)
    out, _, _ = run_passes(capi, n, steps, capi.C128, in_place=True)
    o = oracle_run(oracle, oracle.new(n), steps)
    p = capi.prob_all(out)
    assert maxabs(p, [oracle.prob(o, q) for q in range(n)]) <= ATOL
    _, outs, _ = capi.measure(out, (1 << n) - 1, rolls)
    _, _, want = oracle.measure(o, (1 << n) - 1, rolls)
    assert outs == want


def test_30_qubit_passes_equal_per_gate(capi):
    """Full-size register: a pass plan against plain per-gate application on
    the GPU (both complex128), compared on the first 2^22 amplitudes, on
    2^16 amplitudes spread over the state, and through the norm constant."""
    n = 30
    print("[passes] 30-qubit brickwork depth 2 (two 17 GB states)", flush=True)
    steps = W.brickwork_circuit(303, n, 2) + W.random_circuit(304, n, 40)[0]
    circ = capi.Circuit(n, steps)
    assert circ.passes() > 0
    a = capi.new(n)
    circ.run(a, in_place=True)
    plain = capi.Circuit(n, steps)
    b = capi.new(n)
    plain.run(b, in_place=True)
    ra, na = a.raw(0, 1 << 22)
    rb, nb = b.raw(0, 1 << 22)
    assert maxabs(ra / na, rb / nb) <= ATOL
    rng = np.random.default_rng(7)
    for start in rng.integers(0, (1 << n) - 4096, size=16):
        ra, na = a.raw(int(start), 4096)
        rb, nb = b.raw(int(start), 4096)
        assert maxabs(ra / na, rb / nb) <= ATOL
    assert abs(na - 1.0) <= 1e-12 and abs(nb - 1.0) <= 1e-12


def test_time_ring_records_every_launch(capi):
    """The timing ring holds one duration per gate, block or pass launch,
    up to its capacity, and empties on read."""
    n = 16
    steps, _ = W.random_circuit(8, n, 60)
    circ = capi.Circuit(n, steps)
    circ.passes()
    nsteps = len(circ.plan())
    st = capi.new(n)
    capi.tune("time_ring", 64)
    try:
        circ.run(st, in_place=True)
        ms = capi.time_ring_read()
        assert len(ms) == nsteps and all(m > 0 for m in ms)
        assert capi.time_ring_read() == []
        capi.tune("time_ring", 2)
        circ.run(st, in_place=True)
        assert len(capi.time_ring_read()) == min(2, nsteps)
    finally:
        capi.tune("time_ring", 0)


def zero_state_without_support(capi, n, dtype):
    """|0...0> made from host values: the library knows nothing of its zeros,
    so every pass reads and computes every tile."""
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1.0
    s = capi.from_host(v)
    return s if dtype == capi.C128 else capi.convert(s, dtype)


@pytest.mark.parametrize("n,dtype", [(14, 0), (18, 0), (20, 0), (18, 1)])
def test_zero_tiles_of_fresh_states(capi, oracle, n, dtype):
    """From a new |0...0> state the passes write tiles outside the state's
    support as zeros without reading them or running the ops: the result
    equals the same plan run on an input the library cannot know to be
    sparse (bit for bit, zeros of either sign alike), and the oracle."""
    print(f"[passes] zero tiles n={n} dtype={dtype}", flush=True)
    steps = mixed_steps(700 + n, n, 180)
    circ = capi.Circuit(n, steps)
    circ.passes()
    fresh = circ.run(capi.new_typed(n, dtype), in_place=True)
    dense = circ.run(zero_state_without_support(capi, n, dtype), in_place=True)
    a, b = fresh.values(), dense.values()
    assert np.array_equal(a, b)
    assert fresh.norm == dense.norm
    if dtype == capi.C128:
        want = oracle_run(oracle, oracle.new(n), steps).values()
        assert maxabs(a, want) <= ATOL


def test_zero_tiles_not_assumed_with_non_finite_gates(capi):
    """A pass whose matrices hold an infinity computes every tile, as the
    reference does (0 * inf is NaN): the result equals the dense run's,
    NaN for NaN."""
    n = 14
    rng = np.random.default_rng(3)
    steps = [(W.gate_h(), [q], set(), set()) for q in range(n)]
    bad = np.array([[np.inf, 0.0], [0.0, 1.0]], dtype=np.complex128)
    steps.insert(3, (bad, [int(rng.integers(n))], set(), set()))
    circ = capi.Circuit(n, steps)
    circ.passes()
    fresh = circ.run(capi.new(n), in_place=True).raw()[0]
    dense = circ.run(zero_state_without_support(capi, n, capi.C128), in_place=True).raw()[0]
    assert np.array_equal(fresh, dense, equal_nan=True)


def test_zero_tiles_after_a_non_finite_lone_gate(capi):
    """A multi-qubit gate with a NaN entry, run alone between passes, leaves
    nothing known of the state's zeros: the passes after it compute every
    tile, and the result equals the dense run's, NaN for NaN."""
    n = 16
    bad = np.eye(4, dtype=np.complex128)
    bad[1, 2] = np.nan
    steps = [(W.gate_h(), [q], set(), set()) for q in range(0, n, 2)]
    steps.append((bad, [3, 11], set(), set()))
    steps += [(W.gate_h(), [q], set(), set()) for q in range(1, n, 2)]
    circ = capi.Circuit(n, steps)
    circ.passes()
    assert any(kind == "gate" for kind, _ in circ.plan())
    fresh = circ.run(capi.new(n), in_place=True).raw()[0]
    dense = circ.run(zero_state_without_support(capi, n, capi.C128), in_place=True).raw()[0]
    assert np.array_equal(fresh, dense, equal_nan=True)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 11])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_small_circuits_in_one_workgroup(capi, oracle, n, dtype):
    """Below the pass tile (12 qubits) a planned circuit runs each run of
    one-qubit gates in one launch of one workgroup (the state in LDS), lone
    multi-qubit gates between: against the oracle, out of place (the input
    untouched) and in place, on an input whose norm is not 1."""
    dt = capi.C128 if dtype == "c128" else capi.C64
    tol = ATOL if dt == capi.C128 else C64_TOL
    steps, _ = W.random_circuit(3300 + n, n, 90)
    if n >= 3:
        u = W.haar_unitary(W.rng_for(n), 2)
        steps.insert(40, (u, [n - 1, 0], set(), set()))
        steps.insert(70, (u, [1, n - 1], set(), set()))
    v = W.random_normalised_state(3400 + n, n) * 2.5
    want = oracle_run(oracle, oracle.from_data(n, v), steps).values()
    circ = capi.Circuit(n, steps)
    circ.passes()
    src = capi.from_host(v)
    if dt == capi.C64:
        src = capi.convert(src, dt)
    before = src.values()
    out = circ.run(src, in_place=False)
    assert maxabs(out.values(), want) <= tol
    assert np.array_equal(src.values(), before)
    again = circ.run(src, in_place=True)
    assert again is src and maxabs(src.values(), want) <= tol
