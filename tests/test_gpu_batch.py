"""The batched engine (dk_batch_*): many registers of uneven size in one
persistent launch, checked register by register against the oracle running
the same gates, density requests and rolls.  Complex128: states and
densities within 1e-12, outcomes identical.  Complex64: states within 5e-6,
outcomes identical wherever no roll lies within 1e-4 of its P1."""
import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, oracle_run, reduced_density_einsum

pytestmark = pytest.mark.gpu


def make_batch_specs(seed, count, nmin=4, nmax=16):
    rng = W.rng_for(seed)
    specs = []
    for i in range(count):
        # log-uniform qubit count (p ~ 1/n), then the mixed random circuit
        ns = np.arange(nmin, nmax + 1)
        p = 1.0 / ns
        n = int(rng.choice(ns, p=p / p.sum()))
        depth = int(rng.integers(10, 61))
        steps, rolls = W.random_circuit(seed * 1000 + i, n, depth)
        specs.append((n, steps, rolls))
    return specs


def check_batch(capi, oracle, dtype, specs, tol, **settings):
    b = capi.Batch(dtype, export_states=1, **settings)
    keeps = []
    for n, steps, rolls in specs:
        keep = tuple(range(min(6, n))) if n == 14 else ()
        keeps.append(keep)
        b.add(n, steps, rolls, keep)
    b.run()
    for i, (n, steps, rolls) in enumerate(specs):
        st = oracle_run(oracle, oracle.new(n), steps)
        want = st.values()
        got = b.state(i)
        assert np.max(np.abs(got - want)) <= tol, (i, n)
        outs, p1 = b.outcomes(i)
        rc, left, olst = oracle.measure(st, (1 << n) - 1, rolls)
        if dtype == capi.C128 or min(abs(p1[k] - rolls[k]) for k in range(n)) > 1e-4:
            assert outs == olst, (i, n)
        if keeps[i]:
            rho = b.density(i)
            assert np.max(np.abs(rho - reduced_density_einsum(want, keeps[i]))) <= \
                max(tol, ATOL)
    return b


def test_batch_c128_mixed_sizes(capi, oracle):
    specs = make_batch_specs(11, 70)
    specs += [(14, *W.random_circuit(5001, 14, 40)), (13, *W.random_circuit(5002, 13, 30)),
              (16, *W.random_circuit(5003, 16, 25))]
    check_batch(capi, oracle, capi.C128, specs, ATOL)


def test_batch_c128_small_passes(capi, oracle):
    # few target qubits per pass: many cache-blocked passes per register
    specs = [(15, *W.random_circuit(6000 + i, 15, 30)) for i in range(3)]
    specs += [(9, *W.random_circuit(6100, 9, 30))]
    check_batch(capi, oracle, capi.C128, specs, ATOL, pass_qubits=3)


def test_batch_c64(capi, oracle):
    specs = make_batch_specs(12, 40, 4, 15)
    specs += [(14, *W.random_circuit(7001, 14, 40)), (15, *W.random_circuit(7002, 15, 30))]
    check_batch(capi, oracle, capi.C64, specs, 5e-6)


def test_batch_graph_replays_match(capi, oracle):
    specs = make_batch_specs(13, 30, 4, 14)
    ref = capi.Batch(capi.C128)
    g = capi.Batch(capi.C128, use_graph=1)
    for n, steps, rolls in specs:
        ref.add(n, steps, rolls)
        g.add(n, steps, rolls)
    ref.run()
    want = [ref.outcomes(i) for i in range(len(specs))]
    for rep in range(3):
        g.run()
        for i in range(len(specs)):
            outs, p1 = g.outcomes(i)
            assert outs == want[i][0]
            assert np.array_equal(p1, want[i][1])
        # eager work on the same stream between replays
        s = capi.new(12)
        capi.apply(s, capi.Gate(W.gate_h()), [3], in_place=True)
        capi.synchronize()


@pytest.mark.parametrize("dtype,split,pq", [("c128", 13, 0), ("c128", 13, 3), ("c64", 14, 0)])
def test_batch_split_rounds_equal_single_workgroup(capi, oracle, dtype, split, pq):
    """Registers at or above split_qubits run their gates as grid-wide
    rounds of cache-blocked passes (planned per register, gates merged);
    states, P1 and densities agree with the one-workgroup form to rounding
    (1e-12 complex128, 5e-6 complex64), outcomes are identical up to the
    first roll within 1e-5 of its P1, graph replays repeat each form bit for
    bit, and the mixed batch keeps its small registers in the persistent
    launch."""
    dt = capi.C128 if dtype == "c128" else capi.C64
    rng = W.rng_for(9900 + split + pq)
    specs = []
    for i, n in enumerate([13, 14, 15, 16, 17, 18, 9, 5, 14, 19]):
        steps, rolls = W.random_circuit(9900 + 17 * i + pq, n, int(rng.integers(10, 80)))
        specs.append((n, steps, rolls))
    res = {}
    for mode in (0, split):
        b = capi.Batch(dt, export_states=1, use_graph=1, split_qubits=mode,
                       pass_qubits=pq)
        for n, steps, rolls in specs:
            b.add(n, steps, rolls, tuple(range(6)) if n == 14 else ())
        runs = []
        for rep in range(2):
            b.run()
            runs.append([(b.state(i), b.outcomes(i), b.density(i) if specs[i][0] == 14 else None)
                         for i in range(len(specs))])
        res[mode] = runs
    tol = ATOL if dt == capi.C128 else 5e-6
    for i, (n, _, rolls) in enumerate(specs):
        for mode in (0, split):
            (sa, (oa, pa), da), (sb, (ob, pb), db) = res[mode][0][i], res[mode][1][i]
            assert np.array_equal(sa, sb) and oa == ob and np.array_equal(pa, pb)
            if da is not None:
                assert np.array_equal(da, db)
        (s0, (o0, p0), d0), (s1, (o1, p1), d1) = res[0][0][i], res[split][0][i]
        assert np.max(np.abs(s1 - s0)) <= tol, (i, n)
        assert np.max(np.abs(p1 - p0)) <= tol, (i, n)
        close = [abs(p0[k] - rolls[k]) < 1e-5 for k in range(n)]
        upto = close.index(True) if any(close) else n
        assert o1[:upto] == o0[:upto], (i, n)
        if d0 is not None:
            assert np.max(np.abs(d1 - d0)) <= tol, (i, n)
    if dt == capi.C128:
        for i in (4, 9):
            n, steps, rolls = specs[i]
            want = oracle_run(oracle, oracle.new(n), steps).values()
            assert np.max(np.abs(res[split][1][i][0] - want)) <= ATOL, n


def test_batch_default_split_vs_oracle(capi, oracle):
    """The default threshold (14 qubits: the 16-, 17- and 18-qubit registers
    run as rounds) with the persistent registers beside it, against the
    oracle."""
    specs = [(17, *W.random_circuit(9801, 17, 60)), (18, *W.random_circuit(9802, 18, 45)),
             (12, *W.random_circuit(9803, 12, 30)), (16, *W.random_circuit(9804, 16, 30))]
    check_batch(capi, oracle, capi.C128, specs, ATOL)


@pytest.mark.parametrize("dtype_name", ["c128", "c64"])
def test_batch_densities_scattered_keeps(capi, oracle, dtype_name):
    """Reduced densities over scattered kept qubits (one to six, the top
    qubit among them) of registers held in LDS, in the pool and split into
    rounds, against the contraction of the oracle's final state; the pooled
    and split ones are summed from chunks of the state gathered into LDS."""
    dtype = capi.C128 if dtype_name == "c128" else capi.C64
    tol = ATOL if dtype == capi.C128 else 5e-6
    cases = [(9, (0, 8)), (12, (3, 11)), (13, (1, 4, 9, 12)), (13, (0,)),
             (14, (0, 3, 5, 8, 10, 13)), (14, (13,)), (15, (2, 7)),
             (16, (0, 1, 2, 3, 4, 15)), (17, (5, 6, 16))]
    # NaN over the (uncleared) pool: a read of an unwritten amplitude shows
    b = capi.Batch(dtype, poison_pool=1)
    specs = []
    for i, (n, keep) in enumerate(cases):
        steps, rolls = W.random_circuit(7700 + i, n, 40)
        specs.append((n, steps, keep))
        b.add(n, steps, rolls, keep)
    b.run()
    for i, (n, steps, keep) in enumerate(specs):
        want = oracle_run(oracle, oracle.new(n), steps).values()
        rho = b.density(i)
        assert rho.shape == (1 << len(keep), 1 << len(keep))
        assert np.max(np.abs(rho - reduced_density_einsum(want, keep))) <= tol, (i, n, keep)


def test_batch_nonfinite_pass_over_unwritten_pool(capi, oracle):
    """A gate with an infinite entry, controlled by a qubit no gate sets: as
    the reference computes it, the entry multiplies the zero amplitudes under
    the control (0 * inf is NaN) and the normalised state is NaN throughout.
    Its pass in the rounds runs every tile (a non-finite pass skips none),
    over a pool that is cleared (exported states) or full of NaN (not
    cleared): both runs give the reference's result, NaN for NaN, and the
    same outcomes, probabilities and densities."""
    inf_gate = np.array([[np.inf, 0.0], [0.0, 1.0]], dtype=np.complex128)
    specs = []
    for i, (n, at) in enumerate(((15, 3), (16, 25), (17, 60), (14, 0))):
        rng = W.rng_for(8800 + i)
        steps = [W.random_gate_step(rng, n - 1) for _ in range(70)]
        steps.insert(at, (inf_gate, [int(rng.integers(0, n - 1))], {n - 1}, set()))
        specs.append((n, steps, rng.random(n).tolist()))
    runs = []
    for settings in (dict(export_states=1), dict(poison_pool=1)):
        b = capi.Batch(capi.C128, probs=1, **settings)
        for n, steps, rolls in specs:
            b.add(n, steps, rolls, (0, 2, n - 1))
        b.run()
        runs.append(b)
    clean, poisoned = runs
    for i, (n, steps, rolls) in enumerate(specs):
        want = oracle_run(oracle, oracle.new(n), steps).values()
        assert np.isnan(want).all()
        assert np.array_equal(clean.state(i), want, equal_nan=True), (i, n)
        o0, p0 = clean.outcomes(i)
        o1, p1 = poisoned.outcomes(i)
        assert o1 == o0 and np.array_equal(p1, p0, equal_nan=True), (i, n)
        assert np.array_equal(poisoned.probs(i), clean.probs(i), equal_nan=True), (i, n)
        assert np.array_equal(poisoned.density(i), clean.density(i), equal_nan=True), (i, n)
