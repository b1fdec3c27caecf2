"""The batched engine at full scale: the 4096-register workload (qubit counts
log-uniform on [4, 20], depth U[10, 100], full measurement, 6-qubit densities
on the 14-qubit members) exactly as it is timed, about 500 of its registers
of 17-20 qubits running their passes as grid-wide rounds.

* The split run (rounds of cache-blocked passes, the default) and the
  single-workgroup run (split_qubits = 0) must agree on every register:
  final state, the P1 each outcome was decided on, marginals and densities
  to rounding (1e-12 complex128, 5e-6 complex64), outcomes identical up to
  the first roll within 1e-5 of its P1.
* A sample holding every qubit count 4..20, with at least three 20-qubit
  registers, is checked against the oracle running the same gates and rolls
  (states, marginals and densities within 1e-12, outcomes identical).
"""
import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, oracle_run, reduced_density_einsum

pytestmark = pytest.mark.gpu

SPECS = None


def specs():
    global SPECS
    if SPECS is None:
        SPECS = W.batch_specs(31, 4096)
    return SPECS


def run_batch(capi, dtype, split):
    b = capi.Batch(dtype, export_states=1, probs=1, use_graph=1, split_qubits=split)
    for n, steps, rolls, keep in specs():
        b.add(n, steps, rolls, keep)
    b.run()
    b.run()          # a graph replay: the results checked are the replay's
    return b


def oracle_sample():
    sp = specs()
    pick = {}
    for i, (n, *_rest) in enumerate(sp):
        pick.setdefault(n, []).append(i)
    out = []
    for n in range(4, 21):
        out += pick[n][: (3 if n == 20 else 1)]
    return out


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_full_batch_split_equals_single_workgroup(capi, dtype):
    dt = capi.C128 if dtype == "c128" else capi.C64
    sp = specs()
    assert sum(1 for s in sp if s[0] >= 17) > 400
    split = run_batch(capi, dt, 17)
    single = run_batch(capi, dt, 0)
    print(f"[batch-full] {dtype}: both runs done", flush=True)
    tol = 1e-12 if dtype == "c128" else 5e-6
    cut = 0
    for i, (n, steps, rolls, keep) in enumerate(sp):
        assert np.max(np.abs(split.state(i) - single.state(i))) <= tol, (i, n)
        o1, p1 = split.outcomes(i)
        o0, p0 = single.outcomes(i)
        assert np.max(np.abs(p1 - p0)) <= tol, (i, n)
        close = [abs(p0[k] - rolls[k]) < 1e-5 for k in range(n)]
        upto = close.index(True) if any(close) else n
        cut += upto < n
        assert o1[:upto] == o0[:upto], (i, n)
        assert np.max(np.abs(split.probs(i) - single.probs(i))) <= tol, (i, n)
        if keep:
            assert np.max(np.abs(split.density(i) - single.density(i))) <= tol, (i, n)
    print(f"[batch-full] {dtype}: {len(sp)} registers agree ({cut} cut at a "
          "knife-edge roll)", flush=True)


def test_full_batch_sample_vs_oracle(capi, oracle):
    sp = specs()
    b = run_batch(capi, capi.C128, 17)
    sample = oracle_sample()
    assert sorted({sp[i][0] for i in sample}) == list(range(4, 21))
    assert sum(1 for i in sample if sp[i][0] == 20) >= 3
    for i in sample:
        n, steps, rolls, keep = sp[i]
        st = oracle_run(oracle, oracle.new(n), steps)
        want = st.values()
        assert np.max(np.abs(b.state(i) - want)) <= ATOL, (i, n)
        assert np.max(np.abs(b.probs(i) -
                             [oracle.prob(st, q) for q in range(n)])) <= ATOL, (i, n)
        outs, p1 = b.outcomes(i)
        rc, left, olst = oracle.measure(st, (1 << n) - 1, rolls)
        assert outs == olst, (i, n)
        if keep:
            rho = b.density(i)
            assert np.max(np.abs(rho - reduced_density_einsum(want, keep))) <= ATOL
        print(f"[batch-full] register {i} ({n} qubits) matches the oracle", flush=True)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_full_batch_unwritten_pool_never_read(capi, dtype):
    """Without exported states the pool is not cleared: the rounds leave
    every amplitude outside a register's support unwritten and everything
    after them takes those as zeros without reading them.  With NaN over the
    pool before each run ("poison_pool") every result equals, bit for bit,
    that of a run whose pool was cleared (exported states), in a graph
    replay too."""
    dt = capi.C128 if dtype == "c128" else capi.C64
    sp = specs()
    runs = []
    for settings in (dict(export_states=1), dict(poison_pool=1)):
        b = capi.Batch(dt, probs=1, use_graph=1, **settings)
        for n, steps, rolls, keep in sp:
            b.add(n, steps, rolls, keep)
        b.run()
        b.run()
        runs.append(b)
    clean, poisoned = runs
    for i, (n, steps, rolls, keep) in enumerate(sp):
        o0, p0 = clean.outcomes(i)
        o1, p1 = poisoned.outcomes(i)
        assert o1 == o0 and np.array_equal(p1, p0), (i, n)
        pr = poisoned.probs(i)
        assert not np.isnan(pr).any() and np.array_equal(pr, clean.probs(i)), (i, n)
        if keep:
            assert np.array_equal(poisoned.density(i), clean.density(i)), (i, n)
    print(f"[batch-full] {dtype}: poisoned pool, {len(sp)} registers equal", flush=True)
