"""Deferred execution behind the drop-in module.

registry_apply returns a registry whose gates run when something reads it
(execution_mode "fuse": the chain as cache-blocked passes; "gates": one gate
per call, deferred by one call; "eager": at once, out of place).  Whatever
the mode, every registry must read the reference's value: the golden C0
vectors and the oracle on the same gates, within 1e-12, outcomes identical.
Branches taken from intermediate registries, densities of intermediate
registries and registry_del in the middle of a chain must not disturb any
other registry; chains whose source nothing else can read run in place (no
second state is allocated).
"""
import ctypes as C
import gc

import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, doki_values, load, oracle_run, unpack_steps

pytestmark = pytest.mark.gpu

MODES = ["fuse", "gates", "eager"]


@pytest.fixture
def mode(doki, request):
    prev = doki.execution_mode(request.param, False)
    yield request.param
    doki.execution_mode(prev, False)


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def gates_of(doki, steps):
    return [(doki.gate_new(len(t), np.asarray(m).tolist(), False), list(t), set(c), set(a))
            for m, t, c, a in steps]


def chain(doki, reg, gates):
    for g, t, c, a in gates:
        reg = doki.registry_apply(reg, g, t, c, a, 1, False)
    return reg


def used_bytes(capi):
    """Bytes of live device states; collects first, so that states held by
    reference cycles of earlier tests (exception tracebacks) are gone."""
    gc.collect()
    used = C.c_uint64()
    capi.check(capi.lib().dk_memory_budget(None, C.byref(used)))
    return used.value


@pytest.mark.parametrize("mode", MODES, indirect=True)
@pytest.mark.parametrize("name", ["c0_n12_seed0.npz", "c0_n12_seed1.npz", "c0_n6_seed7.npz"])
def test_c0_golden_each_mode(doki, mode, name):
    z = load(name)
    n = int(z["n"])
    reg = chain(doki, doki.registry_new(n, False), gates_of(doki, unpack_steps(z)))
    assert maxabs(doki_values(doki, reg, n), z["amps"]) <= ATOL
    assert maxabs(doki_values(doki, reg, n, True), z["canon"]) <= ATOL
    assert maxabs([doki.registry_prob(reg, q, 1, False) for q in range(n)],
                  z["probs"]) <= ATOL
    left, lst = doki.registry_measure(reg, (1 << n) - 1, z["rolls"].tolist(), 1, False)
    assert left is None
    assert lst == [None if x < 0 else bool(x) for x in z["measure_all"]]


@pytest.mark.parametrize("mode", MODES, indirect=True)
def test_branches_and_intermediates(doki, oracle, mode):
    """R -> X (30 gates) -> Y (20 more) and X -> Z (20 others); W is a
    pending registry read through a density only.  Every registry reads its
    own value, read in an order that materialises branches first."""
    n = 13
    s1, _ = W.random_circuit(71, n, 30)
    s2, _ = W.random_circuit(72, n, 20)
    s3, _ = W.random_circuit(73, n, 20)
    s4 = [(W.haar_unitary(W.rng_for(74), 3), [0, 5, 11], {2}, set())]
    R = doki.registry_new(n, False)
    X = chain(doki, R, gates_of(doki, s1))
    Y = chain(doki, X, gates_of(doki, s2))
    Z = chain(doki, X, gates_of(doki, s3))
    Wd = chain(doki, Z, gates_of(doki, s4))
    D = doki.registry_density(Wd, False)
    del Wd
    ox = oracle_run(oracle, oracle.new(n), s1)
    oy = oracle_run(oracle, ox, s2)
    oz = oracle_run(oracle, ox, s3)
    ow = oracle_run(oracle, oz, s4)
    assert maxabs(doki_values(doki, Z, n), oz.values()) <= ATOL
    assert maxabs(doki_values(doki, Y, n), oy.values()) <= ATOL
    assert maxabs(doki_values(doki, X, n), ox.values()) <= ATOL
    assert maxabs(doki_values(doki, R, n), oracle.new(n).values()) == 0.0
    w = ow.values()
    for i, j in ((0, 0), (5, 77), (4095, 8191), (1234, 5)):
        assert abs(doki.funmatrix_get(D, i, j, False) - w[i] * np.conj(w[j])) <= ATOL


@pytest.mark.parametrize("mode", ["fuse", "gates"], indirect=True)
def test_dead_source_runs_in_place(doki, capi, mode):
    """A chain whose source registry is referenced by nothing else runs on
    that registry's state: the device memory in use never exceeds one state
    of 64 MiB (22 qubits), and the values are right."""
    n = 22
    steps, _ = W.random_circuit(75, n, 40)
    gates = gates_of(doki, steps)
    capi.synchronize()
    base = used_bytes(capi)
    reg = doki.registry_new(n, False)
    one = used_bytes(capi) - base
    assert one >= 16 << n
    for g, t, c, a in gates:
        reg = doki.registry_apply(reg, g, t, c, a, 1, False)
        assert used_bytes(capi) - base <= one
    doki.registry_sync(reg, False)
    assert used_bytes(capi) - base == one
    probs = [doki.registry_prob(reg, q, 1, False) for q in range(n)]
    # the same gates through the C ABI, gate by gate
    s = capi.new(n)
    for m, t, c, a in steps:
        capi.apply(s, capi.Gate(m), t, c, a, in_place=True)
    assert maxabs(probs, capi.prob_all(s)) <= ATOL
    del reg
    assert used_bytes(capi) - base == one       # only the C ABI state is left


@pytest.mark.parametrize("mode", MODES, indirect=True)
def test_referenced_source_is_kept(doki, oracle, mode):
    n = 10
    steps, _ = W.random_circuit(76, n, 25)
    psi = W.random_normalised_state(77, n)
    R = doki.registry_new_data(n, psi, False)
    X = chain(doki, R, gates_of(doki, steps))
    want = oracle_run(oracle, oracle.from_data(n, psi), steps)
    assert maxabs(doki_values(doki, X, n), want.values()) <= ATOL
    assert maxabs(doki_values(doki, R, n), psi) == 0.0


@pytest.mark.parametrize("mode", MODES, indirect=True)
def test_delete_inside_a_chain(doki, oracle, capi, mode):
    """registry_del of a chain's source, and of a pending registry others
    derive from: the deleted registries raise, the others read their values,
    and the deleted source's memory is returned once nothing needs it."""
    n = 22
    s1, _ = W.random_circuit(78, n, 15)
    s2, _ = W.random_circuit(79, n, 15)
    capi.synchronize()
    base = used_bytes(capi)
    R = doki.registry_new(n, False)
    X = chain(doki, R, gates_of(doki, s1))
    doki.registry_del(R, False)
    Y = chain(doki, X, gates_of(doki, s2))
    doki.registry_del(X, False)
    with pytest.raises(Exception) as e:
        doki.registry_prob(X, 0, 1, False)
    assert "deleted" in str(e.value)
    with pytest.raises(Exception) as e:
        doki.registry_prob(R, 0, 1, False)
    assert "deleted" in str(e.value)
    p = [doki.registry_prob(Y, q, 1, False) for q in range(n)]
    s = capi.new(n)
    for m, t, c, a in s1 + s2:
        capi.apply(s, capi.Gate(m), t, c, a, in_place=True)
    assert maxabs(p, capi.prob_all(s)) <= ATOL
    del s
    capi.synchronize()
    # Y's state is the only one left of the chain
    assert used_bytes(capi) - base == (16 << n) + 256


def test_long_chain_and_mode_switch(doki, oracle):
    """More pending gates than one chain holds (2048), and a chain built
    under one mode and finished under others."""
    n = 8
    steps, rolls = W.random_circuit(80, n, 3000)
    gates = gates_of(doki, steps)
    prev = doki.execution_mode("fuse", False)
    try:
        reg = chain(doki, doki.registry_new(n, False), gates[:2500])
        assert doki.execution_mode("gates", False) == "fuse"
        reg = chain(doki, reg, gates[2500:2800])
        assert doki.execution_mode("eager", False) == "gates"
        reg = chain(doki, reg, gates[2800:])
        assert doki.execution_mode(None, False) == "eager"
    finally:
        doki.execution_mode(prev, False)
    want = oracle_run(oracle, oracle.new(n), steps)
    assert maxabs(doki_values(doki, reg, n), want.values()) <= ATOL
    _, lst = doki.registry_measure(reg, (1 << n) - 1, rolls, 1, False)
    _, _, olst = oracle.measure(want, (1 << n) - 1, rolls)
    assert lst == olst


@pytest.mark.parametrize("mode", MODES, indirect=True)
def test_join_measure_clone_of_pending(doki, oracle, mode):
    sa, rolls = W.random_circuit(81, 7, 30)
    sb, _ = W.random_circuit(82, 6, 30)
    A = chain(doki, doki.registry_new(7, False), gates_of(doki, sa))
    B = chain(doki, doki.registry_new(6, False), gates_of(doki, sb))
    J = doki.registry_join(A, B, 1, False)
    JJ = doki.registry_join(B, B, 1, False)
    oa = oracle_run(oracle, oracle.new(7), sa)
    ob = oracle_run(oracle, oracle.new(6), sb)
    assert maxabs(doki_values(doki, J, 13), oracle.join(oa, ob).values()) <= ATOL
    assert maxabs(doki_values(doki, JJ, 12), oracle.join(ob, ob).values()) <= ATOL
    C2 = doki.registry_clone(chain(doki, A, gates_of(doki, sb[:5])), 1, False)
    oc = oracle_run(oracle, oa, sb[:5])
    assert maxabs(doki_values(doki, C2, 7), oc.values()) <= ATOL
    left, lst = doki.registry_measure(chain(doki, A, gates_of(doki, sb[5:9])), 0b1011,
                                      rolls[:3], 1, False)
    rc, oleft, olst = oracle.measure(oracle_run(oracle, oa, sb[5:9]), 0b1011, rolls[:3])
    assert lst == olst
    assert maxabs(doki_values(doki, left, 4), oleft.values()) <= ATOL
    assert doki.registry_mem(chain(doki, A, gates_of(doki, sb[:2])), False) == 64 + (16 << 7)


def test_execution_mode_errors(doki):
    for bad in ("fast", 3, b"fuse"):
        with pytest.raises(Exception) as e:
            doki.execution_mode(bad, False)
        assert type(e.value).__module__ + "." + type(e.value).__name__ == \
            "qsimov.doki.error"
    assert doki.execution_mode(None, False) in ("fuse", "gates", "eager")
    with pytest.raises(Exception) as e:
        doki.registry_sync(None, False)
    assert "NULL pointer to registry" in str(e.value)


def test_dropped_long_chain_releases_without_recursion(doki, capi):
    """A chain of FUSE_MAX_PENDING - 1 pending registries that nobody reads
    is released link by link when its last registry goes: in a thread with a
    128 KiB stack (a recursive release would need a C frame chain per link),
    and with the source's memory back afterwards."""
    import threading
    prev = doki.execution_mode("fuse", False)
    try:
        base = used_bytes(capi)
        x = doki.gate_new(1, W.gate_x().tolist(), False)
        out = {}

        def work():
            try:
                reg = doki.registry_new(14, False)
                for i in range(2047):
                    reg = doki.registry_apply(reg, x, [i % 14], set(), set(), 1, False)
                del reg
                out["ok"] = True
            except BaseException as e:       # reported below
                out["err"] = e

        old = threading.stack_size(128 * 1024)
        try:
            t = threading.Thread(target=work)
            t.start()
            t.join(120)
        finally:
            threading.stack_size(old)
        assert not t.is_alive()
        assert out.get("ok"), out.get("err")
        assert used_bytes(capi) == base
    finally:
        doki.execution_mode(prev, False)
