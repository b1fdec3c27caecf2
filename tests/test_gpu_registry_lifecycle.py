"""registry_clone and registry_del through the drop-in module, on the GPU.

registry_clone (doki.c:343-398 -> state_clone, qstate.c:90-105) writes the
source's normalised values (raw / norm_const) into a new state whose norm
constant is 1; the oracle's orc_clone restates it.  Checked on an
unnormalised state made by registry_new_data (norm_const 1, so the clone
keeps the unnormalised values), on a collapsed state (norm_const sqrt(P)
!= 1), and on the end state of the golden C0 circuits.

registry_del (doki.c:400-415) frees the state at once; the capsule stays
valid as an object.  Here the freed bytes leave the device-memory budget at
once, and any later use of the capsule raises doki.error (the reference
dereferences the freed struct; the strict error is deliberate).
"""
import ctypes as C

import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, doki_run, doki_values, load, unpack_steps

pytestmark = pytest.mark.gpu


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def used_bytes(capi):
    used = C.c_uint64()
    capi.check(capi.lib().dk_memory_budget(None, C.byref(used)))
    capi.synchronize()
    return used.value


def test_clone_of_unnormalised_data(doki, oracle):
    n = 9
    rng = W.rng_for(61)
    vals = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)) * 0.37
    reg = doki.registry_new_data(n, vals, False)
    cl = doki.registry_clone(reg, 1, False)
    want = oracle.clone(oracle.from_data(n, vals))
    assert want.norm == 1.0
    assert maxabs(doki_values(doki, cl, n), want.values()) <= ATOL
    # Q3: norm_const is 1, so the unnormalised values are copied as they are
    assert maxabs(doki_values(doki, cl, n), vals) <= ATOL
    assert abs(np.linalg.norm(vals) - 1.0) > 0.1
    # the clone behaves as its source under a gate (both read raw / 1)
    u = W.gate_u3(0.3, 1.2, 2.9)
    a = doki_run(doki, cl, [(u, [4], {1}, {7})])
    b = doki_run(doki, reg, [(u, [4], {1}, {7})])
    assert maxabs(doki_values(doki, a, n), doki_values(doki, b, n)) <= ATOL
    o = oracle.apply(want, u, [4], [1], [7])
    assert maxabs(doki_values(doki, a, n), o.values()) <= ATOL


@pytest.mark.parametrize("mask", [0b1010, 0b1])
def test_clone_of_collapsed_state(doki, oracle, capi, mask):
    n = 10
    steps, rolls = W.random_circuit(62 + mask, n, 90)
    reg = doki_run(doki, doki.registry_new(n, False), steps)
    nmeas = bin(mask).count("1")
    left, _ = doki.registry_measure(reg, mask, rolls[:nmeas], 1, False)
    cl = doki.registry_clone(left, 1, False)
    ost = oracle.new(n)
    for m, t, c, a in steps:
        ost = oracle.apply(ost, m, t, sorted(c), sorted(a))
    rc, oleft, _ = oracle.measure(ost, mask, rolls[:nmeas])
    assert abs(oleft.norm - 1.0) > 1e-6    # a collapsed state: norm sqrt(P)
    want = oracle.clone(oleft)
    assert want.norm == 1.0
    m = n - nmeas
    assert maxabs(doki_values(doki, cl, m), want.values()) <= ATOL
    assert maxabs(doki_values(doki, cl, m), doki_values(doki, left, m)) <= ATOL
    # the same through the C ABI: the clone's norm constant is exactly 1
    s = capi.new(n)
    for mm, t, c, a in steps:
        capi.apply(s, capi.Gate(mm), t, c, a, in_place=True)
    cs, _, _ = capi.measure(s, mask, rolls[:nmeas])
    assert abs(cs.norm - oleft.norm) <= ATOL
    cc = capi.clone(cs)
    assert cc.norm == 1.0
    raw, nrm = cc.raw()
    assert nrm == 1.0 and maxabs(raw, want.values()) <= ATOL


@pytest.mark.parametrize("name", ["c0_n12_seed0.npz", "c0_n12_seed2.npz"])
def test_clone_of_c0_golden(doki, name):
    z = load(name)
    n = int(z["n"])
    reg = doki_run(doki, doki.registry_new(n, False), unpack_steps(z))
    cl = doki.registry_clone(reg, -1, False)
    assert maxabs(doki_values(doki, cl, n), z["amps"]) <= ATOL
    assert maxabs(doki_values(doki, cl, n, True), z["canon"]) <= ATOL
    assert maxabs([doki.registry_prob(cl, q, 1, False) for q in range(n)],
                  z["probs"]) <= ATOL
    # deleting the source leaves the clone intact
    doki.registry_del(reg, False)
    assert maxabs(doki_values(doki, cl, n), z["amps"]) <= ATOL


def test_clone_error_texts(doki):
    reg = doki.registry_new(3, False)
    with pytest.raises(Exception) as e:
        doki.registry_clone(reg, 0, False)
    assert type(e.value).__module__ + "." + type(e.value).__name__ == "qsimov.doki.error"
    with pytest.raises(Exception) as e:
        doki.registry_clone(None, 1, False)
    assert "NULL pointer to source registry" in str(e.value)


def test_registry_del_frees_at_once(doki, capi):
    n = 22                                  # 64 MiB: an arena block
    base = used_bytes(capi)
    reg = doki.registry_new(n, False)
    grown = used_bytes(capi) - base
    assert grown >= 16 << n
    assert doki.registry_mem(reg, False) == 64 + (16 << n)
    doki.registry_del(reg, False)
    assert used_bytes(capi) == base         # freed eagerly, capsule still alive
    for call in (lambda: doki.registry_get(reg, 0, False, False),
                 lambda: doki.registry_prob(reg, 0, 1, False),
                 lambda: doki.registry_clone(reg, 1, False),
                 lambda: doki.registry_apply(reg, doki.gate_new(1, W.gate_h().tolist(), False),
                                             [0], set(), set(), 1, False),
                 lambda: doki.registry_measure(reg, 1, [0.5], 1, False),
                 lambda: doki.registry_join(reg, reg, 1, False),
                 lambda: doki.registry_density(reg, False),
                 lambda: doki.registry_mem(reg, False)):
        with pytest.raises(Exception) as e:
            call()
        assert type(e.value).__module__ + "." + type(e.value).__name__ == \
            "qsimov.doki.error"
    # deleting twice is harmless, and dropping the capsule afterwards too
    doki.registry_del(reg, False)
    del reg
    assert used_bytes(capi) == base


def test_registry_del_under_a_density(doki):
    """A density node made before the delete holds the capsule (as the
    reference's node does, qops.c:592) but not the freed state: reading it
    afterwards raises doki.error instead of reading freed memory (the
    reference reads the freed struct)."""
    n = 5
    steps, _ = W.random_circuit(63, n, 40)
    reg = doki_run(doki, doki.registry_new(n, False), steps)
    psi = doki_values(doki, reg, n)
    rho = doki.registry_density(reg, False)
    keep = doki.registry_density(reg, False)
    got = np.array([[doki.funmatrix_get(keep, i, j, False) for j in range(1 << n)]
                    for i in range(1 << n)])
    assert maxabs(got, np.outer(psi, psi.conj())) <= ATOL
    doki.registry_del(reg, False)
    with pytest.raises(Exception) as e:
        doki.funmatrix_get(rho, 0, 0, False)
    assert type(e.value).__module__ + "." + type(e.value).__name__ == \
        "qsimov.doki.error"
    assert "deleted" in str(e.value)
