"""Cache-blocked pass planner (dk_circuit_passes, host code of libdoki_hip.so)
on the CPU: the plan, executed with the pass op semantics by a numpy model
(tests/pass_emu.py), must reproduce the gate-by-gate oracle run within
1e-12, and every pass must be well formed."""
import math

import numpy as np
import pytest

import doki_amd
from doki_amd import workloads as W
from helpers import ATOL, oracle_run, tensor_apply
from pass_emu import OP_EXCH, run_plan


@pytest.fixture(scope="module")
def capi():
    if not doki_amd.built():
        from doki_amd import build
        build.build()
    from doki_amd import capi as A
    return A


def mixed_steps(seed, n, depth, multi=True):
    """Dense and diagonal one-qubit gates (H, X, RX/RY/RZ, U3, Z, phase) with
    up to two controls and one anticontrol, and (when multi) Haar gates on
    two and three targets, some controlled."""
    rng = W.rng_for(seed)
    steps = []
    for _ in range(depth):
        u = rng.random()
        if multi and u < 0.08:
            k = 2 + int(rng.integers(2))
            qs = [int(q) for q in rng.choice(n, size=k + 1, replace=False)]
            ctrl = {qs[k]} if rng.random() < 0.5 else set()
            steps.append((W.haar_unitary(rng, k), qs[:k], ctrl, set()))
            continue
        if u < 0.35:
            kind = int(rng.integers(3))
            th = float(rng.random() * 2 * math.pi)
            m = (W.gate_rz(th), W.gate_z(),
                 np.diag([1.0, complex(math.cos(th), math.sin(th))]))[kind]
            t = int(rng.integers(n))
            free = [q for q in range(n) if q != t]
            nc = int(rng.integers(3))
            cs = [int(q) for q in rng.choice(free, size=nc, replace=False)]
            rest = [q for q in free if q not in cs]
            acs = {int(rng.choice(rest))} if rng.random() < 0.2 else set()
            steps.append((np.asarray(m, dtype=np.complex128), [t], set(cs), acs))
            continue
        steps.append(W.random_gate_step(rng, n))
    return steps


def lone_apply(psi, step):
    m, t, c, a = step
    return tensor_apply(psi, m, list(t), sorted(c), sorted(a))


def check_plan(oracle, capi, n, steps):
    circ = capi.Circuit(n, steps)
    npass = circ.passes()
    plan = circ.plan()
    # every gate is absorbed by exactly one pass or runs alone
    absorbed = sum(circ.pass_info(i)[0] for i in range(npass))
    lone = sum(1 for kind, _ in plan if kind == "gate")
    assert absorbed + lone == len(steps)
    assert sorted(i for kind, i in plan if kind == "pass") == list(range(npass))
    for i in range(npass):
        ng, nops, tq = circ.pass_info(i)
        assert tq[:6] == list(range(6))
        assert len(set(tq)) == 12 and all(6 <= q < n for q in tq[6:])
        ops = circ.pass_ops(i)
        assert len(ops) == nops and ng >= 1
        # gate ops never outnumber gates (merging only shrinks the list)
        assert int(np.sum(ops["kind"] != OP_EXCH)) <= ng
    psi0 = np.zeros(1 << n, dtype=np.complex128)
    psi0[0] = 1
    got = run_plan(circ, steps, n, psi0, lone_apply)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    err = float(np.max(np.abs(got - want)))
    assert err <= ATOL, err
    return circ, npass


@pytest.mark.parametrize("n,seed", [(12, 1), (13, 2), (14, 3), (16, 4)])
def test_random_circuit_plan_equals_gates(oracle, capi, n, seed):
    steps, _ = W.random_circuit(900 + seed, n, 160)
    _, npass = check_plan(oracle, capi, n, steps)
    assert 0 < npass < len(steps) // 4


@pytest.mark.parametrize("n,seed", [(12, 5), (13, 6), (15, 7)])
def test_mixed_diagonal_and_multi_qubit_plan(oracle, capi, n, seed):
    steps = mixed_steps(seed, n, 200)
    check_plan(oracle, capi, n, steps)


def test_brickwork_plan(oracle, capi):
    n = 14
    steps = W.brickwork_circuit(8, n, 8)
    _, npass = check_plan(oracle, capi, n, steps)
    assert npass * 8 < len(steps)


def test_phase_pairs_fold_into_dense_ops(oracle, capi):
    """A phase on |11> of two qubits (CZ, controlled phase) and a later
    uncontrolled dense gate on one of them become one two-matrix op chosen
    by the other qubit, when that qubit is a register bit or a thread-word
    bit >= 6; every such op is followed by its data record."""
    n = 14
    steps = W.brickwork_circuit(21, n, 6)
    circ, npass = check_plan(oracle, capi, n, steps)
    kinds = np.concatenate([circ.pass_ops(i)["kind"] for i in range(npass)])
    mux = np.isin(kinds, [capi.OP_REG_MUX, capi.OP_LANE_MUX])
    assert mux.sum() > 0
    assert np.array_equal(np.flatnonzero(kinds == capi.OP_MDATA), np.flatnonzero(mux) + 1)
    # both selector forms: a register bit, and a qubit outside the tile
    n = 22
    u = W.gate_u3(0.4, 1.3, -0.2)
    cp = np.diag([1, np.exp(0.7j)])
    steps = [(u, [q], set(), set()) for q in (8, 9, 10, 11)]
    steps += [(cp, [8], {9}, set()), (u, [9], set(), set()),
              (W.gate_z(), [10], {20}, set()), (u, [10], set(), set())]
    circ, npass = check_plan(oracle, capi, n, steps)
    ops = np.concatenate([circ.pass_ops(i) for i in range(npass)])
    sels = [int(o["b"]) for o in ops if int(o["kind"]) == capi.OP_REG_MUX]
    assert any(b & capi.SEL_REG for b in sels) and 20 in sels


def test_controls_everywhere(oracle, capi):
    """Every control/anticontrol location (lane, register, wave, outside
    the tile) on dense and diagonal gates."""
    n = 13
    rng = W.rng_for(77)
    steps = []
    for i in range(240):
        t = int(rng.integers(n))
        free = [q for q in range(n) if q != t]
        cs = [int(q) for q in rng.choice(free, size=2, replace=False)]
        a = int(rng.choice([q for q in free if q not in cs]))
        m = W.gate_u3(*(rng.random(3) * 2 * math.pi)) if i % 3 else W.gate_rz(float(i))
        steps.append((m, [t], set(cs[: 1 + i % 2]), {a} if i % 4 == 0 else set()))
    check_plan(oracle, capi, n, steps)


def test_merges_same_target_runs(oracle, capi):
    """Consecutive gates on one target with equal controls become one op."""
    n = 12
    rng = W.rng_for(3)
    steps = [(W.gate_u3(*(rng.random(3) * 6)), [7], set(), set()) for _ in range(10)]
    steps += [(W.gate_rz(0.3 * i), [2], {9}, set()) for i in range(5)]
    circ, npass = check_plan(oracle, capi, n, steps)
    assert npass == 1
    ng, nops, _ = circ.pass_info(0)
    assert ng == 15 and nops == 2


def test_small_register_runs_gate_by_gate(capi):
    n = 10
    steps, _ = W.random_circuit(5, n, 30)
    circ = capi.Circuit(n, steps)
    assert circ.passes() == 0
    assert circ.plan() == [("gate", i) for i in range(30)]


def test_replanning_after_add(oracle, capi):
    n = 12
    steps, _ = W.random_circuit(17, n, 40)
    circ = capi.Circuit(n, steps[:20])
    circ.passes()
    for m, t, c, a in steps[20:]:
        circ.add(m, t, c, a)
    assert circ.plan() == []      # the plan was dropped by add
    circ.passes()
    psi0 = np.zeros(1 << n, dtype=np.complex128)
    psi0[0] = 1
    got = run_plan(circ, steps, n, psi0, lone_apply)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    assert np.max(np.abs(got - want)) <= ATOL


def test_bulk_gate_submission_equals_one_by_one(capi):
    """dk_circuit_add_many (one call for a packed gate list) records what
    dk_circuit_add records gate by gate: the same pass plan and op lists;
    an invalid gate leaves the circuit as it was."""
    n = 14
    steps = mixed_steps(31, n, 120, multi=False)
    one = capi.Circuit(n)
    for m, t, c, a in steps:
        one.add(m, t, c, a)
    many = capi.Circuit(n)
    many.add_packed(capi.pack_steps(steps))
    assert one.passes() == many.passes()
    assert one.plan() == many.plan()
    for i in range(one.passes()):
        assert one.pass_info(i) == many.pass_info(i)
        assert np.array_equal(one.pass_ops(i), many.pass_ops(i))
    before = many.num_gates() if hasattr(many, "num_gates") else None
    bad = capi.pack_steps([steps[0], (W.gate_h(), [n], set(), set())])
    with pytest.raises(capi.DokiHipError):
        many.add_packed(bad)
    cnt = capi.C.c_size_t()
    capi.check(capi.lib().dk_circuit_num_gates(many.ptr, capi.C.byref(cnt)))
    assert cnt.value == len(steps)
    if before is not None:
        assert before == len(steps)
