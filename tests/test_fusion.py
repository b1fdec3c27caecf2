"""Gate fusion planner (dk_circuit_fuse, host code of libdoki_hip.so) on the
CPU: the fused blocks, applied in order as dense gates by the oracle, must
reproduce the gate-by-gate oracle run (1e-12), for every fusion width."""
import numpy as np
import pytest

import doki_amd
from doki_amd import workloads as W
from helpers import ATOL, oracle_run


@pytest.fixture(scope="module")
def capi():
    if not doki_amd.built():
        from doki_amd import build
        build.build()
    from doki_amd import capi as A
    return A


def run_blocks(oracle, circ, nblocks, st):
    for i in range(nblocks):
        qubits, U = circ.block(i)
        st = oracle.apply(st, U, qubits)
    return st


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_random_circuit_fused_equals_unfused(oracle, capi, k):
    n = 9
    steps, _ = W.random_circuit(600 + k, n, 120)
    circ = capi.Circuit(n, steps)
    width = max(len(s[1]) + len(s[2]) + len(s[3]) for s in steps)
    if width > k:
        # a gate wider than the fusion width cannot be planned
        with pytest.raises(capi.DokiHipError):
            circ.fuse(k)
        steps = [s for s in steps if len(s[1]) + len(s[2]) + len(s[3]) <= k]
        circ = capi.Circuit(n, steps)
    nb = circ.fuse(k)
    assert 0 < nb <= len(steps)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    got = run_blocks(oracle, circ, nb, oracle.new(n)).values()
    assert np.max(np.abs(got - want)) <= ATOL
    for i in range(nb):
        qubits, U = circ.block(i)
        assert len(qubits) <= k and qubits == sorted(qubits)
        # every block is unitary
        assert np.allclose(U.conj().T @ U, np.eye(U.shape[0]), atol=1e-13)


def test_brickwork_fusion_block_count(oracle, capi):
    n = 12
    steps = W.brickwork_circuit(5, n, 4)
    circ = capi.Circuit(n, steps)
    nb = circ.fuse(5)
    # far fewer passes than gates
    assert nb * 3 < len(steps)
    want = oracle_run(oracle, oracle.new(n), steps).values()
    got = run_blocks(oracle, circ, nb, oracle.new(n)).values()
    assert np.max(np.abs(got - want)) <= ATOL


def test_fusion_respects_noncommuting_order(oracle, capi):
    # X(0); CX(0->1); H(1); X(0): the last X must not move before the CX
    n = 3
    steps = [(W.gate_x(), [0], set(), set()),
             (W.gate_x(), [1], {0}, set()),
             (W.gate_h(), [1], set(), set()),
             (W.gate_x(), [0], set(), set()),
             (W.gate_h(), [2], set(), set())]
    for k in (2, 3):
        circ = capi.Circuit(n, steps)
        nb = circ.fuse(k)
        want = oracle_run(oracle, oracle.new(n), steps).values()
        got = run_blocks(oracle, circ, nb, oracle.new(n)).values()
        assert np.max(np.abs(got - want)) <= ATOL, k
