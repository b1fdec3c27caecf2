"""The reference's own test scripts (Doki's tests/*.py), restated as GPU tests
of the drop-in `doki` module.

Each test follows one script's check: it drives only `doki` (the product
path, through libdoki_hip.so) and computes the expectation with numpy the
way the script does (kron products for one-qubit gates, controlled matrices
for multi-qubit ones, sin^2 probabilities, classical-state measurements,
kron joins, outer-product densities).  Tolerance: 1e-12 max-abs where the
scripts use np.allclose; exact equality where they compare with ==.
"""
import math

import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, kron_one_qubit, tensor_apply

pytestmark = pytest.mark.gpu


def values(doki, reg, canonical=False):
    return doki.registry_get_all(reg, canonical, False)


def apply(doki, reg, m, targets, controls=(), anticontrols=()):
    g = doki.gate_new(len(targets), np.asarray(m).tolist(), False)
    return doki.registry_apply(reg, g, list(targets), set(controls),
                               set(anticontrols), 1, False)


def test_reg_creation(doki):
    """reg_creation_tests.py:31-44: |0..0> and data round trips compare
    equal (==), from an ndarray and from a list."""
    rng = np.random.default_rng(1)
    for n in range(1, 11):
        z = values(doki, doki.registry_new(n, False))
        want = np.zeros(1 << n, dtype=complex)
        want[0] = 1
        assert np.array_equal(z, want)
        r = rng.random(1 << n)
        r = r / np.linalg.norm(r)
        assert np.array_equal(values(doki, doki.registry_new_data(n, r, False)), r)
        assert np.array_equal(values(doki, doki.registry_new_data(n, list(r), False)), r)


def test_one_gate(doki):
    """one_gate_tests.py:48-71: a U gate on each qubit equals
    I (x) U (x) I applied to the vector."""
    rng = np.random.default_rng(2)
    for n in range(1, 8):
        psi = W.random_normalised_state(100 + n, n)
        for t in range(n):
            u = W.gate_u3(*(np.pi * (rng.random(3) * 2 - 1)))
            if rng.random() < 0.5:
                u = u.conj().T                      # the scripts' inverse flag
            reg = apply(doki, doki.registry_new_data(n, psi, False), u, [t])
            want = kron_one_qubit(n, u, t) @ psi
            assert np.max(np.abs(values(doki, reg) - want)) <= ATOL, (n, t)


def test_multiple_gates(doki):
    """multiple_gate_tests.py (applyCACU): two- and three-qubit gates with
    controls and anticontrols anywhere."""
    rng = np.random.default_rng(3)
    for rep in range(40):
        n = int(rng.integers(3, 8))
        k = int(rng.integers(2, min(3, n - 1) + 1))
        qs = [int(q) for q in rng.permutation(n)]
        targets, rest = qs[:k], qs[k:]
        nc = int(rng.integers(0, len(rest) + 1))
        na = int(rng.integers(0, len(rest) - nc + 1))
        ctl, act = rest[:nc], rest[nc:nc + na]
        u = W.haar_unitary(rng, k)
        psi = W.random_normalised_state(300 + rep, n)
        reg = apply(doki, doki.registry_new_data(n, psi, False), u, targets, ctl, act)
        want = tensor_apply(psi, u, targets, ctl, act)
        assert np.max(np.abs(values(doki, reg) - want)) <= ATOL, (n, targets, ctl, act)


def test_measure_classic(doki):
    """measure_tests.py:222-244: X on a random set of qubits, then measuring
    every qubit gives those values (result list most significant qubit
    first), whatever the rolls; measuring half leaves the basis state of the
    others (compared up to phase)."""
    rng = np.random.default_rng(4)
    x = W.gate_x()
    for n in range(1, 9):
        vals = rng.choice([0, 1], size=n)
        reg = doki.registry_new(n, False)
        for q in range(n):
            if vals[q]:
                reg = apply(doki, reg, x, [q])
        left, res = doki.registry_measure(reg, (1 << n) - 1,
                                          rng.random(n).tolist(), 1, False)
        assert left is None
        assert [int(v) for v in res] == [int(v) for v in vals[::-1]]
        if n < 2:
            continue
        measured = sorted(int(q) for q in rng.choice(n, size=n // 2, replace=False))
        mask = sum(1 << q for q in measured)
        left, res = doki.registry_measure(reg, mask, rng.random(len(measured)).tolist(),
                                          1, False)
        assert len(res) == n
        for q in range(n):
            r = res[n - 1 - q]
            assert (r is None) == (q not in measured)
            if r is not None:
                assert int(r) == int(vals[q])
        kept = [q for q in range(n) if q not in measured]
        idx = sum(int(vals[q]) << i for i, q in enumerate(kept))
        got = values(doki, left)
        assert abs(abs(got[idx]) - 1) <= ATOL and np.sum(np.abs(got) ** 2) - 1 <= ATOL


def test_probability(doki):
    """probability_tests.py:19-41: RY(step * i) on qubit i, then
    P(qubit i = 1) = sin^2(step * i / 2)."""
    for nq in range(1, 10):
        step = 2 * np.pi / nq
        reg = doki.registry_new(nq, False)
        for i in range(nq):
            reg = apply(doki, reg, W.gate_ry(step * i), [i])
        for i in range(nq):
            p = doki.registry_prob(reg, i, 1, False)
            assert abs(p - math.sin(step * i / 2) ** 2) <= 1e-13, (nq, i)


def test_random_join(doki):
    """join_regs_tests.py:10-45: nq one-qubit registers joined in order
    equal the same gates applied to qubits nq-1-i of one register."""
    rng = np.random.default_rng(5)
    for nq in range(2, 9):
        gates = [W.gate_u3(*(np.pi * (rng.random(3) * 2 - 1))) for _ in range(nq)]
        regs = [apply(doki, doki.registry_new(1, False), g, [0]) for g in gates]
        ex = doki.registry_new(nq, False)
        for i, g in enumerate(gates):
            ex = apply(doki, ex, g, [nq - i - 1])
        res = regs[0]
        for r in regs[1:]:
            res = doki.registry_join(res, r, 1, False)
        assert np.max(np.abs(values(doki, res) - values(doki, ex))) <= ATOL


def test_density_and_partial_trace(doki):
    """density_matrix_tests.py:10-54: the density of a state is its outer
    product, and tracing out element 0 sums the two diagonal blocks."""
    for n in range(1, 7):
        steps, _ = W.random_circuit(600 + n, n, 30)
        reg = doki.registry_new(n, False)
        for m, t, c, a in steps:
            reg = apply(doki, reg, m, t, c, a)
        psi = values(doki, reg)
        rho = doki.registry_density(reg, False)
        got = np.array([[doki.funmatrix_get(rho, i, j, False) for j in range(1 << n)]
                        for i in range(1 << n)])
        want = np.outer(psi, psi.conj())
        assert np.max(np.abs(got - want)) <= ATOL
        if n < 2:
            continue
        pt = doki.funmatrix_partialtrace(rho, 0, False)
        assert doki.funmatrix_shape(pt, False) == (1 << (n - 1), 1 << (n - 1))
        got = np.array([[doki.funmatrix_get(pt, i, j, False) for j in range(1 << (n - 1))]
                        for i in range(1 << (n - 1))])
        want_pt = want[0::2, 0::2] + want[1::2, 1::2]
        assert np.max(np.abs(got - want_pt)) <= ATOL
        assert abs(doki.funmatrix_trace(pt, False) - 1) <= ATOL
