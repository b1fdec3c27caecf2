"""Pins the oracle (oracle/doki_oracle.c) before anything is checked against it.

1. Golden vectors written by the reference itself (tests/golden/, made by
   tests/golden/make_golden.py from the reference's own C sources, run with one
   OpenMP thread): the oracle must reproduce them bit for bit.
2. What the mathematics fixes, independently of the loop structure of either
   engine: kron / tensor-contraction gate application, closed-form
   probabilities, product-state measurement, np.kron joins, einsum partial
   traces.
Runs on the CPU only.
"""
import json
import math

import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import (GOLDEN, kron_one_qubit, load, oracle_run,
                     reduced_density_einsum, tensor_apply, unpack_steps)

C0_FIXTURES = ["c0_n12_seed0.npz", "c0_n12_seed1.npz", "c0_n12_seed2.npz",
               "c0_n6_seed7.npz"]


def res_list(arr):
    return [None if x < 0 else bool(x) for x in arr]


@pytest.mark.parametrize("name", C0_FIXTURES)
def test_c0_bit_exact(oracle, name):
    z = load(name)
    n = int(z["n"])
    st = oracle_run(oracle, oracle.new(n), unpack_steps(z))
    assert np.array_equal(st.values(), z["amps"])
    probs = np.array([oracle.prob(st, q) for q in range(n)])
    assert np.array_equal(probs, z["probs"])
    canon = np.array([oracle.get(st, i, True) for i in range(1 << n)])
    assert np.array_equal(canon, z["canon"])
    rc, left, lst = oracle.measure(st, (1 << n) - 1, z["rolls"])
    assert rc == 0 and left is None
    assert lst == res_list(z["measure_all"])
    rc, left, lst = oracle.measure(st, int(z["half_mask"]), z["rolls"][: n // 2])
    assert rc == 0 and lst == res_list(z["measure_half"])
    assert np.array_equal(left.values(), z["half_state"])


def test_c0_generator_reproduces_fixture():
    z = load("c0_n12_seed0.npz")
    steps, rolls = W.random_circuit(0, 12, 200)
    assert np.array_equal(np.array([s[0] for s in steps]), z["mats"])
    assert [s[1][0] for s in steps] == z["targets"].tolist()
    assert np.array_equal(np.array(rolls), z["rolls"])


def test_multi_target_bit_exact(oracle):
    z = load("apply_multi.npz")
    cases = json.load(open(f"{GOLDEN}/apply_multi.json"))
    for i, c in enumerate(cases):
        st = oracle.from_data(c["n"], z[f"in{i}"])
        out = oracle.apply(st, z[f"u{i}"], c["targets"], c["controls"],
                           c["anticontrols"])
        assert np.array_equal(out.values(), z[f"out{i}"]), c


def test_measure_cases_bit_exact(oracle):
    z = load("measure_cases.npz")
    cases = json.load(open(f"{GOLDEN}/measure_cases.json"))
    for i, c in enumerate(cases):
        steps = [(z[f"mats{i}"][j], [int(z[f"tg{i}"][j])],
                  {q for q in range(c["n"]) if (int(z[f"cm{i}"][j]) >> q) & 1},
                  {q for q in range(c["n"]) if (int(z[f"am{i}"][j]) >> q) & 1})
                 for j in range(len(z[f"mats{i}"]))]
        st = oracle_run(oracle, oracle.new(c["n"]), steps)
        rc, left, lst = oracle.measure(st, c["mask"], c["rolls"])
        assert rc == 0
        assert lst == res_list(z[f"res{i}"]), i
        assert (left is not None) == c["has_state"]
        if left is not None:
            assert np.array_equal(left.values(), z[f"state{i}"])


def test_misc_bit_exact(oracle):
    z = load("misc.npz")
    meta = json.load(open(f"{GOLDEN}/misc.json"))
    for i in range(5):
        st = oracle.from_data(2, z[f"canon_in{i}"])
        got = np.array([oracle.get(st, j, True) for j in range(4)])
        assert np.array_equal(got, z[f"canon_out{i}"]), i
    st = oracle.from_data(4, z["unnorm_in"])
    assert np.array_equal(np.array([oracle.prob(st, q) for q in range(4)]),
                          z["unnorm_probs"])
    assert np.array_equal(st.values(), z["unnorm_get"])
    st2 = oracle.apply(st, W.gate_h(), [2])
    assert np.array_equal(st2.values(), z["unnorm_after_h"])
    assert np.array_equal(np.array([oracle.prob(st2, q) for q in range(4)]),
                          z["unnorm_after_h_probs"])
    a = oracle_run(oracle, oracle.new(3), unpack_steps(z, "join_a_"))
    b = oracle_run(oracle, oracle.new(2), unpack_steps(z, "join_b_"))
    assert np.array_equal(oracle.join(a, b).values(), z["join_out"])
    d = oracle_run(oracle, oracle.new(5), unpack_steps(z, "dens_"))
    assert np.array_equal(oracle.density_matrix(d), z["dens_full"])
    for ci, chain in enumerate(meta["dens_chains"]):
        ref = z[f"dens_chain{ci}"]
        dim = ref.shape[0]
        assert [dim, dim] == meta["dens_shapes"][ci]
        got = np.array([[oracle.density_get(d, chain, i, j)[1]
                         for j in range(dim)] for i in range(dim)])
        assert np.array_equal(got, ref), chain
        tr = meta["dens_trace"][ci]
        rc, t = oracle.density_trace(d, chain)
        assert rc == 0 and [t.real, t.imag] == tr
    for n, b in meta["registry_mem"].items():
        assert oracle.mem_size(int(n)) == b


# ---- independent mathematics ------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 5, 7])
def test_one_qubit_matches_kron(oracle, n):
    rng = W.rng_for(10 + n)
    st = oracle.new(n)
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    for t in range(n):
        a = rng.random(3) * 2 * math.pi
        g = W.gate_u3(*a)
        st = oracle.apply(st, g, [t])
        psi = kron_one_qubit(n, g, t) @ psi
        assert np.allclose(st.values(), psi, rtol=0, atol=1e-13)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_controlled_multi_target_matches_tensor(oracle, k):
    rng = W.rng_for(200 + k)
    n = 7
    for rep in range(6):
        psi = W.random_normalised_state(300 + 10 * k + rep, n)
        targets = [int(x) for x in rng.choice(n, size=k, replace=False)]
        rest = [q for q in range(n) if q not in targets]
        rng.shuffle(rest)
        ctrl = rest[: rep % 3]
        actrl = rest[3: 3 + rep % 2]
        u = W.haar_unitary(rng, k)
        got = oracle.apply(oracle.from_data(n, psi), u, targets, ctrl, actrl)
        want = tensor_apply(psi, u, targets, ctrl, actrl)
        assert np.allclose(got.values(), want / np.linalg.norm(want),
                           rtol=0, atol=1e-13)
        assert abs(got.norm - 1.0) < 1e-13


def test_probability_closed_form(oracle):
    # tests/probability_tests.py:19-41: RY(step*i) on qubit i -> sin^2(step*i/2)
    n = 6
    step = 2 * math.pi / n
    st = oracle.new(n)
    for i in range(n):
        st = oracle.apply(st, W.gate_ry(step * i), [i])
    for i in range(n):
        assert abs(oracle.prob(st, i) - math.sin(step * i / 2) ** 2) < 1e-13


def test_measure_classic_states(oracle):
    # tests/measure_tests.py:222-244: X on a random subset; outcomes certain,
    # result list runs from qubit n-1 down to qubit 0.
    rng = W.rng_for(5)
    for n in range(1, 7):
        vals = rng.integers(0, 2, size=n)
        st = oracle.new(n)
        for q in range(n):
            if vals[q]:
                st = oracle.apply(st, W.gate_x(), [q])
        rc, left, lst = oracle.measure(st, (1 << n) - 1, rng.random(n))
        assert rc == 0 and left is None
        assert lst == [bool(v) for v in vals[::-1]]


def test_measure_product_state_collapse(oracle):
    # measuring some qubits of a product state leaves the product of the
    # others (up to phase), as tests/measure_tests.py:152-199 checks
    n = 5
    rng = W.rng_for(9)
    gates = [W.gate_u3(*(rng.random(3) * 6)) for _ in range(n)]
    st = oracle.new(n)
    for q in range(n):
        st = oracle.apply(st, gates[q], [q])
    measured = [1, 3]
    rc, left, lst = oracle.measure(st, (1 << 1) | (1 << 3), [0.3, 0.6])
    kept = [q for q in range(n) if q not in measured]
    psi = np.ones(1, dtype=complex)
    for q in reversed(kept):
        psi = np.kron(psi, gates[q] @ np.array([1, 0]))
    got = left.values()
    ph = np.exp(-1j * (np.angle(got[0]) - np.angle(psi[0])))
    assert np.allclose(got * ph, psi, rtol=0, atol=1e-13)
    # P1 per measured qubit of a product state = |gate[1,0]|^2
    p3 = abs(gates[3][1, 0]) ** 2
    assert lst[n - 1 - 3] == (p3 > 0.3)


def test_measure_roll_order_and_rule(oracle):
    # P1 > roll (strict), rolls consumed from the most significant measured
    # qubit down: qubit 2 certain 1, qubit 0 certain 0.
    st = oracle.new(3)
    st = oracle.apply(st, W.gate_x(), [2])
    rc, left, lst = oracle.measure(st, 0b101, [0.999, 0.0])
    assert lst == [True, None, False]
    # a roll equal to P1 = 1/2 gives 0 (strict comparison), P0 = 1 - P1
    h = oracle.apply(oracle.new(1), W.gate_h(), [0])
    p1 = oracle.prob(h, 0)
    rc, left, lst = oracle.measure(h, 1, [p1])
    assert lst == [False]
    rc, left, lst = oracle.measure(h, 1, [np.nextafter(p1, 0)])
    assert lst == [True]


def test_measure_nan_roll(oracle):
    """A NaN roll passes the reference's range check (roll < 0 || roll >= 1,
    doki.c:1114) and P1 > NaN is false: outcome 0, the kept half renormalised
    by sqrt(1 - P1).  Checked against the compiled reference where it is
    built."""
    n = 4
    steps, _ = W.random_circuit(515, n, 30)
    st = oracle_run(oracle, oracle.new(n), steps)
    p3 = oracle.prob(st, 3)
    rc, one, lst1 = oracle.measure(st, 0b1000, [float("nan")])
    assert rc == 0 and lst1[0] is False
    assert abs(one.norm - math.sqrt(1.0 - p3)) <= 1e-15
    rc, left, lst = oracle.measure(st, 0b1001, [float("nan"), 0.25])
    assert rc == 0 and lst[0] is False
    ref = oracle.load_reference()
    if ref is None:
        pytest.skip("compiled reference not built")
    reg = ref.registry_new(n, False)
    for m, t, c, a in steps:
        g = ref.gate_new(1, np.asarray(m).tolist(), False)
        reg = ref.registry_apply(reg, g, list(t), set(c), set(a), 1, False)
    rleft, rlst = ref.registry_measure(reg, 0b1001, [float("nan"), 0.25], 1, False)
    assert rlst == lst
    got = np.array([ref.registry_get(rleft, i, False, False) for i in range(1 << (n - 2))])
    assert np.array_equal(got, left.values())


def test_join_is_kron(oracle):
    a = W.random_normalised_state(1, 3)
    b = W.random_normalised_state(2, 2)
    got = oracle.join(oracle.from_data(3, a), oracle.from_data(2, b)).values()
    assert np.allclose(got, np.kron(a, b), rtol=0, atol=1e-15)


def test_density_and_partial_traces_match_einsum(oracle):
    n = 5
    psi = W.random_normalised_state(21, n)
    st = oracle.from_data(n, psi)
    full = oracle.density_matrix(st)
    assert np.allclose(full, np.outer(psi, psi.conj()), rtol=0, atol=1e-15)
    # chain [3, 0, 1] traces qubit 3, then qubit 0, then (remaining 1,2,4 ->
    # position 1) qubit 2: kept {1, 4}
    got = oracle.density_matrix(st, [3, 0, 1])
    want = reduced_density_einsum(psi, [1, 4])
    assert np.allclose(got, want, rtol=0, atol=1e-14)


def test_canonical_phase_rules(oracle):
    # phase of the first nonzero amplitude only when its imaginary part is
    # nonzero (negative reals keep their sign)
    st = oracle.from_data(1, [-0.6, 0.8])
    assert oracle.phase(st) == 0.0
    assert oracle.get(st, 0, True) == complex(-0.6, 0)
    st = oracle.from_data(1, [0, (1 + 1j) / math.sqrt(2)])
    assert abs(oracle.phase(st) - math.pi / 4) < 1e-15
    assert abs(oracle.get(st, 1, True) - 1) < 1e-15
