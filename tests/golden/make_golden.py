"""Writes the golden fixtures in this folder from the reference itself.

Run in a container that holds the reference tree:
    make -C oracle            # builds oracle/_ref/doki*.so from the sources
    python tests/golden/make_golden.py
The reference's own C extension (compiled from its sources, never copied) is
driven with one OpenMP thread, so its floating-point results are
deterministic.  Each fixture stores its inputs and the reference's outputs;
nothing here is read at test time except the .npz/.json files it writes.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

sys.path.insert(0, os.path.dirname(HERE))
from oracle import oracle  # noqa: E402
from doki_amd import workloads as W  # noqa: E402
import fm_exprs  # noqa: E402

ref = oracle.load_reference()
if ref is None:
    raise SystemExit("oracle/_ref is not built: run make -C oracle first")
T = 1  # OpenMP threads for every reference call


def values(reg, n, canonical=False):
    return np.array([ref.registry_get(reg, i, canonical, False)
                     for i in range(1 << n)], dtype=np.complex128)


def new_data(v):
    n = int(np.log2(len(v)))
    return ref.registry_new_data(n, np.asarray(v, dtype=np.complex128), False)


def apply_steps(reg, steps):
    for m, t, c, a in steps:
        g = ref.gate_new(len(t), np.asarray(m).tolist(), False)
        reg = ref.registry_apply(reg, g, list(t), set(c), set(a), T, False)
    return reg


def mask_of(qs):
    return sum(1 << q for q in qs)


def pack_steps(steps):
    """Circuit as arrays: matrices (k=1 only), target, control/anticontrol masks."""
    mats = np.array([s[0] for s in steps], dtype=np.complex128)
    tg = np.array([s[1][0] for s in steps], dtype=np.int64)
    cm = np.array([mask_of(s[2]) for s in steps], dtype=np.int64)
    am = np.array([mask_of(s[3]) for s in steps], dtype=np.int64)
    return mats, tg, cm, am


def res_array(lst):
    return np.array([-1 if x is None else int(x) for x in lst], dtype=np.int8)


def c0_fixture(seed, n=12, depth=200):
    steps, rolls = W.random_circuit(seed, n, depth)
    reg = apply_steps(ref.registry_new(n, False), steps)
    amps = values(reg, n)
    canon = values(reg, n, True)
    probs = np.array([ref.registry_prob(reg, q, T, False) for q in range(n)])
    st, res = ref.registry_measure(reg, (1 << n) - 1, rolls, T, False)
    assert st is None
    # half-mask measurement: every other qubit
    half = mask_of(range(0, n, 2))
    st2, res2 = ref.registry_measure(reg, half, rolls[: n // 2], T, False)
    mats, tg, cm, am = pack_steps(steps)
    np.savez_compressed(
        os.path.join(HERE, f"c0_n{n}_seed{seed}.npz"),
        n=n, mats=mats, targets=tg, cmask=cm, amask=am,
        rolls=np.array(rolls), amps=amps, canon=canon, probs=probs,
        measure_all=res_array(res), half_mask=half,
        measure_half=res_array(res2), half_state=values(st2, n - n // 2))


def multi_fixture():
    rng = W.rng_for(1234)
    cases = []
    arrays = {}
    idx = 0
    for n in (3, 5, 6):
        for k in (2, 3):
            if k > n:
                continue
            for rep in range(3):
                v = W.random_normalised_state(1000 + idx, n)
                targets = [int(x) for x in rng.choice(n, size=k, replace=False)]
                rest = [q for q in range(n) if q not in targets]
                ctrl, actrl = set(), set()
                if rest and rep >= 1:
                    ctrl = {int(rest[0])}
                if len(rest) > 1 and rep == 2:
                    actrl = {int(rest[-1])}
                u = W.haar_unitary(rng, k)
                reg = new_data(v)
                out = apply_steps(reg, [(u, targets, ctrl, actrl)])
                arrays[f"in{idx}"] = v
                arrays[f"u{idx}"] = u
                arrays[f"out{idx}"] = values(out, n)
                cases.append(dict(n=n, targets=targets, controls=sorted(ctrl),
                                  anticontrols=sorted(actrl)))
                idx += 1
    np.savez_compressed(os.path.join(HERE, "apply_multi.npz"), **arrays)
    with open(os.path.join(HERE, "apply_multi.json"), "w") as f:
        json.dump(cases, f, indent=1)


def measure_fixture():
    arrays = {}
    cases = []
    specs = [  # (n, seed, mask, nrolls)
        (4, 1, 0b1111, 4), (5, 2, 0b10110, 3), (5, 3, 0b00001, 1),
        (6, 4, 0b101010, 3), (6, 5, 0, 0), (3, 6, 0b11111000 | 0b010, 1),
        (7, 7, 0b1000001, 2), (1, 8, 1, 1),
    ]
    for idx, (n, seed, mask, nr) in enumerate(specs):
        steps, _ = W.random_circuit(seed, n, 30)
        rng = W.rng_for(500 + seed)
        rolls = rng.random(nr).tolist()
        reg = apply_steps(ref.registry_new(n, False), steps)
        st, res = ref.registry_measure(reg, mask, rolls, T, False)
        mats, tg, cm, am = pack_steps(steps)
        arrays.update({f"mats{idx}": mats, f"tg{idx}": tg, f"cm{idx}": cm,
                       f"am{idx}": am, f"res{idx}": res_array(res)})
        left = n - sum(1 for x in res if x is not None)
        if st is not None:
            arrays[f"state{idx}"] = values(st, left)
        cases.append(dict(n=n, mask=mask, rolls=rolls, has_state=st is not None,
                          left=left))
    np.savez_compressed(os.path.join(HERE, "measure_cases.npz"), **arrays)
    with open(os.path.join(HERE, "measure_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)


def misc_fixture():
    """Canonical phase quirks, unnormalised data, join, density, memory."""
    out = {}
    arrays = {}
    # canonical get: first nonzero amplitude negative real / positive real /
    # complex / located past index 0
    canon_inputs = [
        np.array([-0.6, 0.8j, 0, 0]),
        np.array([0.6, -0.8j, 0, 0]),
        np.array([0, 0, (0.6 - 0.0j) * np.exp(1j * 2.1), 0.8]),
        np.array([0, -0.8, 0.6j, 0]),
        np.array([(0.3 + 0.4j), 0, 0.5j, -np.sqrt(0.5)]),
    ]
    for i, v in enumerate(canon_inputs):
        reg = new_data(v)
        arrays[f"canon_in{i}"] = v
        arrays[f"canon_out{i}"] = values(reg, 2, True)
    # unnormalised new_data: prob, get, then an apply (which renormalises)
    v = 2.0 * W.random_normalised_state(77, 4)
    reg = new_data(v)
    arrays["unnorm_in"] = v
    arrays["unnorm_probs"] = np.array([ref.registry_prob(reg, q, T, False)
                                       for q in range(4)])
    arrays["unnorm_get"] = values(reg, 4)
    h = W.gate_h()
    reg2 = apply_steps(reg, [(h, [2], set(), set())])
    arrays["unnorm_after_h"] = values(reg2, 4)
    arrays["unnorm_after_h_probs"] = np.array(
        [ref.registry_prob(reg2, q, T, False) for q in range(4)])
    # join of states with norm constants != 1
    sa, _ = W.random_circuit(31, 3, 20)
    sb, _ = W.random_circuit(32, 2, 20)
    ra = apply_steps(ref.registry_new(3, False), sa)
    rb = apply_steps(ref.registry_new(2, False), sb)
    rj = ref.registry_join(ra, rb, T, False)
    for name, steps in (("a", sa), ("b", sb)):
        mats, tg, cm, am = pack_steps(steps)
        arrays.update({f"join_{name}_mats": mats, f"join_{name}_tg": tg,
                       f"join_{name}_cm": cm, f"join_{name}_am": am})
    arrays["join_out"] = values(rj, 5)
    # density matrices and partial-trace chains of a 5-qubit state
    sd, _ = W.random_circuit(41, 5, 40)
    rd = apply_steps(ref.registry_new(5, False), sd)
    mats, tg, cm, am = pack_steps(sd)
    arrays.update({"dens_mats": mats, "dens_tg": tg, "dens_cm": cm,
                   "dens_am": am})
    rho = ref.registry_density(rd, False)
    arrays["dens_full"] = np.array(
        [[ref.funmatrix_get(rho, i, j, False) for j in range(32)]
         for i in range(32)])
    chains = [[0], [4], [2], [1, 1], [3, 0, 1], [0, 0, 0, 0], [2, 3], [5]]
    out["dens_chains"] = chains
    out["dens_shapes"] = []
    out["dens_mem"] = []
    out["dens_trace"] = []
    for ci, ch in enumerate(chains):
        m = rho
        for e in ch:
            m = ref.funmatrix_partialtrace(m, e, False)
        r, c = ref.funmatrix_shape(m, False)
        out["dens_shapes"].append([r, c])
        out["dens_mem"].append(ref.funmatrix_mem(m, False))
        try:
            tr = ref.funmatrix_trace(m, False)
            out["dens_trace"].append([tr.real, tr.imag])
        except Exception as e:  # noqa: BLE001
            out["dens_trace"].append(str(e))
        mat = np.full((r, c), np.nan + 0j)
        for i in range(r):
            for j in range(c):
                try:
                    mat[i, j] = ref.funmatrix_get(m, i, j, False)
                except Exception:  # noqa: BLE001
                    pass
        arrays[f"dens_chain{ci}"] = mat
    out["dens_full_trace"] = [ref.funmatrix_trace(rho, False).real,
                              ref.funmatrix_trace(rho, False).imag]
    out["dens_full_mem"] = ref.funmatrix_mem(rho, False)
    out["registry_mem"] = {n: ref.registry_mem(ref.registry_new(n, False), False)
                           for n in range(1, 11)}
    np.savez_compressed(os.path.join(HERE, "misc.npz"), **arrays)
    with open(os.path.join(HERE, "misc.json"), "w") as f:
        json.dump(out, f, indent=1)


def errors_fixture():
    """Error messages of invalid calls (exception type qsimov.doki.error)."""
    g1 = ref.gate_new(1, W.gate_x().tolist(), False)
    g2 = ref.gate_new(2, np.eye(4).tolist(), False)
    r3 = ref.registry_new(3, False)
    calls = {
        "registry_new_zero": lambda: ref.registry_new(0, False),
        "registry_new_too_many": lambda: ref.registry_new(63, False),
        "gate_new_zero": lambda: ref.gate_new(0, [[1]], False),
        "gate_new_not_list": lambda: ref.gate_new(1, 5, False),
        "gate_new_wrong_size": lambda: ref.gate_new(1, [[1, 0]], False),
        "gate_new_bad_row": lambda: ref.gate_new(1, [[1, 0], [0]], False),
        "gate_new_bad_elem": lambda: ref.gate_new(1, [[1, 0], [0, "a"]], False),
        "apply_threads": lambda: ref.registry_apply(r3, g1, [0], None, None, 0, False),
        "apply_targets_not_list": lambda: ref.registry_apply(r3, g1, 0, None, None, 1, False),
        "apply_wrong_ntargets": lambda: ref.registry_apply(r3, g1, [0, 1], None, None, 1, False),
        "apply_controls_type": lambda: ref.registry_apply(r3, g1, [0], [1], None, 1, False),
        "apply_anticontrols_type": lambda: ref.registry_apply(r3, g1, [0], None, [1], 1, False),
        "apply_control_range": lambda: ref.registry_apply(r3, g1, [0], {5}, None, 1, False),
        "apply_anticontrol_range": lambda: ref.registry_apply(r3, g1, [0], set(), {5}, 1, False),
        "apply_anticontrol_none_controls": lambda: ref.registry_apply(r3, g1, [0], None, {1}, 1, False),
        "apply_control_is_anticontrol": lambda: ref.registry_apply(r3, g1, [0], {1}, {1}, 1, False),
        "apply_target_is_control": lambda: ref.registry_apply(r3, g1, [1], {1}, None, 1, False),
        "apply_target_is_anticontrol": lambda: ref.registry_apply(r3, g1, [1], set(), {1}, 1, False),
        "apply_target_range": lambda: ref.registry_apply(r3, g1, [3], None, None, 1, False),
        "apply_target_type": lambda: ref.registry_apply(r3, g1, ["a"], None, None, 1, False),
        "apply_2q_wrong_ntargets": lambda: ref.registry_apply(r3, g2, [0], None, None, 1, False),
        "measure_rolls_type": lambda: ref.registry_measure(r3, 1, (0.5,), 1, False),
        "measure_roll_range": lambda: ref.registry_measure(r3, 1, [1.0], 1, False),
        "measure_roll_negative": lambda: ref.registry_measure(r3, 1, [-0.1], 1, False),
        "measure_missing_roll": lambda: ref.registry_measure(r3, 3, [0.5], 1, False),
        "measure_threads": lambda: ref.registry_measure(r3, 1, [0.5], -2, False),
        "prob_threads": lambda: ref.registry_prob(r3, 0, 0, False),
        "join_threads": lambda: ref.registry_join(r3, r3, 0, False),
        "clone_threads": lambda: ref.registry_clone(r3, 0, False),
        "new_data_zero": lambda: ref.registry_new_data(0, [1], False),
        "new_data_list_size": lambda: ref.registry_new_data(2, [1, 0], False),
        "new_data_array_size": lambda: ref.registry_new_data(2, np.zeros(3), False),
        "new_data_type": lambda: ref.registry_new_data(1, (1, 0), False),
        "new_data_not_number": lambda: ref.registry_new_data(1, np.array(["a", "b"]), False),
        "get_not_capsule": lambda: ref.registry_get(5, 0, False, False),
        "funmatrix_get_oob": lambda: ref.funmatrix_get(ref.registry_density(r3, False), 8, 0, False),
    }
    out = {}
    for name, fn in calls.items():
        try:
            fn()
            out[name] = None
        except Exception as e:  # noqa: BLE001
            out[name] = [type(e).__module__ + "." + type(e).__name__, str(e)]
    with open(os.path.join(HERE, "errors.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def funmatrix_fixture(seeds=(0, 1, 2)):
    """The lazy matrix algebra: every chain of fm_exprs.build on seeded
    inputs (shape, mem, all elements, trace), and the edge/error cases.
    projection and apply are left out: the reference's compiled parsers for
    them write an int through a 1-byte bool ("p" format,
    src/doki/doki.c:2020-2023 and 2226-2228), which clobbers a neighbouring
    argument in this build."""
    arrays = {}
    meta = {"seeds": list(seeds), "chains": {}}
    for seed in seeds:
        inp = fm_exprs.random_inputs(seed)
        for name, fm in fm_exprs.build(ref, inp).items():
            shape, mem, vals, tr = fm_exprs.snapshot(ref, fm)
            arrays[f"s{seed}_{name}"] = vals
            meta["chains"][f"s{seed}_{name}"] = {"shape": list(shape),
                                                 "mem": mem, "trace": tr}
    meta["errors"] = fm_exprs.run_error_cases(ref)
    np.savez_compressed(os.path.join(HERE, "funmatrix.npz"), **arrays)
    with open(os.path.join(HERE, "funmatrix.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


FIXTURES = {
    "c0": lambda: ([c0_fixture(s) for s in (0, 1, 2)],
                   c0_fixture(7, n=6, depth=60)),
    "multi": multi_fixture,
    "measure": measure_fixture,
    "misc": misc_fixture,
    "errors": errors_fixture,
    "funmatrix": funmatrix_fixture,
}

if __name__ == "__main__":
    # python tests/golden/make_golden.py [fixture ...]   (default: all)
    for name in sys.argv[1:] or list(FIXTURES):
        FIXTURES[name]()
    print("golden fixtures written to", HERE)
