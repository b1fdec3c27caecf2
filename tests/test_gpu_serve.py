"""The serving engine and the execution lanes under it.

Each request's result must be the same whether it runs alone or in the
concurrent mix (bit for bit: batched registers run in one workgroup in a
fixed order, lane work launches the same kernels with the same grids), and
must agree with the oracle running the same gates: probabilities and
densities within 1e-12, outcomes identical wherever no roll lies within 1e-9
of the P1 it was decided on."""
import threading

import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, oracle_run, reduced_density_einsum

pytestmark = pytest.mark.gpu


def oracle_result(O, spec):
    states = [oracle_run(O, O.new(n), steps) for n, steps in spec["regs"]]
    op = spec["op"]
    st = states[0]
    if op == "join":
        st = O.join(states[0], states[1])
    if op in ("prob", "join"):
        return {"probs": np.array([O.prob(st, q) for q in range(st.nq)])}
    if op == "measure":
        n = st.nq
        rc, left, outs = O.measure(st, (1 << n) - 1, spec["rolls"])
        return {"outcomes": outs}
    return {"rho": reduced_density_einsum(st.values(), spec["keep"])}


def assert_same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        if k == "outcomes":
            assert a[k] == b[k]
        else:
            assert np.array_equal(a[k], b[k]), k


def assert_close(got, want, spec):
    if "probs" in want:
        assert np.max(np.abs(got["probs"] - want["probs"])) <= ATOL
    elif "rho" in want:
        assert np.max(np.abs(got["rho"] - want["rho"])) <= ATOL
    else:
        near = any(abs(p - r) < 1e-9 for p, r in zip(got["p1"], spec["rolls"]))
        if not near:
            assert got["outcomes"] == want["outcomes"]


def small_mix(seed, count):
    """The served distribution with every band capped at 18 qubits, shorter
    circuits, and joins of at most 20 qubits (so that the oracle finishes in
    seconds)."""
    specs = W.serving_requests(seed, 3 * count, qps=1000.0, nmax=18, depth=(20, 80))
    fit = [s for s in specs
           if s["op"] != "join" or s["regs"][0][0] + s["regs"][1][0] <= 20]
    return fit[:count]


def test_lanes_concurrent_match_sequential(capi):
    specs = [(18, W.random_circuit(900 + i, 18, 60)[0]) for i in range(4)]

    def run(n, steps):
        st = capi.new(n)
        c = capi.Circuit(n, steps)
        c.fuse(5)
        c.run(st, in_place=True)
        return capi.prob_all(st), st.values()

    want = [run(n, s) for n, s in specs]
    got = [None] * len(specs)

    def worker(i):
        lane = capi.Lane()
        lane.bind()
        got[i] = run(*specs[i])
        capi.Lane.unbind()

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(len(specs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for (pw, vw), (pg, vg) in zip(want, got):
        assert np.array_equal(pw, pg)
        assert np.array_equal(vw, vg)


def test_state_handed_between_lanes(capi, oracle):
    steps = W.random_circuit(77, 16, 80)[0]
    box = {}

    def make():
        lane = capi.Lane()
        lane.bind()
        st = capi.new(16)
        c = capi.Circuit(16, steps)
        c.run(st, in_place=True)
        box["st"] = st  # still queued on this lane when handed over
        capi.Lane.unbind()

    t = threading.Thread(target=make)
    t.start()
    t.join()
    want = oracle_run(oracle, oracle.new(16), steps)
    got = box.pop("st")
    assert np.max(np.abs(got.values() - want.values())) <= ATOL
    assert np.max(np.abs(capi.prob_all(got) -
                         [oracle.prob(want, q) for q in range(16)])) <= ATOL


def test_serve_alone_equals_mix_and_oracle(capi, oracle):
    from doki_amd import serve
    specs = small_mix(31, 48)
    assert {s["op"] for s in specs} == set(W.SERVE_OPS)
    with serve.Server(workers=3, small_qubits=14) as srv:
        alone = {s["rid"]: srv.submit(s).wait(120) for s in specs}
    with serve.Server(workers=3, small_qubits=14) as srv:
        hs, _ = serve.serve_trace(srv, specs, time_scale=0.0)
        mix = {h.spec["rid"]: h.wait(0) for h in hs}
        assert srv.batches >= 1 and srv.batched >= 1
    for s in specs:
        assert_same(alone[s["rid"]], mix[s["rid"]])
        assert_close(mix[s["rid"]], oracle_result(oracle, s), s)


def test_serve_plain_agrees(capi):
    from doki_amd import serve
    specs = small_mix(32, 16)
    plain, _ = serve.run_plain(specs, time_scale=0.0)
    with serve.Server(workers=2) as srv:
        hs, _ = serve.serve_trace(srv, specs, time_scale=0.0)
    for s, (res, _lat), h in zip(specs, plain, hs):
        assert not isinstance(res, Exception)
        assert_close(h.wait(0), res, s)


def test_serve_rejects_over_budget(capi):
    from doki_amd import serve
    steps = W.random_circuit(5, 12, 20)[0]
    spec = dict(rid=0, arrival=0.0, op="join", regs=[(12, steps), (12, steps)],
                rolls=None, keep=())
    with serve.Server(workers=1, budget=1 << 20) as srv:
        h = srv.submit(spec)
        with pytest.raises(serve.Rejected):
            h.wait(10)
        # a request that fits still runs
        ok = dict(rid=1, arrival=0.0, op="prob", regs=[(12, steps)], rolls=None,
                  keep=())
        assert srv.submit(ok).wait(60)["probs"].shape == (12,)


def _product_expect(spec):
    """Closed forms for a register run by uncontrolled one-qubit gates only
    (a product state): P(q = 1) = |U_q[1, 0]|^2 with U_q the product of the
    gates on q; measuring one qubit leaves the others' marginals unchanged,
    so outcome(q) = P1_q > its roll; the reduced density is the kron of the
    kept qubits' 2x2 densities, highest kept qubit leftmost."""
    n, steps = spec["regs"][0]
    U = [np.eye(2, dtype=np.complex128) for _ in range(n)]
    for m, t, c, a in steps:
        assert not c and not a
        U[t[0]] = np.asarray(m) @ U[t[0]]
    col = [u[:, 0] for u in U]
    p1 = np.array([abs(c[1]) ** 2 for c in col])
    if spec["op"] == "prob":
        return {"probs": p1}
    if spec["op"] == "measure":
        # list slot i is qubit n-1-i and takes roll i
        return {"outcomes": [bool(p1[n - 1 - i] > r) for i, r in enumerate(spec["rolls"])],
                "p1": np.array([p1[n - 1 - i] for i in range(n)])}
    rho = np.ones((1, 1), dtype=np.complex128)
    for q in spec["keep"]:
        rho = np.kron(np.outer(col[q], col[q].conj()), rho)
    return {"rho": rho}


def test_serve_large_band_and_big_joins(capi):
    """The served mix's 27-30-qubit band (worker lanes, cache-blocked passes)
    and joins to 31 and 32 qubits, which no oracle run reaches in seconds:
    product-state requests against closed forms, mixed-circuit requests
    against plain per-gate execution, join marginals against the marginals
    of their two operands (a kron join's marginals are its operands'), and
    every result the same alone and in the concurrent mix."""
    from doki_amd import serve
    rng = W.rng_for(2730)
    specs, kinds = [], {}

    def add(op, regs, kind):
        n0 = regs[0][0]
        rolls = rng.random(n0).tolist() if op == "measure" else None
        keep = tuple(sorted(int(q) for q in rng.choice(n0, size=3, replace=False))) \
            if op == "density" else ()
        specs.append(dict(rid=len(specs), arrival=0.0, op=op, regs=regs,
                          rolls=rolls, keep=keep))
        kinds[len(specs) - 1] = kind

    for n, op in ((27, "measure"), (28, "prob"), (29, "density"), (30, "prob")):
        add(op, [(n, [W.random_gate_step(rng, n, 0.0, 0.0) for _ in range(60)])],
            "product")
    for n, op in ((27, "prob"), (30, "density"), (28, "measure")):
        add(op, [(n, [W.random_gate_step(rng, n) for _ in range(40)])], "mixed")
    for na, nb in ((16, 15), (16, 16)):
        add("join", [(na, [W.random_gate_step(rng, na) for _ in range(50)]),
                     (nb, [W.random_gate_step(rng, nb) for _ in range(50)])], "join")

    with serve.Server(workers=2) as srv:
        alone = {}
        for s in specs:
            alone[s["rid"]] = srv.submit(s).wait(300)
            print(f"[serve-large] request {s['rid']} ({s['op']}, "
                  f"{[r[0] for r in s['regs']]} qubits) alone done", flush=True)
    with serve.Server(workers=3) as srv:
        hs, _ = serve.serve_trace(srv, specs, time_scale=0.0)
        mix = {h.spec["rid"]: h.wait(0) for h in hs}
    for s in specs:
        assert_same(alone[s["rid"]], mix[s["rid"]])
    print("[serve-large] alone == mix", flush=True)

    mixed = [s for s in specs if kinds[s["rid"]] == "mixed"]
    plain, _ = serve.run_plain(mixed, time_scale=0.0)
    for s, (res, _lat) in zip(mixed, plain):
        assert not isinstance(res, Exception), res
        assert_close(mix[s["rid"]], res, s)
    for s in specs:
        got = mix[s["rid"]]
        if kinds[s["rid"]] == "product":
            want = _product_expect(s)
            if "outcomes" in want:
                assert np.max(np.abs(got["p1"] - want["p1"])) <= ATOL
            assert_close(got, want, s)
        elif kinds[s["rid"]] == "join":
            (na, sa), (nb, sb) = s["regs"]
            pa = capi.prob_all(serve_plain_state(capi, na, sa))
            pb = capi.prob_all(serve_plain_state(capi, nb, sb))
            assert got["probs"].shape == (na + nb,)
            assert np.max(np.abs(got["probs"] - np.concatenate([pb, pa]))) <= ATOL


def serve_plain_state(capi, n, steps):
    st = capi.new(n)
    for m, t, c, a in steps:
        st = capi.apply(st, capi.Gate(m), t, c, a)
    return st
