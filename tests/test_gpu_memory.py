"""Device memory: arena-sized states (2 GiB and more of amplitudes, mapped
onto 2 GiB physical chunks with their scalars in a block of their own), chunk
reuse across sizes, and the budget."""
import numpy as np
import pytest

from doki_amd import workloads as W

pytestmark = pytest.mark.gpu


def check_ghz_like(capi, st, n):
    # in place: H on the top qubit, then X on qubit 0 controlled by it
    h = capi.Gate(W.gate_h())
    x = capi.Gate(W.gate_x())
    b = capi.apply(st, h, [n - 1], in_place=True)
    capi.apply(b, x, [0], controls=[n - 1], in_place=True)
    p = capi.prob_all(b)
    want = np.zeros(n)
    want[0] = want[n - 1] = 0.5
    assert np.max(np.abs(p - want)) <= 1e-12
    head, nrm = b.raw(0, 4)
    assert abs(head[0] / nrm - np.sqrt(0.5)) <= 1e-12
    tail, _ = b.raw((1 << (n - 1)) | 1, 1)
    assert abs(tail[0] / nrm - np.sqrt(0.5)) <= 1e-12
    return b


def test_arena_states_and_reuse(capi):
    for n in (27, 28, 27):
        st = capi.new(n)
        check_ghz_like(capi, st, n)
        del st
    capi.synchronize()
    stats = capi.memory_stats()
    assert stats["arena_chunks"] >= 2


def pinned_chunks(capi):
    """Chunks that live blocks hold, after everything else went back."""
    capi.synchronize()
    capi.memory_trim()
    st = capi.memory_stats()
    assert st["arena_free_chunks"] == 0 and st["cached_bytes"] == 0
    return st["arena_chunks"]


def test_arena_chunks_move_between_sizes(capi):
    """With room for 8 more chunks, a 30-qubit state (8 chunks) takes the
    chunks of released 29- and 28-qubit blocks; one more state beyond the
    cap fails cleanly and leaves the library usable."""
    # reduction scratch of this thread's lane sized for 30 qubits first
    check_ghz_like(capi, capi.new(30), 30)
    cap = pinned_chunks(capi) + 8
    capi.tune("arena_max_chunks", cap)
    try:
        x = check_ghz_like(capi, capi.new(29), 29)
        y = check_ghz_like(capi, capi.new(28), 28)
        del x, y
        z = check_ghz_like(capi, capi.new(30), 30)
        with pytest.raises(capi.DokiHipError) as ei:
            capi.new(27)
        assert ei.value.code == 2  # DK_E_ALLOC
        del z
        check_ghz_like(capi, capi.new(27), 27)
        assert capi.memory_stats()["arena_chunks"] == cap
    finally:
        capi.tune("arena_max_chunks", 0)


def test_whole_blocks_carved_and_merged(capi):
    """A released 16 GiB block serves smaller whole-chunk states from its
    front without new chunks (8, 4 and 4 GiB, distinct and intact side by
    side); released again in any order, its parts merge back and serve a 16
    GiB state once more, within the same chunks."""
    check_ghz_like(capi, capi.new(30), 30)
    cap = pinned_chunks(capi) + 8
    capi.tune("arena_max_chunks", cap)
    try:
        a = capi.new(30)
        del a
        capi.synchronize()
        assert capi.memory_stats()["cached_bytes"] == 16 << 30
        parts = [capi.new(n) for n in (29, 28, 28)]
        st = capi.memory_stats()
        assert st["arena_chunks"] == cap and st["cached_bytes"] == 0
        parts = [check_ghz_like(capi, s_, n) for s_, n in zip(parts, (29, 28, 28))]
        # each still holds its own state after the others were written
        for s_, n in zip(parts, (29, 28, 28)):
            p = capi.prob_all(s_)
            assert abs(p[0] - 0.5) <= 1e-12 and abs(p[n - 1] - 0.5) <= 1e-12
            assert np.max(np.abs(p[1:n - 1])) <= 1e-12
        for i in (1, 0, 2):
            parts[i] = None
        del s_
        capi.synchronize()
        assert capi.memory_stats()["cached_bytes"] == 16 << 30
        check_ghz_like(capi, capi.new(30), 30)
        assert capi.memory_stats()["arena_chunks"] == cap
    finally:
        capi.tune("arena_max_chunks", 0)


def test_budget_refuses_and_recovers(capi):
    import ctypes as C
    L = capi.lib()
    b0, u0 = C.c_uint64(), C.c_uint64()
    capi.check(L.dk_memory_budget(C.byref(b0), C.byref(u0)))
    capi.check(L.dk_set_memory_budget(u0.value + (1 << 24)))
    try:
        with pytest.raises(capi.DokiHipError) as ei:
            capi.new(22)  # 64 MiB over a 16 MiB allowance
        assert ei.value.code == 9  # DK_E_BUDGET
        s = capi.new(18)
        assert capi.prob_all(s).shape == (18,)
    finally:
        capi.check(L.dk_set_memory_budget(b0.value))


def test_pieces_and_chunk_recycling(capi):
    """States of 64 MiB to 1 GiB of amplitudes are buddy pieces of split
    chunks; once released, their chunks serve whole-chunk states again."""
    live = [check_ghz_like(capi, capi.new(n), n) for n in (22, 23, 24, 25, 26, 22, 26, 24)]
    for st in live:
        p = capi.prob_all(st)
        assert abs(p[0] - 0.5) <= 1e-12
    del live, st
    cap = pinned_chunks(capi) + 3
    capi.tune("arena_max_chunks", cap)
    held = []
    try:
        # 3 chunks' worth of pieces, released into the piece cache
        pieces = [capi.new(n) for n in (26, 26, 25, 25, 24, 24, 23, 23, 22, 22, 26)]
        del pieces
        while len(held) <= 3:
            try:
                held.append(capi.new(27))
            except capi.DokiHipError as e:
                assert e.code == 2
                break
        assert len(held) == 3
        check_ghz_like(capi, held[-1], 27)
    finally:
        del held
        capi.tune("arena_max_chunks", 0)


def test_scrambled_chunk_pieces_are_invisible(capi):
    """Chunks made of 16 MiB physical pieces mapped in a scrambled order
    (dk_tune "arena_sub_mib") hold states exactly as whole chunks do: the
    same gates on a whole-chunk state (27 qubits) and on buddy pieces (25
    qubits; four fill a chunk) give bit-identical amplitudes on either kind
    of chunk.  The scrambled states come first, after a trim, so that they
    take newly created chunks."""
    rng = np.random.default_rng(11)
    gates = [(capi.Gate(W.gate_u3(*rng.uniform(0, 2 * np.pi, 3))), int(t))
             for t in rng.permutation(27)]

    def run(n):
        st = capi.new(n)
        for g, t in gates:
            if t < n:
                capi.apply(st, g, [t], in_place=True)
        return st

    capi.synchronize()
    capi.memory_trim()
    try:
        capi.tune("arena_sub_mib", 16)
        scrambled = [run(n) for n in (27, 25, 25, 25, 25)]
    finally:
        capi.tune("arena_sub_mib", 0)
    plain = [run(n) for n in (27, 25)]
    for st in scrambled:
        ref = plain[0] if st.nq == 27 else plain[1]
        inner, mx = capi.compare(ref, st)
        assert mx == 0.0, (st.nq, mx)
        assert abs(inner - 1.0) <= 1e-12
    del plain, scrambled, st, ref
    capi.synchronize()
    capi.memory_trim()


def _poison_workload(capi):
    """Each path once on fresh device blocks: passes (from |0...0>, so tiles
    outside the support are skipped, and on a dense input), per-gate
    execution with a dense 5-qubit gate, marginals, a measurement of every
    qubit, a join with its marginals, a reduced density, and a batch."""
    rng = W.rng_for(5150)
    out = {}
    steps = W.random_circuit(5151, 20, 120)[0]
    st = capi.new(20)
    c = capi.Circuit(20, steps)
    c.passes()
    c.run(st, in_place=True)
    out["passes_zero"] = st.values()
    c.run(st, in_place=True)
    out["passes_dense"] = st.values()
    g = capi.Gate(W.haar_unitary(rng, 5))
    b = capi.apply(st, g, [3, 9, 14, 1, 18])
    for m, t, cc, a in steps[:40]:
        b = capi.apply(b, capi.Gate(m), t, cc, a)
    out["gates"] = b.values()
    out["probs"] = capi.prob_all(b)
    left, outs, p1 = capi.measure(b, (1 << 20) - 1, rng.random(20).tolist())
    out["outcomes"] = np.array(outs, dtype=float)
    out["p1"] = p1
    x = capi.new(11)
    for m, t, cc, a in W.random_circuit(5152, 11, 50)[0]:
        x = capi.apply(x, capi.Gate(m), t, cc, a)
    j, pj = capi.join_prob_all(b, x)
    out["join_probs"] = pj
    out["join_head"] = j.raw(0, 4096)[0]
    out["rho"] = capi.reduced_density(b, [0, 7, 19]).ravel()
    bt = capi.Batch(probs=1)
    for i in range(64):
        n = 6 + i % 9
        bt.add(n, W.random_circuit(5200 + i, n, 30)[0], density_keep=(0, n - 1))
    bt.run()
    for i in range(0, 64, 7):
        out[f"batch_probs{i}"] = bt.probs(i)
        out[f"batch_rho{i}"] = bt.density(i).ravel()
    capi.synchronize()
    return out


def test_poisoned_allocations_change_no_result(capi):
    """With every block the library hands out first filled with NaN bytes
    (states, scratch, gate matrices, plans), no result changes by a bit: no
    kernel on these paths reads memory that it or an earlier step did not
    write."""
    want = _poison_workload(capi)
    capi.memory_trim()
    capi.tune("poison_alloc", 1)
    try:
        got = _poison_workload(capi)
    finally:
        capi.tune("poison_alloc", 0)
    for k in want:
        assert np.isfinite(got[k]).all(), k
        assert np.array_equal(got[k], want[k]), k
