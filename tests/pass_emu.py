"""Host-side check of cache-blocked pass plans (test infrastructure).

Executes a pass plan from dk_circuit_passes with numpy, following the op
semantics the pass kernel implements: a tile is the 2^12 amplitudes whose
index bits 0..5 and the pass's six high qubits vary; register, wave and lane
bits name tile slots under the mapping current at each op; exchanges change
the mapping only.  This checks the planner (gate order, locations, control
splitting, merged products) on the CPU; the kernel's register and LDS
mechanics are checked on the GPU against the oracle.
"""
import numpy as np

LANE_BITS, HIGH_BITS, REG_BITS, WAVE_BITS = 6, 6, 4, 2
(OP_REG, OP_LANE, OP_DIAG_REG, OP_DIAG_THREAD, OP_EXCH, OP_REG_MUX, OP_LANE_MUX,
 OP_MDATA) = range(8)
SEL_REG = 0x100


def _tile_index(n, tile_qubits):
    """idx[t, s]: state index of slot s of tile t (slot bit b = tile_qubits[b])."""
    others = [q for q in range(n) if q not in tile_qubits]
    t = np.arange(1 << len(others), dtype=np.int64)
    base = np.zeros_like(t)
    for j, q in enumerate(others):
        base |= ((t >> j) & 1) << q
    s = np.arange(1 << len(tile_qubits), dtype=np.int64)
    off = np.zeros_like(s)
    for b, q in enumerate(tile_qubits):
        off |= ((s >> b) & 1) << q
    return base[:, None] | off[None, :]


def run_pass(psi, n, tile_qubits, ops):
    """One pass on the normalised state psi (complex128); returns the new
    (unnormalised) state.  Asserts what the kernel relies on: exchanges are
    permutations of the high slots under which a register bit keeps its
    slot, and no thread-word field (tcm, tam, a diagonal op's qubit) names a
    qubit held in registers."""
    idx = _tile_index(n, tile_qubits)
    A = psi[idx].copy()
    hpos = tile_qubits[LANE_BITS:]
    rs = list(range(REG_BITS))

    def reg_qubit(j):
        return hpos[rs[j]]

    def bit(q):
        return ((idx >> q) & 1).astype(bool)

    ops = list(ops)
    for i, op in enumerate(ops):
        kind = int(op["kind"])
        if kind == OP_MDATA:
            # read by the mux op before it
            assert i > 0 and int(ops[i - 1]["kind"]) in (OP_REG_MUX, OP_LANE_MUX)
            continue
        if kind == OP_EXCH:
            new_rs = [(int(op["a"]) >> (3 * j)) & 7 for j in range(REG_BITS)]
            ws = [(int(op["b"]) >> (3 * b)) & 7 for b in range(WAVE_BITS)]
            assert sorted(new_rs + ws) == list(range(HIGH_BITS))
            # the kernel splits an exchange in two on a register bit that
            # keeps its slot; every slot kept in registers keeps its bit
            assert any(new_rs[j] == rs[j] for j in range(REG_BITS))
            assert all(new_rs[j] == rs[j] for j in range(REG_BITS)
                       if rs[j] in new_rs)
            rs = new_rs
            continue
        m = op["m"].view(np.complex128).reshape(2, 2)
        held = sum(1 << reg_qubit(j) for j in range(REG_BITS))
        tcm, tam, rmask = int(op["tcm"]), int(op["tam"]), int(op["rmask"])
        assert not (tcm & held) and not (tam & held)
        ok = np.ones(A.shape, dtype=bool)
        for q in range(n):
            if (tcm >> q) & 1:
                ok &= bit(q)
            if (tam >> q) & 1:
                ok &= ~bit(q)
        for j in range(REG_BITS):
            if (rmask >> j) & 1:
                ok &= bit(reg_qubit(j))
            if (rmask >> (8 + j)) & 1:
                ok &= ~bit(reg_qubit(j))
        a = int(op["a"])
        if kind in (OP_REG, OP_LANE, OP_REG_MUX, OP_LANE_MUX):
            if kind in (OP_REG, OP_REG_MUX):
                q, sb = reg_qubit(a), LANE_BITS + rs[a]
            else:
                assert a < LANE_BITS
                q, sb = a, a
            slots = np.arange(A.shape[1])
            partner = slots ^ (1 << sb)
            upper = bit(q)
            x = np.where(upper, A[:, partner], A)   # lower amplitude
            y = np.where(upper, A, A[:, partner])   # upper amplitude
            if kind in (OP_REG_MUX, OP_LANE_MUX):
                # m where the selector qubit is 0, the next record's where 1
                assert not (tcm | tam | rmask)
                m1 = ops[i + 1]["m"].view(np.complex128).reshape(2, 2)
                assert int(ops[i + 1]["kind"]) == OP_MDATA
                b = int(op["b"])
                if b & SEL_REG:
                    assert kind == OP_REG_MUX and (b & 0xff) < REG_BITS and (b & 0xff) != a
                    s = reg_qubit(b & 0xff)
                else:
                    assert LANE_BITS <= b < n and not ((held >> b) & 1)
                    s = b
                sel = bit(s)
                o0 = np.where(sel, m1[0, 0] * x + m1[0, 1] * y, m[0, 0] * x + m[0, 1] * y)
                o1 = np.where(sel, m1[1, 0] * x + m1[1, 1] * y, m[1, 0] * x + m[1, 1] * y)
                A = np.where(upper, o1, o0)
                continue
            o = np.where(upper, m[1, 0] * x + m[1, 1] * y, m[0, 0] * x + m[0, 1] * y)
            A = np.where(ok, o, A)
        else:
            if kind == OP_DIAG_REG:
                q = reg_qubit(a)
            else:
                assert kind == OP_DIAG_THREAD and not ((held >> a) & 1)
                q = a
            f = np.where(bit(q), m[1, 1], m[0, 0])
            A = np.where(ok, A * f, A)
    out = np.empty_like(psi)
    out[idx] = A
    return out


def run_plan(circ, steps, n, psi, apply_gate):
    """Runs circ's pass plan (circ.passes() already called) on psi:
    passes through run_pass, lone gates through apply_gate(psi, step).
    The state is renormalised after every step, as the kernels do."""
    for kind, i in circ.plan():
        if kind == "pass":
            _, _, tq = circ.pass_info(i)
            psi = run_pass(psi, n, tq, circ.pass_ops(i))
        else:
            psi = apply_gate(psi, steps[i])
        psi = psi / np.linalg.norm(psi)
    return psi
