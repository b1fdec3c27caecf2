"""ctypes binding of the C ABI in include/doki_hip.h (libdoki_hip.so).

This is the binding a non-Python host (or a test) uses to drive the device
core directly; the `doki` extension calls the same entry points from C.
"""
import ctypes as C
import os
import re

import numpy as np

from . import LIB_PATH

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "include", "doki_hip.h")

_lib = None


class DokiHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def header_functions():
    """Names of the functions declared in include/doki_hip.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dk_\w+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libdoki_hip.so is not built: run `python -m doki_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    PP = C.POINTER(C.c_void_p)
    dp = C.POINTER(C.c_double)
    u64 = C.c_uint64
    sig = {
        "dk_init": (C.c_int, [C.c_int]),
        "dk_last_error": (C.c_char_p, []),
        "dk_synchronize": (C.c_int, []),
        "dk_memory_budget": (C.c_int, [C.POINTER(u64), C.POINTER(u64)]),
        "dk_set_memory_budget": (C.c_int, [u64]),
        "dk_state_new": (C.c_int, [C.c_uint, C.c_int, PP]),
        "dk_state_from_host": (C.c_int, [C.c_uint, dp, PP]),
        "dk_state_clone": (C.c_int, [P, PP]),
        "dk_state_free": (None, [P]),
        "dk_state_num_qubits": (C.c_uint, [P]),
        "dk_state_mem_size": (u64, [P]),
        "dk_state_device_bytes": (u64, [P]),
        "dk_state_norm": (C.c_int, [P, dp]),
        "dk_gate_new": (C.c_int, [C.c_uint, dp, PP]),
        "dk_gate_free": (None, [P]),
        "dk_apply": (C.c_int, [P, P, C.POINTER(C.c_uint), u64, u64, C.c_int, PP]),
        "dk_apply_noalloc": (C.c_int, [P, P, C.POINTER(C.c_uint), u64, u64, P]),
        "dk_prob": (C.c_int, [P, C.c_uint, dp]),
        "dk_prob_all": (C.c_int, [P, dp]),
        "dk_measure": (C.c_int, [P, u64, dp, C.c_size_t,
                                 C.POINTER(C.c_int8), dp, PP]),
        "dk_join": (C.c_int, [P, P, PP]),
        "dk_join_prob_all": (C.c_int, [P, P, PP, dp]),
        "dk_first_nonzero": (C.c_int, [P, C.POINTER(u64)]),
        "dk_read_raw": (C.c_int, [P, u64, u64, dp, dp]),
        "dk_read": (C.c_int, [P, u64, u64, dp]),
        "dk_state_compare": (C.c_int, [P, P, dp, dp]),
        "dk_reduced_density": (C.c_int, [P, u64, dp]),
        "dk_reduced_density_elements": (C.c_int, [P, u64, C.c_size_t,
                                                  C.POINTER(u64), C.POINTER(u64), dp]),
        "dk_event_record": (C.c_int, [C.c_int]),
        "dk_event_elapsed": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_float)]),
        "dk_time_ring_read": (C.c_int, [C.POINTER(C.c_float), C.c_size_t,
                                        C.POINTER(C.c_size_t)]),
        "dk_tune": (C.c_int, [C.c_char_p, C.c_int64]),
        "dk_device_info": (C.c_int, [C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
        "dk_state_new_typed": (C.c_int, [C.c_uint, C.c_int, C.c_int, PP]),
        "dk_state_dtype": (C.c_int, [P]),
        "dk_state_convert": (C.c_int, [P, C.c_int, PP]),
        "dk_circuit_new": (C.c_int, [C.c_uint, PP]),
        "dk_circuit_free": (None, [P]),
        "dk_circuit_add": (C.c_int, [P, C.c_uint, dp, C.POINTER(C.c_uint), u64, u64]),
        "dk_circuit_add_many": (C.c_int, [P, C.c_size_t, dp, C.POINTER(C.c_uint), C.POINTER(u64), C.POINTER(u64)]),
        "dk_circuit_fuse": (C.c_int, [P, C.c_uint, C.POINTER(C.c_size_t)]),
        "dk_circuit_num_gates": (C.c_int, [P, C.POINTER(C.c_size_t)]),
        "dk_circuit_block": (C.c_int, [P, C.c_size_t, C.POINTER(C.c_uint),
                                       C.POINTER(C.c_uint), dp]),
        "dk_circuit_run": (C.c_int, [P, P, C.c_int, PP]),
        "dk_circuit_passes": (C.c_int, [P, C.POINTER(C.c_size_t)]),
        "dk_circuit_pass": (C.c_int, [P, C.c_size_t, C.POINTER(C.c_uint),
                                      C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
        "dk_circuit_pass_ops": (C.c_int, [P, C.c_size_t, P, C.c_size_t,
                                          C.POINTER(C.c_size_t)]),
        "dk_circuit_plan_step": (C.c_int, [P, C.c_size_t, C.POINTER(C.c_int),
                                           C.POINTER(C.c_size_t)]),
        "dk_batch_new": (C.c_int, [C.c_int, PP]),
        "dk_batch_free": (None, [P]),
        "dk_batch_set": (C.c_int, [P, C.c_char_p, C.c_int64]),
        "dk_batch_add": (C.c_int, [P, C.c_uint, C.c_size_t, dp, C.POINTER(C.c_uint),
                                   C.POINTER(u64), C.POINTER(u64), dp, u64,
                                   C.POINTER(C.c_size_t)]),
        "dk_batch_run": (C.c_int, [P]),
        "dk_batch_size": (C.c_int, [P, C.POINTER(C.c_size_t)]),
        "dk_batch_outcomes": (C.c_int, [P, C.c_size_t, C.POINTER(C.c_int8), dp]),
        "dk_batch_density": (C.c_int, [P, C.c_size_t, dp]),
        "dk_batch_state": (C.c_int, [P, C.c_size_t, dp]),
        "dk_batch_probs": (C.c_int, [P, C.c_size_t, dp]),
        "dk_batch_clear": (C.c_int, [P]),
        "dk_memory_stats": (C.c_int, [C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
        "dk_memory_trim": (C.c_int, []),
        "dk_lane_new": (C.c_int, [PP]),
        "dk_lane_bind": (C.c_int, [P]),
        "dk_lane_free": (None, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise DokiHipError(rc, lib().dk_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class State:
    """A device state handle (freed on GC)."""

    def __init__(self, ptr):
        self.ptr = C.c_void_p(ptr)

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and _lib is not None:
            _lib.dk_state_free(self.ptr)
            self.ptr = None

    @property
    def nq(self):
        return lib().dk_state_num_qubits(self.ptr)

    def raw(self, first=0, count=None):
        n = (1 << self.nq) if count is None else count
        out = np.empty(2 * n)
        nrm = C.c_double()
        check(lib().dk_read_raw(self.ptr, first, n, _dp(out), C.byref(nrm)))
        return out.view(np.complex128), nrm.value

    def values(self):
        raw, nrm = self.raw()
        return raw / nrm

    @property
    def norm(self):
        v = C.c_double()
        check(lib().dk_state_norm(self.ptr, C.byref(v)))
        return v.value


class Gate:
    def __init__(self, matrix):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
        k = int(np.log2(m.shape[0]))
        p = C.c_void_p()
        check(lib().dk_gate_new(k, _dp(m.view(np.float64).ravel()), C.byref(p)))
        self.ptr = p
        self.k = k

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and _lib is not None:
            _lib.dk_gate_free(self.ptr)
            self.ptr = None


def new(nq):
    p = C.c_void_p()
    check(lib().dk_state_new(nq, 1, C.byref(p)))
    return State(p.value)


def from_host(values):
    v = np.ascontiguousarray(np.asarray(values, dtype=np.complex128).ravel())
    nq = int(np.log2(v.size))
    p = C.c_void_p()
    check(lib().dk_state_from_host(nq, _dp(v.view(np.float64)), C.byref(p)))
    return State(p.value)


def clone(s):
    p = C.c_void_p()
    check(lib().dk_state_clone(s.ptr, C.byref(p)))
    return State(p.value)


def mask(qs):
    m = 0
    for q in qs:
        m |= 1 << int(q)
    return m


def apply(s, gate, targets, controls=(), anticontrols=(), in_place=False):
    t = (C.c_uint * len(targets))(*targets)
    p = C.c_void_p()
    check(lib().dk_apply(s.ptr, gate.ptr, t, mask(controls), mask(anticontrols),
                         int(in_place), C.byref(p)))
    if in_place:
        return s
    return State(p.value)


def apply_noalloc(s, gate, targets, controls=(), anticontrols=(), out=None):
    """dk_apply_noalloc: the gate applied to s, written into `out` (a state
    of s's size and type; s itself when None), with no device allocation."""
    out = s if out is None else out
    t = (C.c_uint * len(targets))(*targets)
    check(lib().dk_apply_noalloc(s.ptr, gate.ptr, t, mask(controls),
                                 mask(anticontrols), out.ptr))
    return out


def prob(s, q):
    v = C.c_double()
    check(lib().dk_prob(s.ptr, q, C.byref(v)))
    return v.value


def prob_all(s):
    out = np.empty(s.nq)
    check(lib().dk_prob_all(s.ptr, _dp(out)))
    return out


def measure(s, mask_, rolls):
    n = s.nq
    r = np.ascontiguousarray(np.asarray(rolls, dtype=np.float64))
    res = (C.c_int8 * n)()
    p1 = np.zeros(n)
    p = C.c_void_p()
    check(lib().dk_measure(s.ptr, mask_, _dp(r) if r.size else None, r.size,
                           res, _dp(p1), C.byref(p)))
    st = State(p.value) if p.value else None
    return st, [None if res[i] < 0 else bool(res[i]) for i in range(n)], p1


def join(a, b):
    p = C.c_void_p()
    check(lib().dk_join(a.ptr, b.ptr, C.byref(p)))
    return State(p.value)


def join_prob_all(a, b):
    """join(a, b) and prob_all of the result, made in one pass over it."""
    p = C.c_void_p()
    out = np.empty(a.nq + b.nq)
    check(lib().dk_join_prob_all(a.ptr, b.ptr, C.byref(p), _dp(out)))
    return State(p.value), out


def compare(a, b):
    """(<a|b>, max_i |a_i - b_i|) over normalised amplitudes of two states of
    one size (either amplitude type); |<a|b>|^2 is the fidelity."""
    inner = np.zeros(2)
    mx = np.zeros(1)
    check(lib().dk_state_compare(a.ptr, b.ptr, _dp(inner), _dp(mx)))
    return complex(inner[0], inner[1]), float(mx[0])


def first_nonzero(s):
    v = C.c_uint64()
    check(lib().dk_first_nonzero(s.ptr, C.byref(v)))
    return v.value


def reduced_density(s, keep):
    a = len(keep)
    out = np.empty(2 << (2 * a))
    check(lib().dk_reduced_density(s.ptr, mask(keep), _dp(out)))
    return out.view(np.complex128).reshape(1 << a, 1 << a)


def reduced_density_elements(s, keep, rows, cols):
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    c = np.ascontiguousarray(cols, dtype=np.uint64)
    out = np.empty(2 * r.size)
    u = C.POINTER(C.c_uint64)
    check(lib().dk_reduced_density_elements(
        s.ptr, mask(keep), r.size, r.ctypes.data_as(u), c.ctypes.data_as(u),
        _dp(out)))
    return out.view(np.complex128)


C128, C64 = 0, 1


def new_typed(nq, dtype):
    p = C.c_void_p()
    check(lib().dk_state_new_typed(nq, 1, dtype, C.byref(p)))
    return State(p.value)


def convert(s, dtype):
    p = C.c_void_p()
    check(lib().dk_state_convert(s.ptr, dtype, C.byref(p)))
    return State(p.value)


def dtype_of(s):
    return lib().dk_state_dtype(s.ptr)


# One op of a cache-blocked pass (struct PassOp of doki_amd/csrc/dk_pass.h).
PASS_OP = np.dtype([
    ("kind", "<u4"), ("a", "<u4"), ("b", "<u4"), ("rmask", "<u4"),
    ("tcm", "<u8"), ("tam", "<u8"), ("m", "<f8", (8,)),
])
assert PASS_OP.itemsize == 96
(OP_REG, OP_LANE, OP_DIAG_REG, OP_DIAG_THREAD, OP_EXCH, OP_REG_MUX, OP_LANE_MUX,
 OP_MDATA) = range(8)
SEL_REG = 0x100   # mux selector: register bit (low bits), else a thread-word qubit


class Circuit:
    """A gate list with optional fusion into dense blocks (C ABI dk_circuit_*)."""

    def __init__(self, nq, steps=()):
        p = C.c_void_p()
        check(lib().dk_circuit_new(nq, C.byref(p)))
        self.ptr = p
        self.nq = nq
        for m, t, c, a in steps:
            self.add(m, t, c, a)

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and _lib is not None:
            _lib.dk_circuit_free(self.ptr)
            self.ptr = None

    def add(self, matrix, targets, controls=(), anticontrols=()):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
        t = (C.c_uint * len(targets))(*targets)
        check(lib().dk_circuit_add(self.ptr, len(targets),
                                   _dp(m.view(np.float64).ravel()), t,
                                   mask(controls), mask(anticontrols)))

    def add_packed(self, packed):
        """1-qubit gates packed as by pack_steps, in one call."""
        mf, tg, cm, am, ng = checked_packed(packed)
        u = C.POINTER(C.c_uint64)
        check(lib().dk_circuit_add_many(
            self.ptr, ng, _dp(mf), tg.ctypes.data_as(C.POINTER(C.c_uint)),
            cm.ctypes.data_as(u), am.ctypes.data_as(u)))

    def fuse(self, k):
        n = C.c_size_t()
        check(lib().dk_circuit_fuse(self.ptr, k, C.byref(n)))
        return n.value

    def passes(self):
        """Plans cache-blocked passes (dk_circuit_passes); returns their count."""
        n = C.c_size_t()
        check(lib().dk_circuit_passes(self.ptr, C.byref(n)))
        return n.value

    def plan(self):
        """The pass plan in execution order: ("pass", i) or ("gate", i)."""
        out = []
        L = lib()
        isp, idx = C.c_int(), C.c_size_t()
        i = 0
        while L.dk_circuit_plan_step(self.ptr, i, C.byref(isp), C.byref(idx)) == 0:
            out.append(("pass" if isp.value else "gate", idx.value))
            i += 1
        return out

    def pass_info(self, i):
        """(gates absorbed, ops, 12 tile qubits) of pass i."""
        ng, no = C.c_uint(), C.c_uint()
        tq = (C.c_uint * 12)()
        check(lib().dk_circuit_pass(self.ptr, i, C.byref(ng), C.byref(no), tq))
        return ng.value, no.value, list(tq)

    def pass_ops(self, i):
        """The op records of pass i as a structured numpy array (PASS_OP)."""
        cnt = C.c_size_t()
        check(lib().dk_circuit_pass_ops(self.ptr, i, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, dtype=PASS_OP)
        check(lib().dk_circuit_pass_ops(self.ptr, i, out.ctypes.data_as(C.c_void_p),
                                        cnt.value, C.byref(cnt)))
        return out

    def block(self, i):
        m = C.c_uint()
        qs = (C.c_uint * 64)()
        check(lib().dk_circuit_block(self.ptr, i, C.byref(m), qs, None))
        dim = 1 << m.value
        out = np.empty(2 * dim * dim)
        check(lib().dk_circuit_block(self.ptr, i, None, None, _dp(out)))
        return [qs[j] for j in range(m.value)], out.view(np.complex128).reshape(dim, dim)

    def run(self, s, in_place=False):
        p = C.c_void_p()
        check(lib().dk_circuit_run(self.ptr, s.ptr, int(in_place), C.byref(p)))
        return s if in_place else State(p.value)


def checked_packed(packed):
    """A packed gate list (pack_steps' layout, possibly from a client) as
    contiguous arrays of the types the C ABI reads: float64 matrices of at
    least 8 * ng values, and ng uint32 targets and uint64 masks at least.
    Anything else raises ValueError before a pointer reaches C."""
    mf, tg, cm, am, ng = packed
    ng = int(ng)
    if ng < 0:
        raise ValueError("negative gate count")
    mf = np.ascontiguousarray(mf)
    tg = np.ascontiguousarray(tg)
    cm = np.ascontiguousarray(cm)
    am = np.ascontiguousarray(am)
    for name, a, dt in (("matrices", mf, np.float64), ("targets", tg, np.uint32),
                        ("control masks", cm, np.uint64),
                        ("anticontrol masks", am, np.uint64)):
        if a.dtype != dt:
            raise ValueError(f"packed {name} must be {np.dtype(dt).name}, "
                             f"not {a.dtype.name}")
    rows = max(ng, 1)
    if mf.size < 8 * rows or tg.size < rows or cm.size < rows or am.size < rows:
        raise ValueError(f"packed arrays hold fewer than {ng} gates")
    return mf, tg, cm, am, ng


def pack_steps(steps):
    """(2x2 matrix, [target], controls, anticontrols) tuples -> the packed
    arrays dk_batch_add takes."""
    ng = len(steps)
    mats = np.zeros((max(ng, 1), 4), dtype=np.complex128)
    tg = np.zeros(max(ng, 1), dtype=np.uint32)
    cm = np.zeros(max(ng, 1), dtype=np.uint64)
    am = np.zeros(max(ng, 1), dtype=np.uint64)
    for i, (m, t, c, a) in enumerate(steps):
        if len(t) != 1:
            raise ValueError("batched registers take 1-qubit gates")
        mats[i] = np.asarray(m, dtype=np.complex128).ravel()
        tg[i] = t[0]
        cm[i] = mask(c)
        am[i] = mask(a)
    return np.ascontiguousarray(mats.view(np.float64)), tg, cm, am, ng


class Batch:
    """Many small registers run in one persistent launch (dk_batch_*)."""

    def __init__(self, dtype=C128, **settings):
        p = C.c_void_p()
        check(lib().dk_batch_new(dtype, C.byref(p)))
        self.ptr = p
        self.nqs = []
        self.keeps = []
        for k, v in settings.items():
            check(lib().dk_batch_set(self.ptr, k.encode(), int(v)))

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and _lib is not None:
            _lib.dk_batch_free(self.ptr)
            self.ptr = None

    def add(self, nq, steps, rolls=None, density_keep=()):
        """steps: (2x2 matrix, [target], controls, anticontrols) tuples."""
        return self.add_packed(nq, pack_steps(steps), rolls, mask(density_keep))

    def add_packed(self, nq, packed, rolls=None, keep_mask=0):
        """packed: (matrices as float64 [g, 8], targets uint32 [g],
        control masks uint64 [g], anticontrol masks uint64 [g], g); the
        arrays hold at least one row."""
        mf, tg, cm, am, ng = checked_packed(packed)
        r = None
        if rolls is not None:
            r = np.ascontiguousarray(np.asarray(rolls, dtype=np.float64))
            assert r.size == nq
        idx = C.c_size_t()
        u = C.POINTER(C.c_uint64)
        check(lib().dk_batch_add(
            self.ptr, nq, ng, _dp(mf), tg.ctypes.data_as(C.POINTER(C.c_uint)),
            cm.ctypes.data_as(u), am.ctypes.data_as(u),
            _dp(r) if r is not None else None, int(keep_mask), C.byref(idx)))
        self.nqs.append(nq)
        self.keeps.append(bin(int(keep_mask)).count("1"))
        return idx.value

    def clear(self):
        check(lib().dk_batch_clear(self.ptr))
        self.nqs = []
        self.keeps = []

    def run(self):
        check(lib().dk_batch_run(self.ptr))

    def probs(self, i):
        out = np.empty(self.nqs[i])
        check(lib().dk_batch_probs(self.ptr, i, _dp(out)))
        return out

    def outcomes(self, i):
        n = self.nqs[i]
        res = (C.c_int8 * n)()
        p1 = np.zeros(n)
        check(lib().dk_batch_outcomes(self.ptr, i, res, _dp(p1)))
        return [None if res[k] < 0 else bool(res[k]) for k in range(n)], p1

    def density(self, i):
        d = 1 << self.keeps[i]
        out = np.empty(2 * d * d)
        check(lib().dk_batch_density(self.ptr, i, _dp(out)))
        return out.view(np.complex128).reshape(d, d)

    def state(self, i):
        out = np.empty(2 << self.nqs[i])
        check(lib().dk_batch_state(self.ptr, i, _dp(out)))
        return out.view(np.complex128)


def synchronize():
    check(lib().dk_synchronize())


def memory_stats():
    a, f, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(lib().dk_memory_stats(C.byref(a), C.byref(f), C.byref(c)))
    return {"arena_chunks": a.value, "arena_free_chunks": f.value,
            "cached_bytes": c.value}


def memory_trim():
    check(lib().dk_memory_trim())


class Lane:
    """An execution lane (HIP stream plus scratch).  bind() makes it the
    calling thread's lane; ctypes calls release the GIL, so threads bound to
    different lanes keep the device busy concurrently."""

    def __init__(self):
        p = C.c_void_p()
        check(lib().dk_lane_new(C.byref(p)))
        self.ptr = p

    def bind(self):
        check(lib().dk_lane_bind(self.ptr))

    @staticmethod
    def unbind():
        check(lib().dk_lane_bind(None))

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and _lib is not None:
            _lib.dk_lane_free(self.ptr)
            self.ptr = None


def event_record(slot):
    check(lib().dk_event_record(slot))


def event_elapsed_ms(a, b):
    v = C.c_float()
    check(lib().dk_event_elapsed(a, b, C.byref(v)))
    return v.value


def time_ring_read():
    """Durations (ms) of the gate-kernel launches recorded in the calling
    thread's timing ring (tune("time_ring", capacity)) since the last read."""
    n = C.c_size_t()
    check(lib().dk_time_ring_read(None, 0, C.byref(n)))
    out = (C.c_float * max(1, n.value))()
    check(lib().dk_time_ring_read(out, n.value, C.byref(n)))
    return [out[i] for i in range(n.value)]


def tune(key, value):
    check(lib().dk_tune(key.encode(), int(value)))


def device_info():
    n = C.c_int()
    buf = C.create_string_buffer(64)
    check(lib().dk_device_info(C.byref(n), buf, 64))
    return n.value, buf.value.decode()
