"""Request serving over the state-vector core on one GPU: a Python client of
the library's native server (dk_server_*, dk_submit_circuit, dk_poll in
include/doki_hip.h).

The server itself is C++ in libdoki_hip.so: it admits requests by the device
memory their states need, runs small registers together in one batched
launch per round (continuous batching: arrivals during a launch join the
next one) and large registers and joins on worker threads with execution
lanes of their own, as cache-blocked passes in place.  Session registers
stay resident on the device between requests, so later requests can apply
more gates to them, read them or join two of them.  No Python runs between
a request's submission and its result, and ctypes releases the GIL while
the client waits.

Requests are dicts as workloads.serving_requests makes them (rid, arrival,
op, regs [(nq, steps)], rolls, keep), with optional "packed" gate lists per
register (capi.pack_steps), "keep_result" (the request's register -- the
joined one for a join -- becomes a session; op "keep" only keeps it) and
registers given as ("session", id, steps) to use a live session register.

``run_plain`` is the baseline: one request at a time in arrival order, each
gate applied out of place as registry_apply does (doki.c:722-941), on the
library's default lane.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import capi

SMALL_QUBITS = 14
OPS = {"prob": 0, "measure": 1, "density": 2, "join": 3, "keep": 4}
E_BUDGET, E_PENDING = 9, 10


def state_bytes(nq, dtype=capi.C128):
    return (16 if dtype == capi.C128 else 8) << nq


def request_need(spec, dtype=capi.C128):
    """Peak device bytes of a request's new states: the register, plus the
    collapse chain of a measurement, or both registers and their join."""
    n = [r[0] for r in spec["regs"]]
    if spec["op"] == "join":
        return (state_bytes(n[0], dtype) + state_bytes(n[1], dtype) +
                state_bytes(n[0] + n[1], dtype))
    if spec["op"] == "measure":
        return 2 * state_bytes(n[0], dtype)
    return state_bytes(n[0], dtype)


class Rejected(Exception):
    pass


class _RegSpec(C.Structure):
    _fields_ = [("num_qubits", C.c_uint), ("session", C.c_uint64),
                ("num_gates", C.c_size_t), ("matrices", C.POINTER(C.c_double)),
                ("targets", C.POINTER(C.c_uint)),
                ("ctrl_masks", C.POINTER(C.c_uint64)),
                ("actrl_masks", C.POINTER(C.c_uint64))]


class _Request(C.Structure):
    _fields_ = [("op", C.c_int), ("num_regs", C.c_uint), ("regs", _RegSpec * 2),
                ("rolls", C.POINTER(C.c_double)), ("keep_mask", C.c_uint64),
                ("keep_result", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("probs", C.POINTER(C.c_double)), ("outcomes", C.POINTER(C.c_int8)),
                ("p1", C.POINTER(C.c_double)), ("rho", C.POINTER(C.c_double)),
                ("session", C.c_uint64), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double),
                ("error", C.c_char * 256)]


def _bind():
    L = capi.lib()
    if getattr(L, "_serve_bound", False):
        return L
    P, u64, PP = C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)
    for name, res, args in (
            ("dk_server_new", C.c_int, [C.c_int, C.c_uint, C.c_int, u64, PP]),
            ("dk_server_set", C.c_int, [P, C.c_char_p, C.c_int64]),
            ("dk_submit_circuit", C.c_int, [P, C.POINTER(_Request), C.POINTER(u64)]),
            ("dk_poll", C.c_int, [P, u64, C.c_int, C.POINTER(_Result)]),
            ("dk_session_free", C.c_int, [P, u64]),
            ("dk_server_stats", C.c_int, [P] + [C.POINTER(u64)] * 4),
            ("dk_server_free", None, [P])):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._serve_bound = True
    return L


class Handle:
    """One submitted request: wait() returns its result dict (probs, or
    outcomes and p1, or rho; "session" when a register was kept)."""

    def __init__(self, server, spec, ticket, sizes):
        self.server = server
        self.spec = spec
        self.ticket = ticket
        self.sizes = sizes
        self.result = None
        self.error = None
        self.t_submit = self.t_start = self.t_done = None
        self._done = False
        self._lock = threading.Lock()

    def _fetch(self, timeout):
        with self._lock:
            if self._done:
                return True
            op, n, a = self.sizes
            probs = np.empty(n) if op in ("prob", "join") else None
            outs = np.zeros(n, dtype=np.int8) if op == "measure" else None
            p1 = np.empty(n) if op == "measure" else None
            rho = np.empty(2 << (2 * a)) if op == "density" else None
            res = _Result()
            if probs is not None:
                res.probs = probs.ctypes.data_as(C.POINTER(C.c_double))
            if outs is not None:
                res.outcomes = outs.ctypes.data_as(C.POINTER(C.c_int8))
                res.p1 = p1.ctypes.data_as(C.POINTER(C.c_double))
            if rho is not None:
                res.rho = rho.ctypes.data_as(C.POINTER(C.c_double))
            ms = -1 if timeout is None else max(0, int(timeout * 1000))
            rc = _bind().dk_poll(self.server.ptr, self.ticket, ms, C.byref(res))
            if rc == E_PENDING:
                return False
            self.t_submit, self.t_start, self.t_done = res.t_submit, res.t_start, res.t_done
            if rc == E_BUDGET:
                self.error = Rejected("request %s: %s" % (self.spec.get("rid"),
                                                          res.error.decode()))
            elif rc:
                self.error = capi.DokiHipError(rc, res.error.decode())
            else:
                r = {}
                if probs is not None:
                    r["probs"] = probs
                if outs is not None:
                    r["outcomes"] = [bool(x) for x in outs]
                    r["p1"] = p1
                if rho is not None:
                    r["rho"] = rho.view(np.complex128).reshape(1 << a, 1 << a)
                if res.session:
                    r["session"] = int(res.session)
                self.result = r
            self._done = True
            return True

    def wait(self, timeout=None):
        if not self._fetch(timeout):
            raise TimeoutError("request %s still running" % self.spec.get("rid"))
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def done(self):
        return self._fetch(0)


class Server:
    """The native continuous-batching server on the library's device.
    plan: "auto"/"passes" (cache-blocked passes from 12 qubits), 0 (gate by
    gate) or a dense fusion width 3..6."""

    def __init__(self, workers=6, small_qubits=SMALL_QUBITS, plan="auto",
                 budget=None, max_batch=2048, dtype=capi.C128):
        L = _bind()
        capi.check(L.dk_init(0))
        self.dtype = dtype
        p = C.c_void_p()
        capi.check(L.dk_server_new(workers, small_qubits, dtype, budget or 0,
                                   C.byref(p)))
        self.ptr = p
        planv = -1 if plan in ("auto", "passes") else int(plan)
        capi.check(L.dk_server_set(self.ptr, b"plan", planv))
        capi.check(L.dk_server_set(self.ptr, b"max_batch", int(max_batch)))
        self._handles = []
        self._session_nq = {}

    def _reg(self, entry, packed, keep):
        if entry[0] == "session":
            sid, steps = entry[1], entry[2]
            pk = packed if packed is not None else capi.pack_steps(steps)
            nq = self._session_nq[sid]
        else:
            nq, steps = entry
            sid = 0
            pk = packed if packed is not None else capi.pack_steps(steps)
        mf, tg, cm, am, ng = capi.checked_packed(pk)
        keep.append((mf, tg, cm, am))
        r = _RegSpec()
        r.num_qubits = nq
        r.session = sid
        r.num_gates = ng
        r.matrices = mf.ctypes.data_as(C.POINTER(C.c_double))
        r.targets = tg.ctypes.data_as(C.POINTER(C.c_uint))
        r.ctrl_masks = cm.ctypes.data_as(C.POINTER(C.c_uint64))
        r.actrl_masks = am.ctypes.data_as(C.POINTER(C.c_uint64))
        return r, nq

    def submit(self, spec):
        op = spec["op"]
        req = _Request()
        req.op = OPS[op]
        regs = spec["regs"]
        req.num_regs = len(regs)
        packed = spec.get("packed") or [None] * len(regs)
        keep = []
        nqs = []
        for i, entry in enumerate(regs):
            r, nq = self._reg(entry, packed[i], keep)
            req.regs[i] = r
            nqs.append(nq)
        rolls = None
        if op == "measure":
            rolls = np.ascontiguousarray(np.asarray(spec["rolls"], dtype=np.float64))
            req.rolls = rolls.ctypes.data_as(C.POINTER(C.c_double))
        req.keep_mask = capi.mask(spec.get("keep") or ()) if op == "density" else 0
        req.keep_result = 1 if spec.get("keep_result") or op == "keep" else 0
        ticket = C.c_uint64()
        capi.check(_bind().dk_submit_circuit(self.ptr, C.byref(req), C.byref(ticket)))
        n = sum(nqs) if op == "join" else nqs[0]
        h = Handle(self, spec, ticket.value, (op, n, len(spec.get("keep") or ())))
        self._handles.append(h)
        if req.keep_result:
            # a kept register's size, for later requests naming its session
            h._kept_nq = n
        return h

    def session_of(self, handle):
        """The session id a keep_result request made (waits for it)."""
        r = handle.wait()
        sid = r["session"]
        self._session_nq[sid] = handle._kept_nq
        return sid

    def free_session(self, sid):
        capi.check(_bind().dk_session_free(self.ptr, sid))
        self._session_nq.pop(sid, None)

    def _stats(self):
        v = [C.c_uint64() for _ in range(4)]
        capi.check(_bind().dk_server_stats(self.ptr, *[C.byref(x) for x in v]))
        return [x.value for x in v]

    @property
    def batches(self):
        return self._stats()[0]

    @property
    def batched(self):
        return self._stats()[1]

    def close(self):
        if getattr(self, "ptr", None) is None:
            return
        for h in self._handles:
            h._fetch(None)
        _bind().dk_server_free(self.ptr)
        self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def serve_trace(server, specs, time_scale=1.0):
    """Submits the requests at their arrival times (scaled by time_scale) and
    waits for all of them.  Returns (handles, t0); handle.t_done is on the
    clock of time.perf_counter."""
    t0 = time.perf_counter()
    hs = []
    for s in specs:
        due = t0 + s["arrival"] * time_scale
        now = time.perf_counter()
        if due > now:
            time.sleep(due - now)
        hs.append(server.submit(s))
    for h in hs:
        h._fetch(None)
    return hs, t0


def run_operation(spec, states):
    """The request's final operation on its register(s), through the C ABI
    one call at a time (the plain path)."""
    op = spec["op"]
    st = states[0]
    if op == "prob":
        return {"probs": capi.prob_all(st)}
    if op == "measure":
        n = st.nq
        _, outs, p1 = capi.measure(st, (1 << n) - 1, spec["rolls"])
        return {"outcomes": outs, "p1": p1}
    if op == "density":
        return {"rho": capi.reduced_density(st, spec["keep"])}
    if op == "join":
        j = capi.join(states[0], states[1])
        return {"probs": capi.prob_all(j)}
    raise ValueError("unknown operation %r" % op)


def run_plain(specs, dtype=capi.C128, time_scale=1.0):
    """Plain per-request execution: requests one at a time in arrival order,
    gate by gate out of place, on the default lane.  Returns a list of
    (result or exception, latency seconds) and the makespan."""
    budget = C.c_uint64()
    capi.check(capi.lib().dk_memory_budget(C.byref(budget), None))
    t0 = time.perf_counter()
    out = []
    for s in specs:
        due = t0 + s["arrival"] * time_scale
        now = time.perf_counter()
        if due > now:
            time.sleep(due - now)
        try:
            if request_need(s, dtype) > budget.value:
                raise Rejected("request %d exceeds the budget" % s["rid"])
            states = []
            for n, steps in s["regs"]:
                st = capi.new_typed(n, dtype)
                for m, t, c, a in steps:
                    st = capi.apply(st, capi.Gate(m), t, c, a)
                states.append(st)
            res = run_operation(s, states)
            del states, st
        except Exception as e:  # noqa: BLE001 - reported per request
            res = e
        out.append((res, time.perf_counter() - due))
    return out, time.perf_counter() - t0


def latency_stats(lat):
    lat = np.sort(np.asarray(lat, dtype=np.float64))
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None, "mean_ms": None}
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3)}
