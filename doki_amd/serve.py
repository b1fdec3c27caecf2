"""Request serving over the state-vector core on one GPU.

Requests (dicts as made by workloads.serving_requests: op, regs, rolls, keep)
are admitted by the device memory their states need, against a byte budget:

* registers of at most ``small_qubits`` qubits whose operation the batched
  engine covers (marginal probabilities, full measurement, reduced density of
  at most 6 qubits) go to a batcher thread.  It runs everything queued since
  its previous launch as one persistent dk_batch launch, so arrivals during a
  launch join the next one (continuous batching);
* larger registers and joins go to worker threads, each bound to its own
  execution lane (a HIP stream with its own scratch), so that several
  requests run on the device at once.  Their circuits run in place as
  cache-blocked passes (or fused dense blocks, as the tuned plan says), and
  their states stay resident on the device from creation to the final
  operation.  A join and the marginals read from it are made in one pass
  (the joined state is written once and not read back).

ctypes calls release the GIL, so host threads overlap while the device works.
Every request's result depends only on the request: the batched engine runs
each register in one workgroup in a fixed order, and the lane path launches
the same kernels with the same grids whatever else is running.

``run_plain`` is the baseline: one request at a time in arrival order, each
gate applied out of place as registry_apply does (doki.c:722-941), on the
library's default lane.
"""
import collections
import threading
import time

import numpy as np

from . import capi
from . import tune

SMALL_QUBITS = 14


def state_bytes(nq, dtype=capi.C128):
    return (16 if dtype == capi.C128 else 8) << nq


def request_need(spec, dtype=capi.C128):
    """Peak device bytes of a request's states: the register, plus the
    collapse chain of a measurement (at most one more register's worth), or
    both registers and their join."""
    n0 = spec["regs"][0][0]
    if spec["op"] == "join":
        n1 = spec["regs"][1][0]
        return (state_bytes(n0, dtype) + state_bytes(n1, dtype) +
                state_bytes(n0 + n1, dtype))
    if spec["op"] == "measure":
        return 2 * state_bytes(n0, dtype)
    return state_bytes(n0, dtype)


class Rejected(Exception):
    pass


class Handle:
    """One submitted request: wait() returns its result dict."""

    def __init__(self, spec, dtype):
        self.spec = spec
        self.need = request_need(spec, dtype)
        # a client may send its gates packed already (capi.pack_steps)
        self.packed = spec.get("packed") or [capi.pack_steps(steps)
                                             for _, steps in spec["regs"]]
        self.result = None
        self.error = None
        self.t_submit = None
        self.t_start = None   # execution began (batch launch or worker)
        self.t_done = None
        self._ev = threading.Event()

    def wait(self, timeout=None):
        if not self._ev.wait(timeout):
            raise TimeoutError("request %d still running" % self.spec["rid"])
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def done(self):
        return self._ev.is_set()


def _finish(h, result=None, error=None):
    h.result = result
    h.error = error
    h.t_done = time.perf_counter()
    h._ev.set()


def build_register(nq, packed, dtype, plan="auto"):
    """|0...0> of nq qubits with the packed circuit applied in place under
    `plan`: "passes" (cache-blocked passes), a fusion width k (dense blocks
    of up to k qubits, when the register is wide enough for them), 0 (gate
    by gate), or "auto" for the tuned plan of (nq, dtype)."""
    if plan == "auto":
        plan = tune.best_plan(nq, dtype)
    st = capi.new_typed(nq, dtype)
    circ = capi.Circuit(nq)
    circ.add_packed(packed)
    if plan == "passes" or (plan and nq >= plan + 6):
        tune.prepare(circ, plan)
    circ.run(st, in_place=True)
    return st


def run_operation(spec, states, dtype, one_pass_join=False):
    """The request's final operation on its resident register(s).  With
    one_pass_join a join and its marginals are made in one pass over the
    joined state (dk_join_prob_all), else as a join and then prob_all."""
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
        if one_pass_join:
            _, probs = capi.join_prob_all(states[0], states[1])
            return {"probs": probs}
        j = capi.join(states[0], states[1])
        return {"probs": capi.prob_all(j)}
    raise ValueError("unknown operation %r" % op)


class Server:
    """Continuous-batching request server on the library's device."""

    def __init__(self, workers=6, small_qubits=SMALL_QUBITS, plan="auto",
                 budget=None, max_batch=2048, dtype=capi.C128):
        self.dtype = dtype
        self.small_qubits = small_qubits
        self.plan = plan
        self.max_batch = max_batch
        if budget is None:
            b = capi.C.c_uint64()
            capi.check(capi.lib().dk_memory_budget(capi.C.byref(b), None))
            budget = int(b.value)
        self.budget = budget
        self.used = 0
        self._lock = threading.Lock()
        self._pending = collections.deque()
        self._small = collections.deque()
        self._small_cv = threading.Condition(self._lock)
        self._large = collections.deque()
        self._large_cv = threading.Condition(self._lock)
        self._stop = False
        self.batches = 0          # launches of the batched engine
        self.batched = 0          # requests run in them
        self._threads = [threading.Thread(target=self._batcher, daemon=True)]
        self._threads += [threading.Thread(target=self._worker, daemon=True)
                          for _ in range(workers)]
        for t in self._threads:
            t.start()

    # ---- admission -----------------------------------------------------------

    def is_small(self, spec):
        return (spec["op"] in ("prob", "measure", "density")
                and spec["regs"][0][0] <= self.small_qubits
                and len(spec["keep"]) <= 6)

    def submit(self, spec):
        h = Handle(spec, self.dtype)
        h.t_submit = time.perf_counter()
        with self._lock:
            if h.need > self.budget:
                _finish(h, error=Rejected(
                    "request %d needs %d bytes of states; the budget is %d"
                    % (spec["rid"], h.need, self.budget)))
                return h
            self._pending.append(h)
            self._admit_locked()
        return h

    def _admit_locked(self):
        # FIFO: a request waits while the one ahead of it does not fit;
        # batched requests (a few MiB at most) pass a waiting one when they fit
        blocked = False
        keep = collections.deque()
        while self._pending:
            h = self._pending.popleft()
            fits = self.used + h.need <= self.budget
            small = self.is_small(h.spec)
            if fits and (not blocked or small):
                self.used += h.need
                if small:
                    self._small.append(h)
                    self._small_cv.notify()
                else:
                    self._large.append(h)
                    self._large_cv.notify()
            else:
                blocked = blocked or not small
                keep.append(h)
        self._pending = keep

    def _release(self, need):
        with self._lock:
            self.used -= need
            self._admit_locked()

    # ---- execution -------------------------------------------------------------

    def _batcher(self):
        lane = capi.Lane()
        lane.bind()
        b = capi.Batch(self.dtype, probs=1)
        while True:
            with self._lock:
                while not self._small and not self._stop:
                    self._small_cv.wait()
                if self._stop and not self._small:
                    return
                items = []
                while self._small and len(items) < self.max_batch:
                    items.append(self._small.popleft())
            t_start = time.perf_counter()
            for h in items:
                h.t_start = t_start
            try:
                b.clear()
                for h in items:
                    s = h.spec
                    n = s["regs"][0][0]
                    b.add_packed(n, h.packed[0],
                                 s["rolls"] if s["op"] == "measure" else None,
                                 capi.mask(s["keep"]) if s["op"] == "density" else 0)
                b.run()
                self.batches += 1
                self.batched += len(items)
                for i, h in enumerate(items):
                    op = h.spec["op"]
                    if op == "prob":
                        res = {"probs": b.probs(i)}
                    elif op == "measure":
                        outs, p1 = b.outcomes(i)
                        res = {"outcomes": outs, "p1": p1}
                    else:
                        res = {"rho": b.density(i)}
                    self._release(h.need)
                    _finish(h, res)
            except Exception as e:  # noqa: BLE001 - handed to the waiters
                for h in items:
                    if not h.done:
                        self._release(h.need)
                        _finish(h, error=e)

    def _worker(self):
        lane = capi.Lane()
        lane.bind()
        while True:
            with self._lock:
                while not self._large and not self._stop:
                    self._large_cv.wait()
                if self._stop and not self._large:
                    return
                h = self._large.popleft()
            h.t_start = time.perf_counter()
            try:
                res = self._run_large(h)
                err = None
            except Exception as e:  # noqa: BLE001 - handed to the waiter
                res, err = None, e
            self._release(h.need)
            _finish(h, res, err)

    def _run_large(self, h):
        states = [build_register(n, pk, self.dtype, self.plan)
                  for (n, _), pk in zip(h.spec["regs"], h.packed)]
        res = run_operation(h.spec, states, self.dtype, one_pass_join=True)
        del states  # freed on this thread's lane
        return res

    def close(self):
        with self._lock:
            self._stop = True
            self._small_cv.notify_all()
            self._large_cv.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def serve_trace(server, specs, time_scale=1.0):
    """Submits the requests at their arrival times (scaled by time_scale) and
    waits for all of them.  Returns (handles, t0)."""
    t0 = time.perf_counter()
    hs = []
    for s in specs:
        due = t0 + s["arrival"] * time_scale
        now = time.perf_counter()
        if due > now:
            time.sleep(due - now)
        hs.append(server.submit(s))
    for h in hs:
        h._ev.wait()
    return hs, t0


def run_plain(specs, dtype=capi.C128, time_scale=1.0):
    """Plain per-request execution: requests one at a time in arrival order,
    gate by gate out of place, on the default lane.  Returns a list of
    (result or exception, latency seconds) and the makespan."""
    budget = capi.C.c_uint64()
    capi.check(capi.lib().dk_memory_budget(capi.C.byref(budget), None))
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
            res = run_operation(s, states, dtype)
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
