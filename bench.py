#!/usr/bin/env python3
"""Benchmark of the state-vector core on one MI355X.

Workload (the plain per-gate GPU path): a 30-qubit complex128 register, the
brickwork circuit — each layer is U3 (angles ~ U[0, 2pi)) on every qubit then
CZ on alternating neighbour pairs — applied one `registry_apply` at a time
through the drop-in `doki` module in its per-gate execution mode: every gate
is one pass over the state with the norm fused (as the reference's
apply_gate), run in place once the registry it was applied to is no longer
referenced (the caller's loop drops it).  One step = one layer (30 U3 + 15
or 14 CZ).

value = whole-job HBM GB/s per gate = gates * 32 * 2^n bytes / timed seconds
(every unfused gate reads and writes each amplitude once: 16 B + 16 B).

Also reported: the gate kernel's own duration from HIP events recorded on
the library's stream around every gate of the timed region (roofline),
per-target kernel GB/s (low and high stride),
registry_prob on every qubit, and the reference's CPU path timed on the host
cores over a bounded sample (cpu_baseline).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# target sets of the dense kernels' by-position table (low, mixed, middle,
# high and spread index bits)
DENSE_TARGET_SETS = ((0, 1, 2, 3, 4, 5), (3, 4, 5, 6, 7, 8), (10, 11, 12, 13, 14, 15),
                     (20, 21, 22, 23, 24, 25), (24, 25, 26, 27, 28, 29),
                     (1, 7, 13, 19, 25, 29))
METRIC = ("HBM GB/s per gate at 30q complex128; served circuits/s + p99 "
          "latency vs plain")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class phase:
    """Logs the host clock (CLOCK_MONOTONIC and CLOCK_BOOTTIME, ns) at the
    start and end of a bench phase, so that a kernel trace of the run can be
    cut into the phases."""

    def __init__(self, name):
        self.name = name

    @staticmethod
    def now():
        return time.clock_gettime_ns(time.CLOCK_MONOTONIC), \
            time.clock_gettime_ns(time.CLOCK_BOOTTIME)

    def __enter__(self):
        self.t0 = self.now()

    def __exit__(self, *exc):
        t1 = self.now()
        log(f"[bench] phase {self.name} monotonic_ns {self.t0[0]} {t1[0]} "
            f"boottime_ns {self.t0[1]} {t1[1]}")


def run_layer(d, reg, layer):
    for g, t, c in layer:
        reg = d.registry_apply(reg, g, t, c, None, -1, False)
    return reg


def make_layers(d, W, n, count, seed=2025):
    rng = W.rng_for(seed)
    layers = []
    for li in range(count):
        steps = W.brickwork_layer(rng, n, li)
        layers.append([(d.gate_new(1, m.tolist(), False), t, set(c))
                       for m, t, c, a in steps])
    return layers


def gpu_leg(args):
    import doki_amd
    from doki_amd import capi, workloads as W
    d = doki_amd.doki
    d.execution_mode("gates", False)
    n = args.qubits
    bpg = 32 * (1 << n)
    layers = make_layers(d, W, n, args.warmup + args.steps + 1)
    log(f"[bench] n={n} state={16 * (1 << n) / 1e9:.2f} GB, "
        f"steps={args.steps} warmup={args.warmup}")
    reg = d.registry_new(n, False)
    for li in range(args.warmup):
        reg = run_layer(d, reg, layers[li])
    d.registry_sync(reg, False)
    log("[bench] warmup done")
    # every gate kernel of the timed region records a start and an end HIP
    # event on the library's stream (no synchronisation in between)
    ntimed = sum(len(layers[li]) for li in range(args.warmup, args.warmup + args.steps))
    capi.tune("time_ring", ntimed)
    t0 = time.perf_counter()
    ngates = 0
    for li in range(args.warmup, args.warmup + args.steps):
        reg = run_layer(d, reg, layers[li])
        ngates += len(layers[li])
    d.registry_sync(reg, False)      # the last gate runs, then the device drains
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kms = capi.time_ring_read()
    assert len(kms) == ngates, (len(kms), ngates)
    log(f"[bench] timed {ngates} gates in {elapsed:.3f} s")
    # the same durations by gate kind and target
    by_target = {}
    gi = 0
    for li in range(args.warmup, args.warmup + args.steps):
        for g, t, c in layers[li]:
            key = ("cz" if c else "u3") + str(t[0])
            by_target.setdefault(key, []).append(kms[gi])
            gi += 1
    by_target = {k: round(bpg / (sum(v) / len(v) * 1e-3) / 1e9)
                 for k, v in sorted(by_target.items(), key=lambda kv: (kv[0][:2], int(kv[0][2:])))}

    # per-target kernel durations: five back-to-back gates on every target
    per_target = {}
    u3 = layers[-1][0][0]
    for t in range(n):
        for _ in range(5):
            reg = d.registry_apply(reg, u3, [t], set(), None, -1, False)
        d.registry_sync(reg, False)
        ms = capi.time_ring_read()
        per_target[str(t)] = {"kernel_ms": min(ms), "avg_ms": sum(ms) / len(ms),
                              "GBps": bpg / (min(ms) * 1e-3) / 1e9}
    capi.tune("time_ring", 0)
    kavg = sum(kms) / len(kms)
    log(f"[bench] gate kernel avg {kavg:.3f} ms -> {bpg / kavg / 1e6:.0f} GB/s")

    # a plain copy of the state (registry_clone: 16 B read + 16 B written per
    # amplitude, one kernel): the practical streaming ceiling on this device,
    # and the kernel the HBM counters are calibrated on
    copy_ms = []
    for _ in range(3):
        capi.event_record(0)
        c = d.registry_clone(reg, -1, False)
        capi.event_record(1)
        copy_ms.append(capi.event_elapsed_ms(0, 1))
        del c
    log(f"[bench] state copy {min(copy_ms):.3f} ms -> "
        f"{bpg / min(copy_ms) / 1e6:.0f} GB/s")

    # registry_prob on every qubit (end of the circuit)
    tp0 = time.perf_counter()
    probs = [d.registry_prob(reg, q, -1, False) for q in range(n)]
    tp1 = time.perf_counter()
    assert all(0.0 <= p <= 1.0 + 1e-12 for p in probs)

    d.execution_mode("fuse", False)
    return {
        "elapsed": elapsed, "ngates": ngates, "bpg": bpg, "kavg_ms": kavg,
        "kernel_GBps_timed_by_target": by_target, "per_target": per_target,
        "prob_all_ms": (tp1 - tp0) * 1e3, "copy_ms": min(copy_ms),
    }


def join_leg(n):
    """A join to n qubits (registers of n/2 and n - n/2 qubits) followed by
    the marginal probabilities of every qubit of the result: as a join and a
    prob_all pass over the joined state, and in one pass (dk_join_prob_all:
    the joined state written once and not read back)."""
    import numpy as np

    from doki_amd import capi
    from doki_amd import workloads as W
    a = capi.from_host(W.random_normalised_state(5, n // 2))
    b = capi.from_host(W.random_normalised_state(6, n - n // 2))
    two, one = [], []
    for _ in range(3):
        capi.event_record(0)
        j = capi.join(a, b)
        p2 = capi.prob_all(j)
        capi.event_record(1)
        two.append(capi.event_elapsed_ms(0, 1))
        del j
        capi.event_record(0)
        j, p1 = capi.join_prob_all(a, b)
        capi.event_record(1)
        one.append(capi.event_elapsed_ms(0, 1))
        del j
    assert np.max(np.abs(p1 - p2)) <= 1e-14    # the same sums, other orders
    written = 16 << n
    out = {"qubits": n, "join_then_prob_all_ms": round(min(two), 3),
           "one_pass_ms": round(min(one), 3),
           "one_pass_written_GBps": round(written / (min(one) * 1e-3) / 1e9, 1)}
    log(f"[bench] join + marginals at {n} qubits: {out}")
    return out


def big_leg(args):
    """The 33-qubit variant of the brickwork workload (137.4 GB register):
    one layer through the drop-in module, out of place as registry_apply is
    (old and new state together: 274.9 GB), and the same layer in place
    through the C ABI; per-gate HBM rates and the gate kernel's duration."""
    import doki_amd
    from doki_amd import capi, workloads as W
    d = doki_amd.doki
    n = args.big_qubits
    bpg = 32 * (1 << n)
    steps = W.brickwork_layer(W.rng_for(3333), n, 0)
    out = {"qubits": n, "state_GB": round(16 * (1 << n) / 1e9, 2),
           "gates": len(steps), "bytes_per_gate": bpg}
    gates = [(d.gate_new(1, m.tolist(), False), t, set(c)) for m, t, c, a in steps]
    # (loops inline: a caller's reference would keep a third state alive)
    for mode, key in (("eager", "out_of_place"), ("gates", "drop_in_in_place")):
        prev = d.execution_mode(mode, False)
        reg = d.registry_new(n, False)
        g0, t_, c_ = gates[0]
        reg = d.registry_apply(reg, g0, t_, c_, None, -1, False)  # second state
        d.registry_sync(reg, False)
        t0 = time.perf_counter()
        for g, t_, c_ in gates:
            reg = d.registry_apply(reg, g, t_, c_, None, -1, False)
        d.registry_sync(reg, False)
        el = time.perf_counter() - t0
        d.execution_mode(prev, False)
        out[f"{key}_GBps"] = round(len(gates) * bpg / el / 1e9, 1)
        out[f"{key}_ms_per_gate"] = round(el / len(gates) * 1e3, 3)
        del reg
        capi.synchronize()
    s = capi.new(n)
    cg = [(capi.Gate(m), t, c) for m, t, c, a in steps]
    capi.apply(s, cg[0][0], cg[0][1], cg[0][2], in_place=True)
    capi.synchronize()
    capi.tune("time_ring", len(cg))
    t0 = time.perf_counter()
    for g, t, c in cg:
        capi.apply(s, g, t, c, in_place=True)
    capi.synchronize()
    el = time.perf_counter() - t0
    kms = capi.time_ring_read()
    capi.tune("time_ring", 0)
    kavg = sum(kms) / len(kms)
    out["in_place_GBps"] = round(len(cg) * bpg / el / 1e9, 1)
    out["kernel_avg_ms"] = round(kavg, 3)
    out["kernel_GBps"] = round(bpg / (kavg * 1e-3) / 1e9, 1)
    p = capi.prob_all(s)
    assert all(0.0 <= x <= 1.0 + 1e-12 for x in p)
    del s
    capi.synchronize()
    log(f"[bench] {n} qubits: {out}")
    return out


def dropin_leg(args):
    """The brickwork circuit (the headline register, depth args.fused_depth)
    and the marginals of every qubit, through plain `doki` calls only, under
    each execution mode of the drop-in module: "eager" (each registry_apply
    launches at once, out of place, as the reference does), "gates" (one
    pass per gate, in place once the input registry is unreferenced) and
    "fuse" (the default: the gates run as cache-blocked passes when the
    registry is read).  Circuits/s, and each mode's marginals against
    eager's."""
    import numpy as np

    import doki_amd
    from doki_amd import capi, workloads as W
    d = doki_amd.doki
    n = args.qubits
    steps = W.brickwork_circuit(77, n, args.fused_depth)
    gates = [(d.gate_new(1, m.tolist(), False), t, set(c)) for m, t, c, a in steps]
    out = {"qubits": n, "depth": args.fused_depth, "gates": len(steps)}
    probs = {}
    for mode in ("eager", "gates", "fuse"):
        prev = d.execution_mode(mode, False)
        runs = []
        for rep in range(2):            # the first run warms the gate uploads
            reg = d.registry_new(n, False)
            d.registry_sync(reg, False)
            t0 = time.perf_counter()
            for g, t, c in gates:
                reg = d.registry_apply(reg, g, t, c, None, -1, False)
            p = [d.registry_prob(reg, q, -1, False) for q in range(n)]
            runs.append(time.perf_counter() - t0)
            del reg
            capi.synchronize()
        d.execution_mode(prev, False)
        probs[mode] = np.array(p)
        out[mode] = {"seconds": round(runs[-1], 4),
                     "circuits_per_s": round(1.0 / runs[-1], 3)}
        if mode != "eager":
            out[mode]["max_abs_prob_vs_eager"] = float(
                np.max(np.abs(probs[mode] - probs["eager"])))
        log(f"[bench] drop-in {mode}: {out[mode]}")
    for mode in ("gates", "fuse"):
        out[f"{mode}_speedup_vs_eager"] = round(
            out["eager"]["seconds"] / out[mode]["seconds"], 2)
    return out


def fused_leg(args):
    """Fused dense gates on the matrix cores at the same register size:
    per-k kernel rates (FP64 complex128 and FP32 complex64), circuits of
    Haar-random k-qubit blocks in both precisions (FP32 error and fidelity
    against FP64), and the brickwork circuit plain (per gate) vs fused dense
    blocks vs cache-blocked passes (circuits/s), each with its error and
    fidelity against the plain FP64 result."""
    import numpy as np

    from doki_amd import capi, workloads as W
    n = args.qubits
    out = {"dense_kernel": {}, "dense_by_target_ms": {}}
    dense_by_target = out["dense_by_target_ms"]
    rng = W.rng_for(4242)
    hgate = capi.Gate(W.gate_h())

    def dense_state(dtype):
        """|+...+>: every amplitude nonzero (and known to be), so that no
        kernel can skip tiles of zeros."""
        s = capi.new_typed(n, dtype)
        for q in range(n):
            capi.apply(s, hgate, [q], in_place=True)
        return s

    for dtype, name, amp_bytes in ((capi.C128, "f64", 16), (capi.C64, "f32", 8)):
        s = dense_state(dtype)
        capi.tune("time_ring", 16)
        for k in (4, 5, 6):
            g = capi.Gate(W.haar_unitary(rng, k))
            for rep in range(3):
                tg = list(range(rep * 7 % (n - k), rep * 7 % (n - k) + k))
                capi.apply(s, g, tg, in_place=True)
            ms = capi.time_ring_read()
            t = min(ms) * 1e-3
            out["dense_kernel"][f"{name}_k{k}"] = {
                "kernel_ms": round(t * 1e3, 3),
                "GBps": round(2 * amp_bytes * (1 << n) / t / 1e9, 1),
                "TFLOPs": round(8 * (1 << k) * (1 << n) / t / 1e12, 2)}
            # the same gate by target position and kernel form (median of
            # three launches; the k leading targets of each set)
            for form, fname in ((0, "tiles"), (2, "streaming"), (1, "auto")):
                capi.tune("dense_stream", form)
                row = {}
                for ts in DENSE_TARGET_SETS:
                    if max(ts[:k]) >= n:
                        continue
                    for rep in range(3):
                        capi.apply(s, g, list(ts[:k]), in_place=True)
                    row[",".join(map(str, ts[:k]))] = round(
                        float(np.median(capi.time_ring_read())), 3)
                dense_by_target[f"{name}_k{k}_{fname}"] = row
            capi.tune("dense_stream", 1)
        capi.tune("time_ring", 0)
        del s
        capi.synchronize()
    log(f"[bench] dense kernels {out['dense_kernel']}")
    log(f"[bench] dense kernels by target (ms) {dense_by_target}")

    # circuits of Haar-random k-qubit blocks (four layers of blocks across
    # the register, shifted per layer), in FP64 and in complex64/FP32; the
    # FP32 result against the FP64 one over the whole state
    haar = {}
    for k in (4, 5, 6):
        blocks = []
        for layer in range(4):
            off = (layer * (k // 2 + 1)) % k
            for j in range(off, n - k + 1, k):
                blocks.append((capi.Gate(W.haar_unitary(rng, k)), list(range(j, j + k))))
        rk = {"blocks": len(blocks)}
        states = {}
        for dtype, name in ((capi.C128, "f64"), (capi.C64, "f32")):
            s = capi.new_typed(n, dtype)
            for g, tg in blocks[:2]:
                capi.apply(s, g, tg, in_place=True)    # warm (gate uploads)
            s = capi.new_typed(n, dtype)
            capi.synchronize()
            t0 = time.perf_counter()
            for g, tg in blocks:
                capi.apply(s, g, tg, in_place=True)
            capi.synchronize()
            rk[f"{name}_seconds"] = round(time.perf_counter() - t0, 4)
            states[name] = s
            # the same blocks on a dense input (no tiles of zeros to skip)
            d = dense_state(dtype)
            capi.synchronize()
            t0 = time.perf_counter()
            for g, tg in blocks:
                capi.apply(d, g, tg, in_place=True)
            capi.synchronize()
            rk[f"{name}_seconds_dense_input"] = round(time.perf_counter() - t0, 4)
            del d
        inner, mx = capi.compare(states["f64"], states["f32"])
        rk["f32_max_abs_err_vs_f64"] = mx
        rk["f32_fidelity_vs_f64"] = abs(inner) ** 2
        del states, s
        capi.synchronize()
        haar[f"k{k}"] = rk
        log(f"[bench] haar blocks k={k}: {rk}")
    out["haar_block_circuits"] = haar

    depth = args.fused_depth
    steps = W.brickwork_circuit(77, n, depth)
    res = {}
    for label, k, dtype in (("plain_f64", 0, capi.C128), ("fused5_f64", 5, capi.C128),
                            ("fused4_f64", 4, capi.C128), ("fused6_f64", 6, capi.C128),
                            ("fused5_f32", 5, capi.C64), ("passes_f64", "passes", capi.C128),
                            ("passes_f32", "passes", capi.C64)):
        circ = capi.Circuit(n, steps)
        nb = circ.passes() if k == "passes" else circ.fuse(k)
        st = capi.new_typed(n, dtype)
        circ.run(st, in_place=True)      # warm (gate uploads)
        capi.synchronize()
        st = capi.new_typed(n, dtype)
        capi.synchronize()
        t0 = time.perf_counter()
        circ.run(st, in_place=True)
        capi.synchronize()
        el = time.perf_counter() - t0
        res[label] = {"passes": nb or len(steps), "seconds": round(el, 4),
                      "circuits_per_s": round(1.0 / el, 3)}
        if label == "plain_f64":
            ref_state = st
        else:
            # against the plain FP64 per-gate result, whole state (one pass
            # over both on the GPU)
            inner, mx = capi.compare(ref_state, st)
            res[label]["max_abs_err_vs_plain_f64"] = mx
            res[label]["fidelity_vs_plain_f64"] = abs(inner) ** 2
        if label != "plain_f64":
            del st
        capi.synchronize()
        if label != "plain_f64":
            # the same plan on a dense input: from |0...0> the first passes
            # (blocks) find most tiles outside the state's support and skip
            # them; here every tile is computed
            d = dense_state(dtype)
            capi.synchronize()
            t0 = time.perf_counter()
            circ.run(d, in_place=True)
            capi.synchronize()
            res[label]["seconds_dense_input"] = round(time.perf_counter() - t0, 4)
            del d
            capi.synchronize()
        log(f"[bench] {label}: {res[label]}")
    out["brickwork_circuit"] = {"qubits": n, "depth": depth, "gates": len(steps), **res}
    del ref_state
    capi.synchronize()
    out["fused_speedup_vs_plain"] = round(
        res["plain_f64"]["seconds"] / res["fused5_f64"]["seconds"], 2)
    out["passes_speedup_vs_plain"] = round(
        res["plain_f64"]["seconds"] / res["passes_f64"]["seconds"], 2)
    return out


def batch_leg(args):
    """Many small registers: one persistent launch (graph replay) against
    plain per-gate execution of the same registers one at a time."""
    import numpy as np

    from doki_amd import capi
    from doki_amd.workloads import batch_specs
    t0 = time.perf_counter()
    specs = batch_specs(31, args.batch_registers)
    log(f"[bench] batch: {len(specs)} registers generated in "
        f"{time.perf_counter() - t0:.1f} s")
    out = {"registers": len(specs),
           "gates": int(sum(len(s[1]) for s in specs)),
           "qubits_hist": {int(n): int(c) for n, c in
                           zip(*np.unique([s[0] for s in specs], return_counts=True))}}
    kept = None
    for dtype, name in ((capi.C128, "c128"), (capi.C64, "c64")):
        b = capi.Batch(dtype, use_graph=1)
        for n, steps, rolls, keep in specs:
            b.add(n, steps, rolls, keep)
        b.run()                          # upload + capture
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            b.run()
            ts.append(time.perf_counter() - t1)
        t = min(ts)
        out[name] = {"seconds": round(t, 4),
                     "circuits_per_s": round(len(specs) / t, 1)}
        log(f"[bench] batch {name}: {out[name]}")
        if dtype == capi.C128:
            kept = b                     # its results are checked below
        else:
            del b
    # plain: each register through the per-gate C ABI path, one at a time
    rng = np.random.default_rng(5)
    sample = rng.choice(len(specs), size=min(len(specs), args.batch_plain_sample),
                        replace=False)
    gate_cache = {}
    plain = {}
    t1 = time.perf_counter()
    for i in sample:
        n, steps, rolls, keep = specs[i]
        s = capi.new(n)
        for m, t, c, a in steps:
            key = m.tobytes()
            g = gate_cache.get(key)
            if g is None:
                g = gate_cache[key] = capi.Gate(m)
            capi.apply(s, g, t, c, a, in_place=True)
        rho = capi.reduced_density(s, keep) if keep else None
        _, outs, p1 = capi.measure(s, (1 << n) - 1, rolls)
        plain[int(i)] = (outs, p1, rho)
        del s
    capi.synchronize()
    tp = time.perf_counter() - t1
    # the batched c128 results of the sampled registers against the plain
    # path: outcomes identical (except where a roll lies within 1e-9 of its
    # P1), P1 and densities within 1e-12
    worst = 0.0
    for i, (outs, p1, rho) in plain.items():
        bo, bp = kept.outcomes(i)
        rolls = specs[i][2]
        if bo != outs and not any(abs(p - r) < 1e-9 for p, r in zip(p1, rolls)):
            raise AssertionError(f"batched register {i}: outcomes differ from "
                                 "the plain path")
        worst = max(worst, float(np.max(np.abs(bp - p1))))
        if rho is not None:
            worst = max(worst, float(np.max(np.abs(kept.density(i) - rho))))
    if worst > 1e-12:
        raise AssertionError(f"batched vs plain: max deviation {worst:.3e}")
    out["checked_vs_plain"] = {"registers": len(plain), "max_abs": worst}
    del kept
    per = tp / len(sample)
    out["plain_c128"] = {"sample": int(len(sample)), "seconds_per_register": round(per, 6),
                         "circuits_per_s": round(1.0 / per, 1)}
    out["batched_speedup_vs_plain_c128"] = round(
        out["c128"]["circuits_per_s"] / out["plain_c128"]["circuits_per_s"], 1)
    log(f"[bench] batch plain: {out['plain_c128']}")
    return out


def serve_leg(args):
    """Served mix (Poisson arrivals, 8-30 qubits, prob/measure/density/join):
    the continuous-batching server against plain per-request execution of the
    same arrival trace."""
    import numpy as np

    from doki_amd import capi, serve
    from doki_amd import workloads as W
    specs = W.serving_requests(41, args.serve_requests, args.serve_qps)
    for s in specs:
        # the client sends its gates packed
        s["packed"] = [capi.pack_steps(st) for _, st in s["regs"]]
    # warm-up outside the trace: one batched and one lane request
    warm = W.serving_requests(42, 40, 1000.0, nmax=20, depth=(20, 40))
    with serve.Server(workers=args.serve_workers) as srv:
        for h in [srv.submit(s) for s in warm]:
            try:
                h.wait()
            except serve.Rejected:
                pass                 # a join too large for the device
    out = {"requests": len(specs), "qps_offered": args.serve_qps,
           "ops": {op: sum(1 for s in specs if s["op"] == op) for op in W.SERVE_OPS},
           "qubits_max": int(max(s["regs"][0][0] for s in specs))}

    def served(time_scale):
        with serve.Server(workers=args.serve_workers) as srv:
            hs, t0 = serve.serve_trace(srv, specs, time_scale)
            batches, batched = srv.batches, srv.batched
        ok = [h for h in hs if h.error is None]
        rej = [h for h in hs if isinstance(h.error, serve.Rejected)]
        failed = [h for h in hs if h.error is not None and h not in rej]
        if failed:
            raise failed[0].error
        lat = [h.t_done - (t0 + h.spec["arrival"] * time_scale) for h in ok]
        span = max(h.t_done for h in hs) - t0
        return hs, {"completed": len(ok), "rejected_over_budget": len(rej),
                    "seconds": round(span, 3),
                    "throughput_rps": round(len(ok) / span, 2),
                    "batch_launches": batches, "batched_requests": batched,
                    "workers": args.serve_workers, **serve.latency_stats(lat)}

    def plain_run(time_scale):
        res, span = serve.run_plain(specs, time_scale=time_scale)
        pl = [lt for r, lt in res if not isinstance(r, Exception)]
        return res, {"completed": len(pl), "seconds": round(span, 3),
                     "throughput_rps": round(len(pl) / span, 2),
                     **serve.latency_stats(pl)}

    # at the offered rate: latency; all requests at once: capacity
    hs, out["served"] = served(1.0)
    log(f"[bench] serve: {out['served']}")
    plain_res, out["plain"] = plain_run(1.0)
    log(f"[bench] serve plain: {out['plain']}")
    with phase("served_burst"):
        _, out["served_burst"] = served(0.0)
    log(f"[bench] serve burst: {out['served_burst']}")
    with phase("plain_burst"):
        _, out["plain_burst"] = plain_run(0.0)
    log(f"[bench] serve plain burst: {out['plain_burst']}")
    # load curve: the same trace offered at other rates (arrival times
    # scaled), up to saturation; plain beyond its capacity only queues
    curve = []
    for qps in args.serve_curve:
        scale = args.serve_qps / qps
        _, sv = served(scale)
        pt = {"qps_offered": qps, "served": {k: sv[k] for k in (
            "throughput_rps", "p50_ms", "p99_ms", "completed")}}
        if qps <= 2 * args.serve_qps:
            _, pl = plain_run(scale)
            pt["plain"] = {k: pl[k] for k in ("throughput_rps", "p50_ms", "p99_ms")}
        curve.append(pt)
        log(f"[bench] serve curve: {pt}")
    out["curve"] = curve
    out["speedup_p99"] = round(out["plain"]["p99_ms"] / out["served"]["p99_ms"], 2)
    out["speedup_capacity"] = round(out["served_burst"]["throughput_rps"] /
                                    out["plain_burst"]["throughput_rps"], 2)
    # agreement of the two paths on every completed request
    worst = 0.0
    flips = 0
    for h, (res, _) in zip(hs, plain_res):
        if h.error is not None or isinstance(res, Exception):
            continue
        for k in ("probs", "rho"):
            if k in res:
                worst = max(worst, float(np.max(np.abs(h.result[k] - res[k]))))
        if "outcomes" in res and res["outcomes"] != h.result["outcomes"]:
            flips += 1
    out["served_vs_plain_max_abs"] = worst
    out["served_vs_plain_outcome_mismatches"] = flips
    return out


def cpu_model():
    """The host CPU's model name (as lscpu prints it), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(args):
    """The reference's own CPU path (oracle/_ref: its extension compiled from
    its sources) on the host cores, over bounded samples of the workloads:

    * the headline workload itself: brickwork gates (U3 on a low and a high
      target, a CZ) on a 30-qubit register with every host thread the run
      may use: seconds per gate and GB/s (the value reported);
    * one brickwork layer at 24 qubits, with those threads and with one;
    * the 12-qubit, 200-gate parity circuit in full (marginals and a full
      measurement), with those threads and with one;
    * a subset of the batched-registers workload, registers one at a time.

    The thread count goes to the reference through registry_apply's
    num_threads argument (its omp_set_num_threads)."""
    from oracle import oracle as O
    from doki_amd import workloads as W
    ref = O.load_reference()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    info = {"cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    if ref is None:
        # the single-threaded oracle port, 24 qubits
        n = args.cpu_qubits
        steps = W.brickwork_layer(W.rng_for(7), n, 0)
        st = O.new(n)
        t0 = time.perf_counter()
        g = 0
        while time.perf_counter() - t0 < args.cpu_seconds and g < len(steps):
            m, t, c, a = steps[g]
            st = O.apply(st, m, t, sorted(c), sorted(a))
            g += 1
        el = time.perf_counter() - t0
        return {"value": round(g * 32 * (1 << n) / el / 1e9, 3), "unit": "GB/s",
                "cores": 1, "kind": "port",
                "sample": f"{g} brickwork gates at n={n} through the single-threaded "
                          "oracle port", **info}

    def gate(m):
        return ref.gate_new(1, m.tolist(), False)

    def run_gates(reg, gl, nthr, budget=None):
        t0 = time.perf_counter()
        done = 0
        for gg, t, c in gl:
            reg = ref.registry_apply(reg, gg, t, c, set(), nthr, False)
            done += 1
            if budget is not None and time.perf_counter() - t0 > budget:
                break
        return reg, done, time.perf_counter() - t0

    # headline register: 30 qubits, a low-target U3, a high-target U3, a CZ
    n = args.qubits
    lay = W.brickwork_layer(W.rng_for(7), n, 0)
    pick = [lay[0], lay[n - 1], lay[n]]            # U3 t=0, U3 t=n-1, CZ (0,1)
    gl = [(gate(m), t, set(c)) for m, t, c, a in pick]
    reg = ref.registry_new(n, False)
    reg, _, _ = run_gates(reg, gl[:1], threads)    # first touch of both states
    per = {}
    for (gg, t, c), name in zip(gl, ("u3_t0", f"u3_t{n - 1}", "cz_0_1")):
        reg, _, el = run_gates(reg, [(gg, t, c)], threads)
        per[name] = round(el, 3)
    del reg
    sec = sum(per.values()) / len(per)
    value = 32 * (1 << n) / sec / 1e9

    # one brickwork layer at 24 qubits, all threads and one thread
    m24 = args.cpu_qubits
    lay24 = [(gate(m), t, set(c)) for m, t, c, a in W.brickwork_layer(W.rng_for(8), m24, 0)]
    layer24 = {}
    for nthr in (threads, 1):
        reg = ref.registry_new(m24, False)
        reg, done, el = run_gates(reg, lay24, nthr, budget=args.cpu_seconds / 3)
        layer24[f"threads_{nthr}"] = {"gates": done, "seconds_per_gate": round(el / done, 4),
                                      "GBps": round(done * 32 * (1 << m24) / el / 1e9, 2)}
        del reg

    # the 12-qubit parity circuit in full
    steps, rolls = W.random_circuit(0, 12, 200)
    c0g = [(ref.gate_new(1, m.tolist(), False), t, set(c), set(a)) for m, t, c, a in steps]
    c0 = {}
    for nthr in (threads, 1):
        t0 = time.perf_counter()
        reg = ref.registry_new(12, False)
        for gg, t, c, a in c0g:
            reg = ref.registry_apply(reg, gg, t, c, a, nthr, False)
        _ = [ref.registry_prob(reg, q, nthr, False) for q in range(12)]
        ref.registry_measure(reg, (1 << 12) - 1, rolls, nthr, False)
        c0[f"threads_{nthr}_seconds"] = round(time.perf_counter() - t0, 4)

    # batched-registers subset, one register at a time (all threads)
    specs = W.batch_specs(31, args.cpu_batch_registers)
    t0 = time.perf_counter()
    for nq, st_, rl, keep in specs:
        reg = ref.registry_new(nq, False)
        for m, t, c, a in st_:
            reg = ref.registry_apply(reg, ref.gate_new(1, m.tolist(), False), t, c, a,
                                     threads, False)
        if keep:
            rho = ref.registry_density(reg, False)
            for e in range(nq - 1, 5, -1):
                rho = ref.funmatrix_partialtrace(rho, e, False)
            ref.funmatrix_get(rho, 0, 0, False)
        ref.registry_measure(reg, (1 << nq) - 1, rl, threads, False)
    el = time.perf_counter() - t0
    batch = {"registers": len(specs), "seconds": round(el, 3),
             "circuits_per_s": round(len(specs) / el, 1)}
    return {"value": round(value, 3), "unit": "GB/s", "cores": threads, "kind": "reference",
            "sample": f"U3 t=0, U3 t={n - 1} and CZ(0,1) on a {n}-qubit complex128 "
                      f"register, registry_apply of the reference's OpenMP path "
                      f"with {threads} threads, {sec:.2f} s per gate",
            **info, "headline_register": {"qubits": n, "seconds_per_gate": per},
            "brickwork_layer": {"qubits": m24, **layer24},
            "c0_12q_200_gates": c0, "batch_subset": batch}


def traffic_child(n):
    """Workload of the counter passes (run as a child under rocprofv3): a
    clone of an n-qubit complex128 register (the calibration kernel: it
    reads and writes exactly 16 B per amplitude) and one U3 on every target,
    in place (every form of the one-qubit kernel)."""
    import numpy as np
    import doki_amd
    from doki_amd import workloads as W
    d = doki_amd.doki
    d.execution_mode("gates", False)
    reg = d.registry_new(n, False)
    rng = np.random.default_rng(7)
    for t in range(n):
        g = d.gate_new(1, W.gate_u3(*rng.uniform(0, 2 * np.pi, 3)).tolist(), False)
        reg = d.registry_apply(reg, g, [t], set(), None, -1, False)
    d.registry_sync(reg, False)
    c = d.registry_clone(reg, -1, False)
    d.registry_sync(c, False)
    print("traffic child done", flush=True)


def kernel_short_name(full):
    """'void (anonymous namespace)::k_x<HIP_vector_type<double, 2u>, 1>(...)'
    -> 'k_x<double2, 1>'."""
    s = full.replace("(anonymous namespace)::", "")
    s = s.replace("HIP_vector_type<double, 2u>", "double2")
    s = s.replace("HIP_vector_type<float, 2u>", "float2")
    if s.startswith("void "):
        s = s[5:]
    return s.split("(")[0].strip().replace(" >", ">")


def _pmc_pass(counter, n, tmp):
    """Per-dispatch values of one counter over the traffic child's kernels:
    {kernel name: [KiB per dispatch, ...]}, or None."""
    import csv
    import glob
    import shutil
    import subprocess
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    out = os.path.join(tmp, counter.lower())
    cmd = ["timeout", "-s", "KILL", "120", exe, "--pmc", counter,
           "--output-format", "csv", "-d", out, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__), "--traffic-child",
           "--qubits", str(n)]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if r.returncode != 0 or not files:
        log(f"[bench] {counter} pass failed (rc {r.returncode}): {r.stderr[-400:]}")
        return None
    vals = {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            name = kernel_short_name(row["Kernel_Name"])
            vals.setdefault(name, []).append(float(row["Counter_Value"]))
    return vals


def live_traffic(n):
    """HBM bytes per launch of the one-qubit gate kernels, measured by this
    run: two rocprofv3 passes (FETCH_SIZE, then WRITE_SIZE, in KiB per
    dispatch) over the traffic child, each scaled by what the counter reads
    for the state copy kernel (16 * 2^n B read and written).  Returns
    (bytes per launch averaged over every one-qubit dispatch, a summary
    dict), or (None, reason)."""
    import tempfile
    tmp = tempfile.mkdtemp(prefix="dk_traffic_", dir="/tmp")
    fetch = _pmc_pass("FETCH_SIZE", n, tmp)
    write = _pmc_pass("WRITE_SIZE", n, tmp) if fetch else None
    if not fetch or not write:
        return None, "rocprofv3 counter passes unavailable"
    copy = [k for k in fetch if k.startswith("k_normalised_copy")]
    gates = [k for k in fetch if k.startswith("k_apply1_")]
    if not copy or not gates:
        return None, f"kernels not found in the counter passes: {sorted(fetch)}"
    state_kib = 16 * (1 << n) / 1024
    fscale = state_kib / fetch[copy[0]][0]
    wscale = state_kib / write[copy[0]][0]
    per = {}
    tot, cnt = 0.0, 0
    for k in gates:
        fs, ws = fetch[k], write.get(k, [])
        if len(fs) != len(ws):
            return None, f"dispatch counts differ for {k}"
        b = [(f * fscale + w * wscale) * 1024 for f, w in zip(fs, ws)]
        per[k] = {"dispatches": len(b), "bytes_per_launch": round(sum(b) / len(b)),
                  "over_algorithmic": round(sum(b) / len(b) / (32 * (1 << n)), 4)}
        tot += sum(b)
        cnt += len(b)
    return round(tot / cnt), {
        "source": "measured by this run: rocprofv3 --pmc FETCH_SIZE, then "
                  "WRITE_SIZE, over a clone and one U3 per target at "
                  f"{n} qubits (KiB per dispatch, scaled by the copy kernel: "
                  f"fetch x{fscale:.4f}, write x{wscale:.4f})",
        "by_kernel": per}


def measured_traffic(n):
    """HBM bytes per launch of the 1-qubit gate kernel from the committed
    rocprofv3 summary in profiles/ (copy-kernel calibrated) and where they
    come from; (None, None) when no profile covers this register size.  The
    fallback when this run cannot take its own counter passes."""
    name = f"apply1_{n}q_c128_hbm_traffic.txt"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    src = f"profiles/{name} (the committed profile, not measured by this run)"
    return int(float(vals["traffic_bytes_per_launch"])), src


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--cpu-qubits", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=6.0)
    ap.add_argument("--cpu-batch-registers", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--fused-depth", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-registers", type=int, default=4096)
    ap.add_argument("--batch-plain-sample", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--big-qubits", type=int, default=33,
                    help="qubits of the large-register leg (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-qps", type=float, default=60.0)
    ap.add_argument("--serve-workers", type=int, default=6)
    ap.add_argument("--no-traffic", action="store_true",
                    help="take roofline.traffic from the committed profile "
                         "instead of two rocprofv3 counter passes")
    ap.add_argument("--traffic-child", action="store_true",
                    help=argparse.SUPPRESS)
    ap.add_argument("--serve-curve", type=float, nargs="*",
                    default=[30.0, 120.0, 240.0, 480.0],
                    help="further offered rates (requests/s) of the load curve")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the state-vector core has no multi-GPU path in this build",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be >= 1")
    if args.traffic_child:
        traffic_child(args.qubits)
        return

    # the counter passes run first, while this process has not touched the
    # GPU (their counters would otherwise see its kernels too)
    traffic, traffic_info = (None, None) if args.no_traffic else live_traffic(args.qubits)
    if traffic is None:
        if traffic_info:
            log(f"[bench] live traffic: {traffic_info}")
        traffic, traffic_src = measured_traffic(args.qubits)
        traffic_info = {"source": traffic_src}
    else:
        log(f"[bench] live traffic {traffic} B per launch "
            f"({traffic / (32 * (1 << args.qubits)):.4f} x algorithmic)")

    g = gpu_leg(args)
    big = big_leg(args) if args.big_qubits else None
    joins = join_leg(args.qubits + 2)
    dropin = None if args.no_fused else dropin_leg(args)
    fused = None if args.no_fused else fused_leg(args)
    batched = None if args.no_batch else batch_leg(args)
    served = None if args.no_serve else serve_leg(args)
    cpu = None if args.no_cpu else cpu_leg(args)
    if cpu:
        log(f"[bench] cpu baseline {cpu}")
    value = g["ngates"] * g["bpg"] / g["elapsed"] / 1e9
    achieved = g["bpg"] / (g["kavg_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(g["elapsed"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"brickwork circuit, {args.qubits}-qubit complex128 register, "
                        "plain per-gate registry_apply through the drop-in module "
                        "(execution_mode 'gates': one state pass per gate, fused norm, "
                        "in place once the input registry is unreferenced); "
                        "one step = one layer of U3 on every qubit + CZ pairs",
            "qubits": args.qubits,
            "gates_timed": g["ngates"],
            "bytes_per_gate": g["bpg"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_info,
        },
        "cpu_baseline": cpu,
        "extras": {
            "gate_kernel_avg_ms": round(g["kavg_ms"], 4),
            "state_copy_GBps": round(g["bpg"] / (g["copy_ms"] * 1e-3) / 1e9, 1),
            "per_target_kernel": g["per_target"],
            "timed_region_kernel_GBps_by_gate": g["kernel_GBps_timed_by_target"],
            "registry_prob_all_qubits_ms": round(g["prob_all_ms"], 3),
            "large_register": big,
            "join_prob_all": joins,
            "drop_in_modes": dropin,
            "fused": fused,
            "batched": batched,
            "served": served,
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
