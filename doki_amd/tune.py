"""Autotuning of the circuit plan (cache-blocked passes or a fusion width),
kernel block shapes and batched-engine pass width, measured on the device
and cached per (qubits, amplitude type).

The cache is a JSON file (``DOKI_TUNE_CACHE``, default
``~/.cache/doki_amd/tune-<arch>.json``) mapping ``"<n>:<c128|c64>"`` to the
chosen settings and the timings they were chosen on.  Lookups without an
entry fall back to fixed defaults, so untuned use never touches the GPU for
tuning.
"""
import json
import os
import time

from . import capi
from . import workloads as W

DTYPE_NAMES = {capi.C128: "c128", capi.C64: "c64"}
FUSE_WIDTHS = (0, 3, 4, 5, 6)
# circuit plans: cache-blocked passes, or dense fusion of width k (0 = none)
PLANS = ("passes",) + FUSE_WIDTHS
PASS_MIN_QUBITS = 12


def default_fuse_k(n):
    """Fusion width without a tuning entry: 5-qubit blocks once the register
    is wide enough for the dense kernels, none below."""
    return 5 if n >= 11 else 0


def default_plan(n):
    """Plan without a tuning entry: cache-blocked passes once the register
    holds a whole pass tile, gate by gate below."""
    return "passes" if n >= PASS_MIN_QUBITS else 0


def prepare(circ, plan):
    """Plans circ (a capi.Circuit) with `plan`: "passes" or a fusion width."""
    if plan == "passes":
        return circ.passes()
    if plan:
        return circ.fuse(int(plan))
    return 0


def cache_path():
    p = os.environ.get("DOKI_TUNE_CACHE")
    if p:
        return p
    try:
        arch = capi.device_info()[1].split(":")[0]
    except Exception:  # noqa: BLE001 - no device: a generic name
        arch = "gfx950"
    return os.path.join(os.path.expanduser("~"), ".cache", "doki_amd",
                        f"tune-{arch}.json")


def load_cache(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(data, path=None):
    path = path or cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key(n, dtype):
    return f"{int(n)}:{DTYPE_NAMES[dtype]}"


def lookup(n, dtype, path=None):
    return load_cache(path).get(key(n, dtype))


def best_fuse_k(n, dtype=capi.C128, path=None):
    e = lookup(n, dtype, path)
    if e and "fuse_k" in e:
        return int(e["fuse_k"])
    return default_fuse_k(n)


def best_plan(n, dtype=capi.C128, path=None):
    """The tuned circuit plan for (n, dtype): "passes" or a fusion width."""
    e = lookup(n, dtype, path)
    if e and "plan" in e:
        p = e["plan"]
        return p if p == "passes" else int(p)
    return default_plan(n)


def _time_circuit(n, dtype, steps, plan, reps):
    circ = capi.Circuit(n, steps)
    prepare(circ, plan)
    st = capi.new_typed(n, dtype)
    circ.run(st, in_place=True)            # gate uploads, allocation
    capi.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        circ.run(st, in_place=True)
        capi.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune_fusion(n, dtype=capi.C128, depth=None, seed=7, reps=3):
    """Times a mixed random circuit (depth 4n by default) gate by gate, fused
    into blocks of 3..6 qubits and as cache-blocked passes; returns
    {plan: seconds}."""
    steps, _ = W.random_circuit(seed, n, depth or 4 * n)
    out = {}
    for p in PLANS:
        if p == "passes" and n < PASS_MIN_QUBITS:
            continue
        if p != "passes" and p and n < p:
            continue
        out[p] = _time_circuit(n, dtype, steps, p, reps)
    return out


def _kernel_ms(n, dtype, gate, targets, reps=3):
    s = capi.new_typed(n, dtype)
    capi.tune("time_apply", 1)
    try:
        ms = []
        for _ in range(reps + 1):
            capi.apply(s, gate, targets, in_place=True)
            ms.append(capi.event_elapsed_ms(62, 63))
    finally:
        capi.tune("time_apply", 0)
    return min(ms[1:])


def tune_kernels(n, dtype=capi.C128, seed=11):
    """Block shapes of the 1-qubit kernels (pair groups per block) and of the
    4-qubit dense kernel (column sets per block), timed at this size."""
    rng = W.rng_for(seed)
    res = {"apply1_group": {}, "dense_k4_cw": {}}
    u1 = capi.Gate(W.gate_u3(0.3, 0.2, 0.1))
    for grp in (1, 2, 4):
        capi.tune("apply1_group", grp)
        try:
            res["apply1_group"][grp] = sum(
                _kernel_ms(n, dtype, u1, [t]) for t in sorted({0, n // 2, n - 1}))
        finally:
            capi.tune("apply1_group", 1)
    if n >= 13:
        g4 = capi.Gate(W.haar_unitary(rng, 4))
        for cw in (1, 2, 4):
            capi.tune("dense_k4_cw", cw)
            try:
                res["dense_k4_cw"][cw] = _kernel_ms(n, dtype, g4, [0, 1, 2, 3]) + \
                    _kernel_ms(n, dtype, g4, [n - 4, n - 3, n - 2, n - 1])
            finally:
                capi.tune("dense_k4_cw", 0)
    return res


def tune_batch_passes(specs, dtype=capi.C128, widths=(0, 10, 8, 6), reps=2):
    """Times one batch of (n, steps, rolls) registers with cache-blocked
    passes of each width (0 = the full tile); returns {width: seconds}."""
    out = {}
    for w in widths:
        b = capi.Batch(dtype, pass_qubits=w)
        for n, steps, rolls in specs:
            b.add(n, steps, rolls)
        b.run()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            b.run()
            best = min(best, time.perf_counter() - t0)
        out[w] = best
    return out


def autotune(n, dtype=capi.C128, path=None, batch_specs=None):
    """Measures and caches the settings for registers of n qubits."""
    fus = tune_fusion(n, dtype)
    ker = tune_kernels(n, dtype)
    widths = {k: v for k, v in fus.items() if k != "passes"}
    entry = {
        "plan": min(fus, key=fus.get),
        "fuse_k": min(widths, key=widths.get),
        "fusion_seconds": {str(k): v for k, v in fus.items()},
        "apply1_group": min(ker["apply1_group"], key=ker["apply1_group"].get),
        "apply1_group_ms": {str(k): v for k, v in ker["apply1_group"].items()},
        "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S"),
    }
    if ker["dense_k4_cw"]:
        entry["dense_k4_cw"] = min(ker["dense_k4_cw"], key=ker["dense_k4_cw"].get)
        entry["dense_k4_cw_ms"] = {str(k): v for k, v in ker["dense_k4_cw"].items()}
    if batch_specs:
        bp = tune_batch_passes(batch_specs, dtype)
        entry["pass_qubits"] = min(bp, key=bp.get)
        entry["pass_seconds"] = {str(k): v for k, v in bp.items()}
    data = load_cache(path)
    data[key(n, dtype)] = entry
    save_cache(data, path)
    return entry


def apply_kernel_settings(n, dtype=capi.C128, path=None):
    """Sets the library's block-shape tunables from the cache entry for
    (n, dtype), if there is one; returns the entry."""
    e = lookup(n, dtype, path)
    if e:
        if "apply1_group" in e:
            capi.tune("apply1_group", int(e["apply1_group"]))
        if "dense_k4_cw" in e:
            capi.tune("dense_k4_cw", int(e["dense_k4_cw"]))
    return e
