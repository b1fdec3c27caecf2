"""Shared helpers of the test suite: drive the same gate steps through the
product `doki` module, the oracle and dense textbook algebra."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# numeric contract for FP64 results (amplitudes, probabilities, densities)
ATOL = 1e-12


def bits(mask, n):
    return {q for q in range(n) if (int(mask) >> q) & 1}


def unpack_steps(z, prefix=""):
    """Gate steps stored by tests/golden/make_golden.py."""
    mats = z[prefix + "mats"]
    tg = z[prefix + "targets"] if prefix + "targets" in z else z[prefix + "tg"]
    cm = z[prefix + "cmask"] if prefix + "cmask" in z else z[prefix + "cm"]
    am = z[prefix + "amask"] if prefix + "amask" in z else z[prefix + "am"]
    n = 64
    return [(mats[i], [int(tg[i])], bits(cm[i], n), bits(am[i], n))
            for i in range(len(mats))]


def doki_run(doki, reg, steps, nthreads=1):
    for m, t, c, a in steps:
        g = doki.gate_new(len(t), np.asarray(m).tolist(), False)
        reg = doki.registry_apply(reg, g, list(t), set(c), set(a), nthreads, False)
    return reg


def doki_values(doki, reg, n, canonical=False):
    return np.array([doki.registry_get(reg, i, canonical, False)
                     for i in range(1 << n)], dtype=np.complex128)


def oracle_run(O, st, steps):
    for m, t, c, a in steps:
        st = O.apply(st, m, list(t), sorted(c), sorted(a))
    return st


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def tensor_apply(psi, m, targets, controls=(), anticontrols=()):
    """Textbook tensor contraction: the state as an n-axis tensor (axis a is
    qubit n-1-a), the gate contracted over its target axes (gate index bit b
    is targets[b]), then kept only where every control is 1 and every
    anticontrol is 0."""
    n = int(np.log2(psi.size))
    k = len(targets)
    T = psi.reshape((2,) * n)
    axes = [n - 1 - targets[b] for b in reversed(range(k))]
    moved = np.moveaxis(T, axes, list(range(k)))
    shp = moved.shape
    out = (np.asarray(m) @ moved.reshape(1 << k, -1)).reshape(shp)
    out = np.moveaxis(out, list(range(k)), axes).reshape(-1)
    idx = np.arange(psi.size)
    keep = np.ones(psi.size, dtype=bool)
    for c in controls:
        keep &= ((idx >> c) & 1) == 1
    for a in anticontrols:
        keep &= ((idx >> a) & 1) == 0
    return np.where(keep, out, psi)


def kron_one_qubit(n, g, target):
    """I(2^(n-target-1)) (x) g (x) I(2^target), as tests/one_gate_tests.py:48-61
    builds it."""
    return np.kron(np.kron(np.eye(1 << (n - target - 1)), g), np.eye(1 << target))


def reduced_density_einsum(psi, keep):
    """rho over the kept qubits (ascending qubit order = reduced index bit
    order) by tensor contraction of |psi><psi| over the others."""
    n = int(np.log2(psi.size))
    T = psi.reshape((2,) * n)
    keep_axes = [n - 1 - q for q in sorted(keep, reverse=True)]
    tr_axes = [a for a in range(n) if a not in keep_axes]
    A = np.transpose(T, keep_axes + tr_axes).reshape(1 << len(keep), -1)
    return A @ A.conj().T
