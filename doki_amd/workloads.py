"""Seeded synthetic circuits for the state-vector core.

Gate matrices are built once on the host in FP64 and handed unchanged to
every engine that runs them, so the matrices themselves are never a source of
disagreement.  Conventions follow the reference's own test helpers:
U3 as ``U_np(theta, phi, lam, False)`` (tests/one_gate_tests.py:17-34),
RZ as ``RZ_np(angle, False)`` (tests/measure_tests.py:14-23),
RY as in tests/probability_tests.py:12-16, RX the standard rotation.
Every generator draws from ``np.random.default_rng(np.random.SeedSequence(seed))``
as the reference's harness does (tests/timed_test.py:62-98).

A gate step is a tuple ``(matrix, targets, controls, anticontrols)`` with the
argument meaning of ``registry_apply``: ``targets[0]`` is the least
significant bit of the matrix index, rows are output basis states.
"""
import math

import numpy as np

SQ2 = math.sqrt(2.0) / 2.0


def gate_h():
    return np.array([[SQ2, SQ2], [SQ2, -SQ2]], dtype=np.complex128)


def gate_x():
    return np.array([[0, 1], [1, 0]], dtype=np.complex128)


def gate_z():
    return np.array([[1, 0], [0, -1]], dtype=np.complex128)


def gate_rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def gate_ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def gate_rz(t):
    return np.array([[complex(math.cos(-t / 2), math.sin(-t / 2)), 0],
                     [0, complex(math.cos(t / 2), math.sin(t / 2))]],
                    dtype=np.complex128)


def gate_u3(theta, phi, lam):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array(
        [[c, complex(-s * math.cos(lam), -s * math.sin(lam))],
         [complex(s * math.cos(phi), s * math.sin(phi)),
          complex(c * math.cos(phi + lam), c * math.sin(phi + lam))]],
        dtype=np.complex128)


def rng_for(seed):
    return np.random.default_rng(np.random.SeedSequence(seed))


def random_gate_step(rng, n, p_ctrl=0.3, p_actrl=0.1):
    """One gate of the mixed workload: H, X, RX, RY, RZ or U3 on a uniform
    target; with probability p_ctrl 1-2 distinct controls, with probability
    p_actrl one anticontrol, all distinct from the target and each other."""
    kind = int(rng.integers(6))
    if kind == 0:
        m = gate_h()
    elif kind == 1:
        m = gate_x()
    elif kind in (2, 3, 4):
        t = float(rng.random() * 2 * math.pi)
        m = (gate_rx, gate_ry, gate_rz)[kind - 2](t)
    else:
        a = rng.random(3) * 2 * math.pi
        m = gate_u3(float(a[0]), float(a[1]), float(a[2]))
    target = int(rng.integers(n))
    free = [q for q in range(n) if q != target]
    controls = set()
    anticontrols = set()
    if n > 1 and rng.random() < p_ctrl:
        nc = min(len(free), 1 + int(rng.integers(2)))
        controls = set(int(q) for q in rng.choice(free, size=nc, replace=False))
    free2 = [q for q in free if q not in controls]
    if free2 and rng.random() < p_actrl:
        anticontrols = {int(free2[int(rng.integers(len(free2)))])}
    return (m, [target], controls, anticontrols)


def random_circuit(seed, n, depth):
    """The mixed random circuit: ``depth`` gates from random_gate_step, then
    the rolls for measuring every qubit (one per qubit, MSB first)."""
    rng = rng_for(seed)
    steps = [random_gate_step(rng, n) for _ in range(depth)]
    rolls = rng.random(n).tolist()
    return steps, rolls


def brickwork_layer(rng, n, layer):
    """U3 on every qubit, then CZ on alternating neighbour pairs: (0,1),(2,3)..
    on even layers and (1,2),(3,4).. on odd layers.  CZ is a Z on the upper
    qubit controlled by the lower one."""
    steps = []
    for q in range(n):
        a = rng.random(3) * 2 * math.pi
        steps.append((gate_u3(float(a[0]), float(a[1]), float(a[2])), [q],
                      set(), set()))
    z = gate_z()
    for q in range(layer % 2, n - 1, 2):
        steps.append((z, [q + 1], {q}, set()))
    return steps


def brickwork_circuit(seed, n, depth):
    rng = rng_for(seed)
    steps = []
    for layer in range(depth):
        steps += brickwork_layer(rng, n, layer)
    return steps


def haar_unitary(rng, k):
    """Haar-random 2^k x 2^k unitary: QR of a complex Gaussian matrix with the
    phases of R's diagonal moved into Q (Mezzadri, Notices AMS 54, 2007)."""
    d = 1 << k
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return np.ascontiguousarray(q * ph[None, :])


def random_normalised_state(seed, n):
    """A seeded complex-Gaussian state of unit norm (complex128)."""
    rng = rng_for(seed)
    v = rng.standard_normal(2 << n).view(np.complex128)
    v /= np.linalg.norm(v)
    return v


SERVE_OPS = ("prob", "measure", "density", "join")


def serving_qubits(rng, nmax=30):
    """Register size of the served mix: 8-16 qubits with probability 0.70,
    17-26 with 0.25, 27-30 with 0.05 (uniform inside each band); nmax caps
    the top of every band (smaller test runs)."""
    u = rng.random()
    lo, hi = (8, 16) if u < 0.70 else (17, 26) if u < 0.95 else (27, 30)
    hi = min(hi, nmax)
    lo = min(lo, hi)
    return int(rng.integers(lo, hi + 1))


def serving_requests(seed, count, qps, nmax=30, depth=(20, 200)):
    """Poisson arrivals at rate qps.  Each request holds one register (two for
    a join) of serving_qubits() qubits running a random_gate_step circuit of
    depth U[depth] gates from |0...0>, then one operation, uniform over
    SERVE_OPS: the marginal probabilities of every qubit, a full measurement
    with injected rolls, a reduced density over 1-4 random qubits, or the
    join of two registers (first = high qubits) followed by the marginal
    probabilities of the joined register.

    Returns dicts: rid, arrival (s), op, regs [(nq, steps)], rolls, keep."""
    rng = rng_for(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / qps))
        op = SERVE_OPS[int(rng.integers(len(SERVE_OPS)))]
        regs = []
        for _ in range(2 if op == "join" else 1):
            n = serving_qubits(rng, nmax)
            d = int(rng.integers(depth[0], depth[1] + 1))
            regs.append((n, [random_gate_step(rng, n) for _ in range(d)]))
        n0 = regs[0][0]
        rolls = rng.random(n0).tolist() if op == "measure" else None
        keep = ()
        if op == "density":
            a = int(rng.integers(1, 5))
            keep = tuple(sorted(int(q) for q in rng.choice(n0, size=a, replace=False)))
        out.append(dict(rid=rid, arrival=t, op=op, regs=regs, rolls=rolls, keep=keep))
    return out


def batch_specs(seed, count, nmin=4, nmax=20):
    """The batched-registers workload: qubit counts log-uniform (p ~ 1/n) on
    [nmin, nmax], each a mixed random circuit of depth U[10, 100] with rolls
    for a full measurement; members of 14 qubits also ask for the reduced
    density of qubits 0..5.  Returns (n, steps, rolls, keep) tuples."""
    rng = rng_for(seed)
    ns = np.arange(nmin, nmax + 1)
    p = 1.0 / ns
    p /= p.sum()
    specs = []
    for i in range(count):
        n = int(rng.choice(ns, p=p))
        depth = int(rng.integers(10, 101))
        steps, rolls = random_circuit(seed * 100003 + i, n, depth)
        keep = tuple(range(6)) if n == 14 else ()
        specs.append((n, steps, rolls, keep))
    return specs
