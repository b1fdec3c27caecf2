"""ctypes front end of the oracle (liboracle.so, built from doki_oracle.c).

TEST INFRASTRUCTURE ONLY: the CPU restatement the HIP path is compared against.
Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module; the product never does.

Also loads the reference's own extension from ``_ref/`` (compiled from the
reference's sources by this folder's Makefile) under a private module name,
so that it can be timed or used to write golden vectors next to the product
``doki`` module in one process.
"""
import ctypes as C
import glob
import importlib.machinery
import importlib.util
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def build(ref=True):
    """Compile liboracle.so (and the reference extension when present)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint)
        L.orc_new.restype = P
        L.orc_new.argtypes = [C.c_uint, C.c_int]
        L.orc_from_data.restype = P
        L.orc_from_data.argtypes = [C.c_uint, dp]
        L.orc_clone.restype = P
        L.orc_clone.argtypes = [P]
        L.orc_free.argtypes = [P]
        L.orc_num_qubits.restype = C.c_uint
        L.orc_num_qubits.argtypes = [P]
        L.orc_norm.restype = C.c_double
        L.orc_norm.argtypes = [P]
        L.orc_read.argtypes = [P, dp]
        L.orc_read_raw.argtypes = [P, dp]
        L.orc_apply.restype = P
        L.orc_apply.argtypes = [P, C.c_uint, dp, up, up, C.c_uint, up, C.c_uint]
        L.orc_prob.restype = C.c_double
        L.orc_prob.argtypes = [P, C.c_uint]
        L.orc_collapse.restype = P
        L.orc_collapse.argtypes = [P, C.c_uint, C.c_int, C.c_double]
        L.orc_measure.restype = C.c_int
        L.orc_measure.argtypes = [P, C.c_uint64, dp, C.c_int,
                                  C.POINTER(C.c_byte), C.POINTER(P)]
        L.orc_join.restype = P
        L.orc_join.argtypes = [P, P]
        L.orc_phase.restype = C.c_double
        L.orc_phase.argtypes = [P]
        L.orc_get.argtypes = [P, C.c_int64, C.c_int, dp]
        ip = C.POINTER(C.c_int)
        L.orc_density_get.restype = C.c_int
        L.orc_density_get.argtypes = [P, ip, C.c_int, C.c_int64, C.c_int64, dp]
        L.orc_density_matrix.restype = C.c_int
        L.orc_density_matrix.argtypes = [P, ip, C.c_int, dp]
        L.orc_density_trace.restype = C.c_int
        L.orc_density_trace.argtypes = [P, ip, C.c_int, dp]
        L.orc_mem_size.restype = C.c_uint64
        L.orc_mem_size.argtypes = [C.c_uint]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _uarr(xs):
    a = (C.c_uint * max(1, len(xs)))(*xs)
    return a


class State:
    """An oracle state (raw amplitudes + norm constant), freed on GC."""

    def __init__(self, ptr):
        if not ptr:
            raise MemoryError("oracle allocation failed")
        self.ptr = C.c_void_p(ptr)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.orc_free(self.ptr)
            self.ptr = None

    @property
    def nq(self):
        return lib().orc_num_qubits(self.ptr)

    @property
    def norm(self):
        return lib().orc_norm(self.ptr)

    def values(self):
        n = self.nq
        out = np.empty(2 << n, dtype=np.float64)
        lib().orc_read(self.ptr, _dp(out))
        return out.view(np.complex128)

    def raw(self):
        n = self.nq
        out = np.empty(2 << n, dtype=np.float64)
        lib().orc_read_raw(self.ptr, _dp(out))
        return out.view(np.complex128)


def new(nq):
    return State(lib().orc_new(nq, 1))


def from_data(nq, values):
    v = np.ascontiguousarray(np.asarray(values, dtype=np.complex128).ravel())
    return State(lib().orc_from_data(nq, _dp(v.view(np.float64))))


def clone(s):
    return State(lib().orc_clone(s.ptr))


def apply(s, matrix, targets, controls=(), anticontrols=()):
    m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))
    k = len(targets)
    assert m.shape == (1 << k, 1 << k)
    mf = m.view(np.float64).ravel()
    return State(lib().orc_apply(s.ptr, k, _dp(mf), _uarr(targets),
                                 _uarr(sorted(controls)), len(controls),
                                 _uarr(sorted(anticontrols)),
                                 len(anticontrols)))


def prob(s, q):
    return lib().orc_prob(s.ptr, q)


def collapse(s, q, value, p1):
    return State(lib().orc_collapse(s.ptr, q, int(value), p1))


def measure(s, mask, rolls):
    """Returns (rc, state-or-None, list of True/False/None, qubit n-1 first)."""
    n = s.nq
    r = np.ascontiguousarray(np.asarray(rolls, dtype=np.float64))
    res = (C.c_byte * max(1, n))()
    out = C.c_void_p()
    rc = lib().orc_measure(s.ptr, mask, _dp(r) if len(r) else None, len(r),
                           res, C.byref(out))
    if rc:
        return rc, None, None
    st = State(out.value) if out.value else None
    lst = [None if res[i] < 0 else bool(res[i]) for i in range(n)]
    return 0, st, lst


def join(a, b):
    return State(lib().orc_join(a.ptr, b.ptr))


def phase(s):
    return lib().orc_phase(s.ptr)


def get(s, idx, canonical=False):
    out = np.empty(2)
    lib().orc_get(s.ptr, idx, int(canonical), _dp(out))
    return complex(out[0], out[1])


def _chain(chain):
    a = (C.c_int * max(1, len(chain)))(*chain)
    return a


def density_matrix(s, chain=()):
    dim = (1 << s.nq) >> len(chain)
    out = np.empty(2 * dim * dim)
    rc = lib().orc_density_matrix(s.ptr, _chain(chain), len(chain), _dp(out))
    if rc:
        raise ValueError(f"oracle density code {rc}")
    return out.view(np.complex128).reshape(dim, dim)


def density_get(s, chain, i, j):
    out = np.empty(2)
    rc = lib().orc_density_get(s.ptr, _chain(chain), len(chain), i, j, _dp(out))
    return rc, complex(out[0], out[1])


def density_trace(s, chain=()):
    out = np.empty(2)
    rc = lib().orc_density_trace(s.ptr, _chain(chain), len(chain), _dp(out))
    return rc, complex(out[0], out[1])


def mem_size(nq):
    return lib().orc_mem_size(nq)


def load_reference():
    """The reference's compiled extension from _ref/, or None if not built."""
    paths = sorted(glob.glob(os.path.join(HERE, "_ref", "doki*.so")))
    if not paths:
        return None
    loader = importlib.machinery.ExtensionFileLoader("doki", paths[0])
    spec = importlib.util.spec_from_file_location("doki", paths[0],
                                                  loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod
