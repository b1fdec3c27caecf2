"""DLPack interchange of registry values (registry_to_dlpack,
registry_from_dlpack): host tensors from numpy's own DLPack producer, device
tensors through the module's own exports, and the errors for tensors the
module cannot take.

A registry made from a tensor holds the values as given (not normalised),
as registry_new_data does (doki.c:625-720); an exported tensor holds the
values registry_get returns, bit for bit (each raw component divided by the
norm, qstate.h:86), in a copy on the device."""
import numpy as np
import pytest

from doki_amd import workloads as W

pytestmark = pytest.mark.gpu


def random_values(seed, n, normalise=False):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v) if normalise else v


def run_circuit(doki, reg, seed, n, depth):
    steps, _ = W.random_circuit(seed, n, depth)
    for m, t, c, a in steps:
        g = doki.gate_new(len(t), np.asarray(m).tolist(), False)
        reg = doki.registry_apply(reg, g, list(t), set(c), set(a), -1, False)
    return reg


def test_from_numpy_host_tensors(doki):
    """numpy arrays through their own __dlpack__ (host memory): complex128
    (n) and float64 (n, 2); values as given, unnormalised, like
    registry_new_data."""
    v = random_values(1, 9) * 3.0
    r1 = doki.registry_from_dlpack(v, False)
    assert np.array_equal(doki.registry_get_all(r1, False, False), v)
    r2 = doki.registry_from_dlpack(np.ascontiguousarray(v.view(np.float64).reshape(-1, 2)),
                                   False)
    assert np.array_equal(doki.registry_get_all(r2, False, False), v)
    r3 = doki.registry_new_data(9, v, False)
    for q in range(9):
        assert doki.registry_prob(r1, q, -1, False) == doki.registry_prob(r3, q, -1, False)
    assert doki.registry_mem(r1, False) == 64 + 16 * (1 << 9)


def test_export_values_bit_for_bit(doki):
    """An exported registry (after a circuit: its norm is not 1) read back
    through a registry made from the export equals registry_get_all bit for
    bit; the export is a copy, so it outlives registry_del of its source."""
    n = 12
    reg = run_circuit(doki, doki.registry_new(n, False), 5, n, 60)
    want = doki.registry_get_all(reg, False, False)
    cap = doki.registry_to_dlpack(reg, False)
    doki.registry_del(reg, False)
    back = doki.registry_from_dlpack(cap, False)
    got = doki.registry_get_all(back, False, False)
    assert np.array_equal(got, want)
    # the capsule was consumed
    with pytest.raises(Exception, match="DLPack capsule"):
        doki.registry_from_dlpack(cap, False)


def test_export_of_unnormalised_and_pending_registries(doki):
    """Exports read a registry's values: an unnormalised registry_new_data
    state exports its values as given, and a pending chain runs first."""
    v = random_values(2, 10) * 0.25
    r = doki.registry_new_data(10, v, False)
    got = doki.registry_get_all(doki.registry_from_dlpack(doki.registry_to_dlpack(r, False),
                                                          False), False, False)
    assert np.array_equal(got, v)
    prev = doki.execution_mode("fuse", False)
    try:
        p = run_circuit(doki, doki.registry_new(14, False), 6, 14, 40)   # pending
        cap = doki.registry_to_dlpack(p, False)
        assert np.array_equal(
            doki.registry_get_all(doki.registry_from_dlpack(cap, False), False, False),
            doki.registry_get_all(p, False, False))
    finally:
        doki.execution_mode(prev, False)


def test_unconsumed_export_is_released(doki, capi):
    """A capsule nobody consumes gives its device copy back when it goes."""
    import gc
    import ctypes as C
    reg = doki.registry_new(20, False)
    used = C.c_uint64()
    gc.collect()
    capi.check(capi.lib().dk_memory_budget(None, C.byref(used)))
    before = used.value
    cap = doki.registry_to_dlpack(reg, False)
    capi.check(capi.lib().dk_memory_budget(None, C.byref(used)))
    # the copy's amplitudes (and its scalars)
    assert 16 * (1 << 20) <= used.value - before <= 16 * (1 << 20) + 4096
    del cap
    gc.collect()
    capi.check(capi.lib().dk_memory_budget(None, C.byref(used)))
    assert used.value == before


def test_bad_tensors(doki):
    err = Exception
    with pytest.raises(err):
        doki.registry_from_dlpack(np.zeros(12, dtype=np.complex128), False)   # not 2^n
    with pytest.raises(err):
        doki.registry_from_dlpack(np.zeros(8, dtype=np.complex64), False)     # complex64
    with pytest.raises(err):
        doki.registry_from_dlpack(np.zeros(16, dtype=np.complex128)[::2], False)  # strided
    with pytest.raises(err):
        doki.registry_from_dlpack(np.zeros((8, 3)), False)                    # (n, 3)
    with pytest.raises(err):
        doki.registry_from_dlpack(np.zeros(1, dtype=np.complex128), False)    # 2^0
    with pytest.raises(err):
        doki.registry_from_dlpack([1, 2, 3, 4], False)                        # no __dlpack__
    with pytest.raises(err):
        doki.registry_to_dlpack(None, False)
