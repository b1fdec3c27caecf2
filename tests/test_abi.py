"""CPU checks of the drop-in boundary: the C ABI library exports what the
header declares, the `doki` extension has the reference's method table and
exception, and argument validation that precedes any device work matches the
reference's error texts.  No GPU computation happens here."""
import json
import subprocess

import numpy as np
import pytest

import doki_amd
from helpers import GOLDEN

# the reference's method table, src/doki/doki.c:114-192
REFERENCE_METHODS = [
    "gate_new", "gate_get", "registry_new", "registry_new_data",
    "registry_clone", "registry_del", "registry_get", "registry_apply",
    "registry_join", "registry_measure", "registry_prob", "registry_density",
    "registry_mem", "funmatrix_create", "funmatrix_identity",
    "funmatrix_statezero", "funmatrix_densityzero", "funmatrix_hadamard",
    "funmatrix_addcontrol", "funmatrix_get", "funmatrix_add", "funmatrix_sub",
    "funmatrix_scalar_mul", "funmatrix_scalar_div", "funmatrix_matmul",
    "funmatrix_ewmul", "funmatrix_kron", "funmatrix_eyekron",
    "funmatrix_transpose", "funmatrix_dagger", "funmatrix_projection",
    "funmatrix_shape", "funmatrix_partialtrace", "funmatrix_trace",
    "funmatrix_apply", "funmatrix_mem",
]


@pytest.fixture(scope="module")
def built():
    if not doki_amd.built():
        from doki_amd import build
        build.build()
    return True


def test_library_exports_every_header_symbol(built):
    from doki_amd import capi
    names = capi.header_functions()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", doki_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = capi.lib()
    for n in names:
        assert getattr(lib, n) is not None


def test_extension_method_table(built):
    d = doki_amd.doki
    for name in REFERENCE_METHODS:
        assert callable(getattr(d, name)), name
    e = d.error
    assert e.__module__ + "." + e.__name__ == "qsimov.doki.error"
    assert issubclass(e, Exception)


def test_extension_links_device_library(built):
    out = subprocess.run(["ldd", doki_amd.EXT_PATH], capture_output=True,
                         text=True, check=True).stdout
    assert "libdoki_hip.so" in out and "not found" not in out


ERRORS = json.load(open(f"{GOLDEN}/errors.json"))


def _err(fn):
    d = doki_amd.doki
    with pytest.raises(d.error) as ei:
        fn()
    return str(ei.value)


@pytest.mark.parametrize("case,call", [
    ("registry_new_zero", lambda d: d.registry_new(0, False)),
    ("registry_new_too_many", lambda d: d.registry_new(63, False)),
    ("gate_new_zero", lambda d: d.gate_new(0, [[1]], False)),
    ("gate_new_not_list", lambda d: d.gate_new(1, 5, False)),
    ("gate_new_wrong_size", lambda d: d.gate_new(1, [[1, 0]], False)),
    ("gate_new_bad_row", lambda d: d.gate_new(1, [[1, 0], [0]], False)),
    ("gate_new_bad_elem", lambda d: d.gate_new(1, [[1, 0], [0, "a"]], False)),
    ("new_data_zero", lambda d: d.registry_new_data(0, [1], False)),
    ("new_data_list_size", lambda d: d.registry_new_data(2, [1, 0], False)),
    ("new_data_array_size", lambda d: d.registry_new_data(2, np.zeros(3), False)),
    ("new_data_type", lambda d: d.registry_new_data(1, (1, 0), False)),
    ("new_data_not_number",
     lambda d: d.registry_new_data(1, np.array(["a", "b"]), False)),
    ("get_not_capsule", lambda d: d.registry_get(5, 0, False, False)),
])
def test_host_side_errors_match_reference(built, case, call):
    want = ERRORS[case]
    assert want[0] == "qsimov.doki.error"
    assert _err(lambda: call(doki_amd.doki)) == want[1]


def test_gate_round_trip_on_host(built):
    d = doki_amd.doki
    m = [[0.5, 1j], [2, -3.25 + 1e-3j]]
    g = d.gate_new(1, m, False)
    assert d.gate_get(g, False) == [[complex(0.5), 1j], [complex(2), -3.25 + 1e-3j]]


def test_no_device_fails_loudly(built):
    """Without a gfx950 device the product raises; it never computes on the
    CPU."""
    from doki_amd import capi
    if capi.lib().dk_init(0) == 0:
        pytest.skip("a device is present")
    d = doki_amd.doki
    with pytest.raises(d.error):
        d.registry_new(3, False)


def test_bulk_read_entry_validates_on_host(built):
    """registry_get_all (bulk form of registry_get, not in the reference's
    table) checks its arguments before any device work."""
    d = doki_amd.doki
    assert callable(d.registry_get_all)
    assert _err(lambda: d.registry_get_all(5, False, False)) == "NULL pointer to registry"
    assert _err(lambda: d.registry_get_all(5, False)) == (
        "Syntax: registry_get_all(registry, canonical, verbose)")


def test_batch_host_validation(built):
    """The batched engine checks every register on the host, before any
    device work: gate targets and masks (registry_apply's rules), density
    size, rolls in [0, 1), settings; without a device its run fails loudly."""
    from doki_amd import capi
    from doki_amd import workloads as W
    b = capi.Batch(capi.C128, split_qubits=13, pass_qubits=3, use_graph=1)
    steps, rolls = W.random_circuit(3, 6, 20)
    assert b.add(6, steps, rolls, (0, 1)) == 0
    assert b.add(18, [(W.gate_h(), [17], {0}, {5})]) == 1
    h = W.gate_h()
    bad = [
        (4, [(h, [4], set(), set())], None, ()),        # target outside
        (4, [(h, [1], {1}, set())], None, ()),          # control on target
        (4, [(h, [1], {2}, {2})], None, ()),            # control = anticontrol
        (4, [(h, [1], {7}, set())], None, ()),          # control outside
        (8, [], None, tuple(range(7))),                  # density of 7 qubits
        (3, [], [0.5, 1.0, 0.1], ()),                    # roll not below 1
        (27, [], None, ()),                              # too many qubits
    ]
    for nq, st, r, keep in bad:
        with pytest.raises(capi.DokiHipError):
            b.add(nq, st, r, keep)
    with pytest.raises(capi.DokiHipError):
        capi.check(capi.lib().dk_batch_set(b.ptr, b"no_such_setting", 1))
    if capi.lib().dk_init(0) != 0:
        with pytest.raises(capi.DokiHipError):
            b.run()
