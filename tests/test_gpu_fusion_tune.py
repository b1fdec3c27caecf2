"""The autotuner on the device: it measures, caches and its choices run
circuits with the same results as the oracle."""
import numpy as np
import pytest

from doki_amd import workloads as W
from helpers import ATOL, oracle_run

pytestmark = pytest.mark.gpu


def test_autotune_and_use(capi, oracle, tmp_path):
    from doki_amd import tune
    path = str(tmp_path / "tune.json")
    n = 14
    specs = [(10, *W.random_circuit(60 + i, 10, 30)) for i in range(8)]
    entry = tune.autotune(n, capi.C128, path=path, batch_specs=specs)
    assert entry["fuse_k"] in tune.FUSE_WIDTHS
    assert entry["plan"] in tune.PLANS
    assert entry["apply1_group"] in (1, 2, 4)
    assert entry["dense_k4_cw"] in (1, 2, 4)
    assert entry["pass_qubits"] in (0, 10, 8, 6)
    assert tune.lookup(n, capi.C128, path)["fuse_k"] == entry["fuse_k"]
    try:
        tune.apply_kernel_settings(n, capi.C128, path)
        steps, _ = W.random_circuit(99, n, 60)
        for plan in (tune.best_fuse_k(n, capi.C128, path),
                     tune.best_plan(n, capi.C128, path)):
            c = capi.Circuit(n, steps)
            tune.prepare(c, plan)
            st = c.run(capi.new(n))
            want = oracle_run(oracle, oracle.new(n), steps)
            assert np.max(np.abs(st.values() - want.values())) <= ATOL
    finally:
        capi.tune("apply1_group", 1)
        capi.tune("dense_k4_cw", 0)
