"""Autotuner cache handling on the CPU (no device): defaults, lookups and
round trips of the JSON cache."""
import json

from doki_amd import capi, tune


def test_defaults_without_cache(tmp_path):
    path = str(tmp_path / "none.json")
    assert tune.lookup(20, capi.C128, path) is None
    assert tune.best_fuse_k(20, capi.C128, path) == 5
    assert tune.best_fuse_k(8, capi.C128, path) == 0


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "t.json")
    data = {tune.key(24, capi.C128): {"fuse_k": 4},
            tune.key(24, capi.C64): {"fuse_k": 6}}
    tune.save_cache(data, path)
    assert json.load(open(path)) == data
    assert tune.best_fuse_k(24, capi.C128, path) == 4
    assert tune.best_fuse_k(24, capi.C64, path) == 6
    assert tune.best_fuse_k(25, capi.C128, path) == 5


def test_plan_defaults_and_entries(tmp_path):
    path = str(tmp_path / "p.json")
    assert tune.best_plan(30, capi.C128, path) == "passes"
    assert tune.best_plan(12, capi.C64, path) == "passes"
    assert tune.best_plan(11, capi.C128, path) == 0
    tune.save_cache({tune.key(20, capi.C128): {"plan": 5, "fuse_k": 5},
                     tune.key(21, capi.C128): {"plan": "passes", "fuse_k": 4}}, path)
    assert tune.best_plan(20, capi.C128, path) == 5
    assert tune.best_plan(21, capi.C128, path) == "passes"
    # an entry from before plans were tuned: the default plan
    tune.save_cache({tune.key(22, capi.C128): {"fuse_k": 4}}, path)
    assert tune.best_plan(22, capi.C128, path) == "passes"


def test_prepare_plans_circuits():
    import doki_amd
    if not doki_amd.built():
        from doki_amd import build
        build.build()
    from doki_amd import workloads as W
    steps, _ = W.random_circuit(3, 13, 50)
    c = capi.Circuit(13, steps)
    assert tune.prepare(c, "passes") > 0 and c.plan()
    assert tune.prepare(c, 4) > 0 and c.plan() == []
    assert tune.prepare(c, 0) == 0


def test_corrupt_cache_reads_empty(tmp_path):
    path = tmp_path / "bad.json"
    path.write_text("{not json")
    assert tune.load_cache(str(path)) == {}
    assert tune.best_fuse_k(30, capi.C128, str(path)) == 5


def test_env_path(tmp_path, monkeypatch):
    monkeypatch.setenv("DOKI_TUNE_CACHE", str(tmp_path / "env.json"))
    assert tune.cache_path() == str(tmp_path / "env.json")
