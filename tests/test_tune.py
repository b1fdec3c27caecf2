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


def test_corrupt_cache_reads_empty(tmp_path):
    path = tmp_path / "bad.json"
    path.write_text("{not json")
    assert tune.load_cache(str(path)) == {}
    assert tune.best_fuse_k(30, capi.C128, str(path)) == 5


def test_env_path(tmp_path, monkeypatch):
    monkeypatch.setenv("DOKI_TUNE_CACHE", str(tmp_path / "env.json"))
    assert tune.cache_path() == str(tmp_path / "env.json")
