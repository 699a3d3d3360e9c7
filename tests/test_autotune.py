"""CPU checks of the launch-shape tuner's cache and selection rule."""

import json

from magnelio_amd import autotune


def test_choose_fastest_ties_to_smaller_chunk():
    assert autotune.choose({1: 0.30, 2: 0.21, 3: 0.25, 4: 0.22}) == 2
    assert autotune.choose({1: 0.2, 2: 0.2, 4: 0.3}) == 1


def test_cache_roundtrip_and_bad_files(tmp_path):
    path = tmp_path / "sub" / "autotune.json"
    assert autotune._load(path) == {}  # missing
    autotune._store(path, {"a": {"kchunk": 2}})
    assert autotune._load(path) == {"a": {"kchunk": 2}}
    assert not [p for p in path.parent.iterdir() if p.name.startswith(".autotune-")]
    path.write_text("{not json")
    assert autotune._load(path) == {}  # unreadable: re-tune, do not fail
    path.write_text(json.dumps([1, 2]))
    assert autotune._load(path) == {}


def test_cache_path_from_environment(tmp_path, monkeypatch):
    monkeypatch.setenv("MAGNELIO_HIP_AUTOTUNE_CACHE", str(tmp_path / "x.json"))
    assert autotune.cache_path() == tmp_path / "x.json"
