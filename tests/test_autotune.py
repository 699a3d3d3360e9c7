"""CPU checks of the launch-shape tuner's cache and selection rule."""

import json

import pytest

from magnelio_amd import autotune, kernels


def test_choose_fastest_ties_to_smaller_shape():
    shape = kernels.launch_shape
    assert autotune.choose({1: 0.30, 2: 0.21, 3: 0.25, 4: 0.22}) == 2
    assert autotune.choose({1: 0.2, 2: 0.2, 4: 0.3}) == 1
    # ties: shorter chunk, then fewer sets per thread, then the smaller band
    t = {shape(2, 3, 4): 0.2, shape(2, 2, 16): 0.2, shape(4, 1, 4): 0.2, shape(1, 3, 8): 0.3}
    assert autotune.choose(t) == shape(2, 2, 16)


def test_shape_words_and_candidates():
    w = kernels.launch_shape(3, 2, 16)
    assert kernels.shape_parts(w) == (3, 2, 16) and kernels.shape_parts(2) == (2, 0, 0)
    for bad in ((0, 0, 0), (1, 4, 0), (1, 0, 256), (70000, 0, 0)):
        with pytest.raises(ValueError):
            kernels.launch_shape(*bad)
    one, six = autotune.candidates(1), autotune.candidates(6)
    assert len(one) == len(autotune.KCHUNKS) * len(autotune.BANDS)
    assert len(six) == len(autotune.KCHUNKS) * len(autotune.SETS) * len(autotune.BANDS)
    assert {kernels.shape_parts(x)[1] for x in one} == {1}
    assert len(set(six)) == len(six)


def test_batch_candidates_and_choice():
    for n in range(1, 13):
        c = autotune.batch_candidates(n)
        assert c[0] == n and c[-1] == 1 and c == sorted(set(c), reverse=True)
        # every solver count m has its smallest largest batch in the list
        for m in range(1, n + 1):
            assert -(-n // m) in c
    assert autotune.batch_candidates(6) == [6, 3, 2, 1]
    # K = 6 in one solver: 1.0 ms a step beats two solvers of 3 (2 x 0.55)
    best, cost = autotune.choose_batch(6, {6: 1.0, 3: 0.55, 2: 0.4, 1: 0.3})
    assert best == 6 and cost == {6: 1.0, 3: 1.1, 2: pytest.approx(1.2), 1: pytest.approx(1.8)}
    assert autotune.choose_batch(6, {6: 1.3, 3: 0.5, 2: 0.4, 1: 0.3})[0] == 3
    assert autotune.choose_batch(6, {6: 1.0, 3: 0.5, 2: 0.5, 1: 0.3})[0] == 6  # tie: larger batch
    assert autotune.choose_batch(7, {7: 1.0, 4: 0.45, 3: 0.4, 2: 0.3, 1: 0.2})[0] == 4  # 2 x 0.45


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
