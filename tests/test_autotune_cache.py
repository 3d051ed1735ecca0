"""The erosion autotuner's cache entries (CPU only): lookups by step count,
entries written before plans had a step count, and the merge that keeps a
faster tuning of step counts a later search did not cover."""
import json

from nixis_amd import autotune


def _write(tmp_path, entries):
    path = str(tmp_path / "tune.json")
    with open(path, "w") as f:
        json.dump(entries, f)
    return path


def test_lookup_by_step_count(tmp_path):
    V = 10_000_002
    key = autotune.erosion_key(V)
    path = _write(tmp_path, {key: {"chunk": 896, "variant": 0, "graph_chunk": 32, "steps": 4,
                                   "us_per_step": 92.0,
                                   "steps2": {"chunk": 768, "variant": 5, "graph_chunk": 32,
                                              "steps": 2, "us_per_step": 117.0}}})
    assert autotune.lookup_erosion(V, path=path) == autotune.ErosionTuning(896, 0, 32, 4)
    assert autotune.lookup_erosion(V, path=path, steps=4) == autotune.ErosionTuning(896, 0, 32, 4)
    assert autotune.lookup_erosion(V, path=path, steps=2) == autotune.ErosionTuning(768, 5, 32, 2)
    assert autotune.lookup_erosion(V // 4, path=path) is None
    kw = autotune.lookup_erosion(V, path=path).graph_kwargs()
    assert kw == {"chunk": 896, "variant": 0, "graph_chunk": 32, "steps": 4}


def test_entry_without_step_count_is_two_step(tmp_path):
    V = 2_560_002
    path = _write(tmp_path, {autotune.erosion_key(V): {"chunk": 768, "variant": 0,
                                                       "graph_chunk": 64, "us_per_step": 30.0}})
    assert autotune.lookup_erosion(V, path=path) == autotune.ErosionTuning(768, 0, 64, 2)
    assert autotune.lookup_erosion(V, path=path, steps=2).steps == 2
    assert autotune.lookup_erosion(V, path=path, steps=4) is None


def test_merge_keeps_faster_unsearched_step_count():
    four = {"chunk": 896, "variant": 0, "graph_chunk": 32, "steps": 4, "us_per_step": 92.0}
    two = {"chunk": 768, "variant": 5, "graph_chunk": 32, "steps": 2, "us_per_step": 117.0}
    # a two-step-only search (the multi-planet form) keeps the faster four-step entry first
    m = autotune.merge_erosion_entry(four, two, {2: None})
    assert m["steps"] == 4 and m["chunk"] == 896 and m["steps2"] == two
    # a search that covered four-step plans replaces the entry
    assert autotune.merge_erosion_entry(four, two, {4: None, 2: None}) == two
    # a slower cached entry of another step count is replaced
    slow = dict(four, us_per_step=130.0)
    assert autotune.merge_erosion_entry(slow, two, {2: None}) == two
    assert autotune.merge_erosion_entry(None, two, {2: None}) == two
