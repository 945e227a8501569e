"""The autotuner's JSON cache (no GPU): entries are keyed by CU count and
workload, survive a reload, and a damaged file reads as empty."""
from __future__ import annotations

from helio_amd import autotune


def test_store_lookup_round_trip(tmp_path):
    p = tmp_path / "tuned" / "shapes.json"
    assert autotune.lookup(256, "w", p) is None
    autotune.store(256, "w", {"wg": 512, "ipt": 4}, p)
    autotune.store(304, "w", {"wg": 256, "ipt": 2}, p)
    autotune.store(256, "hiz", {"best_tile": 32}, p)
    e = autotune.lookup(256, "w", p)
    assert (e["wg"], e["ipt"]) == (512, 4) and "tuned_at" in e
    assert autotune.lookup(304, "w", p)["wg"] == 256
    assert autotune.lookup(256, "hiz", p)["best_tile"] == 32
    autotune.store(256, "w", {"wg": 128, "ipt": 8}, p)  # a retune replaces the entry
    assert autotune.lookup(256, "w", p)["wg"] == 128
    assert autotune.key(256, "w") == "cu256:w"


def test_damaged_cache_reads_as_empty(tmp_path):
    p = tmp_path / "shapes.json"
    p.write_text("{not json")
    assert autotune.lookup(256, "w", p) is None
    autotune.store(256, "w", {"wg": 64, "ipt": 1}, p)
    assert autotune.lookup(256, "w", p)["ipt"] == 1
