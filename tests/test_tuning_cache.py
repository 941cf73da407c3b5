"""The shipped autotuner cache (bpm_amd/tuning_cache.json) on the CPU: every
entry names variants the C ABI accepts (rows 0-5, columns 0-9: launch
shape plus radix schedule), its best choice is the fastest entry of its own table, and lookup
falls back to the nearest cached batch of the same grid."""
import json

from bpm_amd import tuning


def test_cache_entries_are_valid_and_consistent():
    with open(tuning.CACHE) as f:
        cache = json.load(f)
    assert cache
    for key, entry in cache.items():
        gfx, grid, batch = key.split(":")
        assert gfx == "gfx950"
        nx, ny = (int(v) for v in grid.split("x"))
        assert nx >= 16 and ny >= 16 and nx & (nx - 1) == 0 and ny & (ny - 1) == 0
        assert int(batch) >= 1
        rv, cv = entry["best"]
        assert 0 <= rv <= 5 and 0 <= cv <= 9, key
        table = entry["table"]
        assert "%d,%d" % (rv, cv) in table, key
        fastest = min(table, key=lambda k: table[k]["ms"])
        assert fastest == "%d,%d" % (rv, cv), key
        for k, v in table.items():
            r, c = (int(x) for x in k.split(","))
            assert 0 <= r <= 5 and 0 <= c <= 9 and v["ms"] > 0, (key, k)


def test_lookup_uses_nearest_cached_batch(monkeypatch):
    fake = {
        "gfx950:1024x1024:1": {"best": [0, 1], "table": {}},
        "gfx950:1024x1024:16": {"best": [0, 3], "table": {}},
        "gfx950:512x512:1": {"best": [1, 0], "table": {}},
    }
    monkeypatch.setattr(tuning, "_cache", fake)
    assert tuning.lookup(1024, 1024, 1) == (0, 1)
    assert tuning.lookup(1024, 1024, 16) == (0, 3)
    assert tuning.lookup(1024, 1024, 12) == (0, 3)
    assert tuning.lookup(1024, 1024, 3) == (0, 1)
    assert tuning.lookup(512, 512, 64) == (1, 0)
    assert tuning.lookup(256, 256, 1) is None
