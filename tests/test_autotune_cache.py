"""Autotuner host logic: candidate shapes per n and the JSON cache (CPU)."""
import json

from tomos_amd import autotune


def test_candidates_cover_kernel_limits():
    for n in (1, 64, 512, 1000, 1024, 1536, 2048):
        c = autotune.candidates(n)
        assert c, n
        for g, t, b in c:
            assert n <= autotune.MAX_N[(g, t)]
    assert all(g != 8 for g, t, _ in autotune.candidates(1500))  # G=8 shapes stop at 1024
    assert autotune.candidates(2049) == []


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "device_key", lambda: "gfx950-256cu")
    path = tmp_path / "tune.json"
    c = autotune.TuneCache(str(path))
    entry = {"row_groups": 8, "threads": 512, "blocks_per_cu": 0, "ms": 0.7, "GBps": 5700.0}
    c.put(5, 1000, entry)
    c.save()
    data = json.loads(path.read_text())
    assert data == {"gfx950-256cu": {"fused/5/1000": entry}}
    c2 = autotune.TuneCache(str(path))
    assert c2.get(5, 1000) == entry
    assert c2.get(1, 1000) is None
    monkeypatch.setattr(autotune, "device_key", lambda: "gfx942-304cu")
    assert c2.get(5, 1000) is None  # entries are per device
