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


def test_installed_entries_are_per_format(capi):
    """Setting or erasing one (a_type, n) entry leaves the others alone."""
    from tomos_amd import _capi as C
    C.call("tomos_fused_config_clear")
    try:
        C.call("tomos_fused_config_set", C.FP32_BF16, 1000, 8, 0, 512)
        C.call("tomos_fused_config_set", C.F32, 1000, 4, 2, 256)
        assert autotune.installed(C.FP32_BF16, 1000) == {"row_groups": 8, "blocks_per_cu": 0, "threads": 512}
        C.call("tomos_fused_config_set", C.F32, 1000, 0, 0, 0)  # erase one entry
        assert autotune.installed(C.F32, 1000) is None
        assert autotune.installed(C.FP32_BF16, 1000)["threads"] == 512
        assert autotune.installed(C.FP32_BF16, 999) is None
    finally:
        C.call("tomos_fused_config_clear")
    assert autotune.installed(C.FP32_BF16, 1000) is None
