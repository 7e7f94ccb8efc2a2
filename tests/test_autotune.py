"""Autotuner bookkeeping on the CPU: cache keys, the batch bucket, the
user database round trip and the shipped table's format (the timing itself
runs on the GPU, test_gpu_autotune.py)."""

import json
import os

import pytest


def test_keys_and_buckets():
    from jaqsi_amd import autotune

    assert [autotune.batch_bucket(b) for b in (1, 2, 3, 4, 5, 255, 256, 257)] == [1, 2, 4, 4, 8, 256, 256, 512]
    k1 = autotune.entry_key("gfx950", ("sv", 3, ("op", (0,), 0)), 0, 3)
    assert k1.startswith("gfx950|c64|b4|")
    assert k1 != autotune.entry_key("gfx950", ("sv", 3, ("op", (1,), 0)), 0, 3)
    assert k1 != autotune.entry_key("gfx950", ("sv", 3, ("op", (0,), 0)), 1, 3)


def test_database_round_trip(tmp_path, monkeypatch):
    from jaqsi_amd import autotune

    monkeypatch.setenv("JAQSI_AMD_TUNE_DB", str(tmp_path / "db.json"))
    monkeypatch.setattr(autotune, "_db", None)
    autotune._store("gfx950|c64|b1|abc", {"low_bits": 4, "variant_flags": 2, "ms": 1.0})
    monkeypatch.setattr(autotune, "_db", None)
    assert json.load(open(tmp_path / "db.json"))["gfx950|c64|b1|abc"]["low_bits"] == 4
    assert autotune._load_db()["gfx950|c64|b1|abc"]["variant_flags"] == 2


def test_modes():
    from jaqsi_amd import autotune

    prev = autotune.get_mode()
    try:
        autotune.set_mode("off")
        with pytest.raises(ValueError):
            autotune.set_mode("sometimes")
    finally:
        autotune.set_mode(prev)


def test_shipped_tables_are_well_formed():
    from jaqsi_amd import _native as nat

    d = os.path.join(os.path.dirname(nat.__file__), "tuning")
    for fn in os.listdir(d) if os.path.isdir(d) else []:
        if not fn.endswith(".json"):
            continue
        table = json.load(open(os.path.join(d, fn)))
        for k, rec in table.items():
            parts = k.split("|")
            arch, dt, b, h = parts[:4]
            assert arch == fn[:-5] and dt in ("c64", "c128") and b.startswith("b") and len(h) == 20
            assert parts[4:] in ([], ["fold"])  # plans measured through their observables' folded tail
            assert rec["low_bits"] in (0, 3, 4, 5, 6) and rec["variant_flags"] & ~15 == 0
            assert rec.get("tile_bits", 0) in (0, 13, 14)
