"""Autotuner host logic (CPU): class keys, cache round trip, server table."""
import importlib
import os


def test_cache_and_tables(tmp_path):
    at = importlib.import_module("4dzb_amd.autotune")
    assert at.default_layout(32) in at.CANDIDATES and at.default_layout(64) in at.CANDIDATES
    assert at.default_layout(32) == (16, 8, 4) and at.default_layout(33) == (16, 8, 8)
    assert len(at.CANDIDATES) == len(at.BINS) * len(at.TILES) * len(at.WAVES)
    cache = {at.class_key(0, 512, 512, 32): dict(bins=32, tile=16, waves=2, order=at.ORDER_AREA, ms=1.5, default_ms=2.0,
                                                 candidates={}),
             at.class_key(2, 1920, 1080, 16): dict(bins=8, tile=8, ms=3.0, default_ms=3.1, candidates={})}
    p = str(tmp_path / "tune.json")
    at.save(cache, p)
    assert at.load(p) == cache
    assert at.load(str(tmp_path / "missing.json")) == {}
    assert at.layout(cache, 0, 512, 512, 32) == (32, 16, 2)
    assert at.layout(cache, 2, 1920, 1080, 16) == (8, 8, 4)   # an entry without waves: 4
    assert at.layout(cache, 0, 256, 256, 32) == (16, 8, 4)   # unseen class: default layout
    assert at.layout(cache, 0, 256, 256, 64) == (16, 8, 8)
    assert at.order(cache, 0, 512, 512, 32) == at.ORDER_AREA
    assert at.order(cache, 2, 1920, 1080, 16) == 0   # an entry without an order: automatic
    assert at.order(cache, 0, 256, 256, 32) == 0
    tab = sorted(at.server_table(cache), key=lambda t: t["scene"])
    assert tab[0] == dict(scene=0, w=512, h=512, res=32, bins=32, tile=16, waves=2, order=at.ORDER_AREA, ms=1.5)
    assert tab[1]["order"] == 0
    assert tab[1]["w"] == 1920 and tab[1]["h"] == 1080 and tab[1]["bins"] == 8
    assert not os.path.exists(p + ".tmp")
