"""Autotuner host logic (CPU): class keys, cache round trip, server table."""
import importlib
import os


def test_cache_and_tables(tmp_path):
    at = importlib.import_module("4dzb_amd.autotune")
    assert at.default_layout(32) in at.CANDIDATES and at.default_layout(64) in at.CANDIDATES
    assert at.default_layout(32) == (16, 8, 4) and at.default_layout(33) == (16, 8, 8)
    # frames of up to 8192 work items default to 8 waves, larger ones with 32-bit rows to 4
    assert at.default_layout(32, 512, 512) == (16, 8, 8) and at.default_layout(32, 1024, 1024) == (16, 8, 4)
    assert at.work_items(704, 704) <= at.WAVES8_ITEMS < at.work_items(720, 720)
    assert at.default_waves(16, 704, 704) == 8 and at.default_waves(16, 720, 720) == 4
    assert at.default_waves(64, 4096, 4096) == 8
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
    assert at.layout(cache, 0, 256, 256, 32) == (16, 8, 8)   # unseen class: default layout (a small frame)
    assert at.layout(cache, 0, 1024, 1024, 32) == (16, 8, 4)
    assert at.layout(cache, 0, 256, 256, 64) == (16, 8, 8)
    assert at.order(cache, 0, 512, 512, 32) == at.ORDER_AREA
    assert at.order(cache, 2, 1920, 1080, 16) == 0   # an entry without an order: automatic
    assert at.order(cache, 0, 256, 256, 32) == 0
    tab = sorted(at.server_table(cache), key=lambda t: t["scene"])
    assert tab[0] == dict(scene=0, w=512, h=512, res=32, bins=32, tile=16, waves=2, order=at.ORDER_AREA, split=0,
                          ms=1.5)
    assert tab[1]["order"] == 0
    assert tab[1]["w"] == 1920 and tab[1]["h"] == 1080 and tab[1]["bins"] == 8
    assert not os.path.exists(p + ".tmp")


class _FakeCtx:
    """Stands in for fd4d.Context.autotune: layout candidates get fixed
    times; walk-order pairs get the list / area times of the instance."""

    def __init__(self, layout_ms, order_ms, split_ms=None):
        self.layout_ms = layout_ms
        self.order_ms = order_ms
        self.split_ms = split_ms or {}
        self.order_calls = []
        self.split_calls = []

    def autotune(self, item, cands, reps=3):
        if len(cands[0]) == 5:   # (bins, tile, waves, order, split)
            self.split_calls.append(tuple(cands))
            ms = [self.split_ms[c[4]] for c in cands]
        elif len(cands[0]) == 4:   # (bins, tile, waves, order) per order
            self.order_calls.append((item["id"], tuple(cands)))
            ms = list(self.order_ms[item["id"]])
        else:
            ms = [self.layout_ms.get(tuple(c), -1.0) for c in cands]
        ok = [m for m in ms if m >= 0]
        return (ms.index(min(ok)) if ok else -1), ms


class _Light:
    res_w = 16


def _items(n, w=512):
    return [dict(id=k, w=w, h=w, light=_Light()) for k in range(n)]


def test_order_choice_over_class_instances():
    at = importlib.import_module("4dzb_amd.autotune")
    cands = [(16, 8, 4), (8, 8, 4)]
    layout_ms = {(16, 8, 4): 2.0, (8, 8, 4): 1.5}
    # (list, area, mesh) per instance: area fastest on the first instances,
    # slower in sum; the class takes the fastest order in sum (mesh)
    order_ms = {0: (1.0, 0.8, 0.9), 1: (1.0, 0.8, 0.9), 2: (1.0, 0.8, 0.9), 3: (1.0, 2.0, 0.95)}
    ctx = _FakeCtx(layout_ms, order_ms)
    cache = at.tune_classes(ctx, 0, _items(4), candidates=cands, orders=True)
    e = cache[at.class_key(0, 512, 512, 16)]
    assert (e["bins"], e["tile"], e["waves"]) == (8, 8, 4)
    assert e["order_ms"] == {"list": 4.0, "area": 4.4, "mesh": 3.65, "instances": 4} and e["order"] == at.ORDER_MESH
    assert e["ms"] == 0.9   # the chosen order's time on the class's representative
    # every order triple is timed at the chosen layout
    assert all(c == tuple((8, 8, 4, o) for o in at.ORDERS) for _, c in ctx.order_calls)
    for times, want in (((1.0, 0.97, 0.99), at.ORDER_AREA), ((0.9, 0.97, 0.99), at.ORDER_LIST)):
        ctx = _FakeCtx(layout_ms, {k: times for k in range(12)})
        e = at.tune_one(ctx, _items(1)[0], cands, order_items=_items(12)[:at.ORDER_SAMPLES])
        assert e["order"] == want and e["order_ms"]["instances"] == at.ORDER_SAMPLES
    # at most ORDER_SAMPLES instances per class decide; without orders=True, no order pairs run
    ctx = _FakeCtx(layout_ms, {k: (1.0, 0.5, 0.7) for k in range(12)})
    at.tune_classes(ctx, 0, _items(12), candidates=cands, orders=True)
    assert sorted(i for i, _ in ctx.order_calls) == list(range(at.ORDER_SAMPLES))
    ctx = _FakeCtx(layout_ms, {})
    cache = at.tune_classes(ctx, 0, _items(3), candidates=cands)
    assert not ctx.order_calls and cache[at.class_key(0, 512, 512, 16)]["order"] == 0


def test_split_choice():
    at = importlib.import_module("4dzb_amd.autotune")
    cands = [(16, 8, 4), (8, 8, 4)]
    layout_ms = {(16, 8, 4): 2.0, (8, 8, 4): 1.5}
    for two, want in ((0.96, 2), (0.98, 0)):   # split 2 vs the automatic rule (0) at 1.0 ms
        split_ms = {0: 1.0, 1: 1.3, 2: two, 4: 1.01, 8: 1.2, 16: -1.0}
        ctx = _FakeCtx(layout_ms, {0: (1.0, 2.0, 1.5)}, split_ms)
        e = at.tune_one(ctx, _items(1)[0], cands, order_items=_items(1), splits=at.SPLITS)
        assert e["order"] == at.ORDER_LIST and e["split"] == want
        assert e["ms"] == (two if want else 1.0)   # a kept split's time; else the order pair's
        # the splits are timed at the chosen layout and order, the automatic rule first
        (call,) = ctx.split_calls
        assert call[0] == (8, 8, 4, at.ORDER_LIST, 0) and [c[4] for c in call[1:]] == list(at.SPLITS)
        assert e["split_ms"]["16"] is None   # not exact: never chosen
    cache = {at.class_key(0, 64, 64, 8): e}
    assert at.split(cache, 0, 64, 64, 8) == 0 and at.split(cache, 0, 32, 32, 8) == 0
    cache[at.class_key(0, 64, 64, 8)] = dict(e, split=4)
    assert at.split(cache, 0, 64, 64, 8) == 4 and at.server_table(cache)[0]["split"] == 4
