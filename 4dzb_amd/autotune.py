"""Autotuner of the batched and served forms.

For one instance class -- scene, receiver size, light grid -- it runs a
representative instance under every candidate layout (bins per light-plane
side x receiver tile width x waves per occlusion work item) through the
library's fd_autotune, which times each on the device and compares every
output bit with the default layout (16x16 bins, 8x8 tiles, the library's
default waves).  The fastest bit-exact layout wins.  Optionally (orders=True)
the walk orders of the bins' triangle lists -- the reference's, largest
projected triangles first, and the mesh order (largest first except the
very largest, last) -- are then timed at that layout on up to ORDER_SAMPLES
instances of the class, and the fastest in sum is kept (an order's effect
changes with the light pose by up to 60%, so one instance is not enough to
decide a class).  Optionally (splits=True) the number of stretches per bin
triangle list is then timed at that layout and order -- the library's
automatic rule against fixed counts -- and a fixed count is kept when it is
at least 3% faster.  The bin grid
also sets the receiver sort's key width (one counting-sort pass over
2*log2(bins) + 1 bits), so picking the grid picks the sort's radix width.

Results persist in a JSON cache keyed by class; a cache entry records every
candidate's time (or "not exact"), so a later run on the same device and
scene can reuse it.
"""
from __future__ import annotations

import json
import os
import time

BINS = (8, 16, 32)
TILES = (8, 16, 32)
WAVES = (4, 8)
CANDIDATES = [(b, t, w) for w in WAVES for b in BINS for t in TILES]
ORDER_LIST, ORDER_AREA, ORDER_MESH = 1, 2, 3   # FD_TRI_ORDER_LIST / _AREA / _MESH
ORDERS = (ORDER_LIST, ORDER_AREA, ORDER_MESH)
ORDER_SAMPLES = 8               # instances of a class that decide its walk order
SPLITS = (1, 2, 4, 8, 16)       # stretches per bin list tried against the automatic rule (0)


WAVES8_ITEMS = 8192   # FD_WAVES8_ITEMS: frames of up to this many work items default to 8 waves


def work_items(w, h, bins=16):
    """Occlusion work items of a w x h frame (64 receivers of a bin each, at
    most: the library's chunk slots)."""
    return (w * h + 63) // 64 + bins * bins + 1


def default_waves(res, w=None, h=None, bins=16):
    """The library's default waves per occlusion work item for a res x res
    light grid (fd_set_occlusion_waves): 8 for 64-bit rows and for frames of
    up to WAVES8_ITEMS work items, else 4 (a frame of unknown size counts
    as large)."""
    if res > 32:
        return 8
    return 8 if w is not None and h is not None and work_items(w, h, bins) <= WAVES8_ITEMS else 4


def default_layout(res, w=None, h=None):
    return (16, 8, default_waves(res, w, h))


def class_key(scene, w, h, res):
    return f"{scene}:{w}x{h}:{res}"


def tune_one(ctx, item, candidates=CANDIDATES, reps=3, order_items=None, splits=None):
    """Tune one instance class on `ctx`'s scene.  item: a Batch item (pos,
    nrm, w, h, light, shadow).  order_items: instances of the class on which
    to time the two walk orders at the chosen layout (None: no order
    choice, the automatic one); the entry's ms is then the first of them
    under the chosen order.  splits: stretch counts to time against the
    automatic rule on item (None: the automatic rule).  Returns the cache
    entry."""
    t0 = time.perf_counter()
    best, ms = ctx.autotune(item, candidates, reps)
    dl = default_layout(item["light"].res_w, item["w"], item["h"])
    default = ms[candidates.index(dl)] if dl in candidates else None
    b, t, w = candidates[best][:3] if best >= 0 else dl   # nothing bit-exact besides the reference: the default
    best_ms = ms[best] if best >= 0 else default
    # the walk order at that layout, summed over the given instances
    order, order_ms = 0, {}
    if order_items:
        tot = [0.0] * len(ORDERS)
        first = None
        for it in order_items:
            _, oms = ctx.autotune(it, [(b, t, w, o) for o in ORDERS], reps)
            if min(oms) < 0:   # not exact or no fit (cannot happen: the order changes no bit)
                tot = None
                break
            first = first or oms
            tot = [x + m for x, m in zip(tot, oms)]
        names = {ORDER_LIST: "list", ORDER_AREA: "area", ORDER_MESH: "mesh"}
        order_ms = {names[o]: (round(x, 4) if tot else None) for o, x in zip(ORDERS, tot or [0.0] * len(ORDERS))}
        order_ms["instances"] = len(order_items)
        if tot:
            k = min(range(len(ORDERS)), key=lambda i: tot[i])
            order = ORDERS[k]
            best_ms = first[k]
    # stretches per bin list at that layout and order
    split, split_ms = 0, {}
    if splits:
        cands = [(b, t, w, order, 0)] + [(b, t, w, order, sp) for sp in splits]
        sb, sms = ctx.autotune(item, cands, reps)
        split_ms = {str(c[4]): (round(m, 4) if m >= 0 else None) for c, m in zip(cands, sms)}
        if sb > 0 and sms[0] > 0 and sms[sb] < 0.97 * sms[0]:
            split = cands[sb][4]
            best_ms = sms[sb]
    return dict(bins=b, tile=t, waves=w, order=order, split=split, ms=best_ms, default_ms=default,
                candidates={"x".join(str(x) for x in c[:3]): (round(m, 4) if m >= 0 else ("not exact" if m == -1 else "pair arena too small"))
                            for c, m in zip(candidates, ms)},
                order_ms=order_ms, split_ms=split_ms,
                tune_s=round(time.perf_counter() - t0, 3))


def tune_classes(ctx, scene, items, cache=None, candidates=CANDIDATES, reps=3, orders=False, splits=False):
    """items: Batch items; one representative per class (scene, w, h, res) is
    tuned unless the cache already holds it (with orders, up to
    ORDER_SAMPLES of the class's items decide its walk order).  Returns the
    cache (dict)."""
    cache = {} if cache is None else cache
    groups = {}
    for it in items:
        groups.setdefault(class_key(scene, it["w"], it["h"], it["light"].res_w), []).append(it)
    for key, its in groups.items():
        if key in cache:
            continue
        cache[key] = tune_one(ctx, its[0], candidates, reps, order_items=its[:ORDER_SAMPLES] if orders else None,
                              splits=SPLITS if splits else None)
    return cache


def layout(cache, scene, w, h, res):
    """(bins, tile, waves) of a class, or the default layout."""
    e = cache.get(class_key(scene, w, h, res))
    return (e["bins"], e["tile"], e.get("waves", default_waves(res, w, h, e["bins"]))) if e else default_layout(res, w, h)


def order(cache, scene, w, h, res):
    """Walk order of a class (ORDER_LIST / _AREA / _MESH), 0 = automatic."""
    e = cache.get(class_key(scene, w, h, res))
    return e.get("order", 0) if e else 0


def split(cache, scene, w, h, res):
    """Stretches per bin list of a class, 0 = the automatic rule."""
    e = cache.get(class_key(scene, w, h, res))
    return e.get("split", 0) if e else 0


def server_table(cache):
    """The cache as the server's per-class table (fd_server_config.tune)."""
    out = []
    for key, e in cache.items():
        scene, wh, res = key.split(":")
        w, h = wh.split("x")
        out.append(dict(scene=int(scene), w=int(w), h=int(h), res=int(res), bins=e["bins"], tile=e["tile"],
                        waves=e.get("waves", default_waves(int(res), int(w), int(h), e["bins"])), order=e.get("order", 0),
                        split=e.get("split", 0), ms=e["ms"] or 0.0))
    return out


def load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save(cache, path):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
