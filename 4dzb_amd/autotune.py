"""Autotuner of the batched and served forms.

For one instance class -- scene, receiver size, light grid -- it runs a
representative instance under every candidate layout (bins per light-plane
side x receiver tile width x waves per occlusion work item) through the
library's fd_autotune, which times each on the device and compares every
output bit with the default layout (16x16 bins, 8x8 tiles, the library's
default waves).  The fastest bit-exact layout wins.  The bin grid
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


def default_waves(res):
    """The library's default waves per occlusion work item for a res x res
    light grid (fd_set_occlusion_waves)."""
    return 8 if res > 32 else 4


def default_layout(res):
    return (16, 8, default_waves(res))


def class_key(scene, w, h, res):
    return f"{scene}:{w}x{h}:{res}"


def tune_one(ctx, item, candidates=CANDIDATES, reps=3):
    """Tune one instance class on `ctx`'s scene.  item: a Batch item (pos,
    nrm, w, h, light, shadow).  Returns the cache entry."""
    t0 = time.perf_counter()
    best, ms = ctx.autotune(item, candidates, reps)
    dl = default_layout(item["light"].res_w)
    default = ms[candidates.index(dl)] if dl in candidates else None
    b, t, w = candidates[best] if best >= 0 else dl   # nothing bit-exact besides the reference: the default
    return dict(bins=b, tile=t, waves=w, ms=(ms[best] if best >= 0 else default), default_ms=default,
                candidates={f"{cb}x{ct}x{cw}": (round(m, 4) if m >= 0 else ("not exact" if m == -1 else "pair arena too small"))
                            for (cb, ct, cw), m in zip(candidates, ms)},
                tune_s=round(time.perf_counter() - t0, 3))


def tune_classes(ctx, scene, items, cache=None, candidates=CANDIDATES, reps=3):
    """items: Batch items; one representative per class (scene, w, h, res) is
    tuned unless the cache already holds it.  Returns the cache (dict)."""
    cache = {} if cache is None else cache
    seen = set()
    for it in items:
        key = class_key(scene, it["w"], it["h"], it["light"].res_w)
        if key in cache or key in seen:
            continue
        seen.add(key)
        cache[key] = tune_one(ctx, it, candidates, reps)
    return cache


def layout(cache, scene, w, h, res):
    """(bins, tile, waves) of a class, or the default layout."""
    e = cache.get(class_key(scene, w, h, res))
    return (e["bins"], e["tile"], e.get("waves", default_waves(res))) if e else default_layout(res)


def server_table(cache):
    """The cache as the server's per-class table (fd_server_config.tune)."""
    out = []
    for key, e in cache.items():
        scene, wh, res = key.split(":")
        w, h = wh.split("x")
        out.append(dict(scene=int(scene), w=int(w), h=int(h), res=int(res), bins=e["bins"], tile=e["tile"],
                        waves=e.get("waves", default_waves(int(res))), ms=e["ms"] or 0.0))
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
