"""The batched and served workloads at their full sizes (MI355X only).

  batch-64       64 instances over the heightfield-1k scene (200k triangles),
                 receivers from {256^2, 512^2, 1024^2, 1920x1080}, light grids
                 from {8^2, 16^2, 32^2}, poses U[-2,2]^2 x [4,8]
                 (workloads.batch_instances)
  serve-poisson  Poisson arrivals at 500 req/s over the four resident scenes
                 of 100k / 300k / 1M / 2M triangles (workloads.serve_scenes,
                 workloads.poisson_requests)

Every instance / request must give the same bits in the mix as alone
through the plain per-frame path (fd_launch), for the batch also under a
captured HIP graph and its replay, and sampled instances / requests must
give the oracle's bits on sampled receivers.
"""
import numpy as np
import pytest
import xxhash

pytestmark = pytest.mark.gpu


def _digest(bufs):
    h = xxhash.xxh3_128()
    for b in bufs:
        h.update(b.download(np.uint8))
    return h.hexdigest()


def _poison(bufs):
    for b in bufs:
        b.upload(np.full(b.nbytes, 0xA5, np.uint8))


def _oracle_subset(oracle, frame, shadow_buf, masks_buf, n, seed):
    WH = frame.rx.w * frame.rx.h
    ids = np.sort(np.random.default_rng(seed).choice(WH, n, replace=False))
    o = oracle.run(frame, subset=ids)
    assert o.exit == 0
    R = frame.light.res_w
    per = R * (1 if R <= 32 else 2)
    sh = shadow_buf.download(np.float32)
    mk = masks_buf.download(np.uint32).reshape(WH, per)
    assert np.array_equal(sh[ids].view(np.uint32), o.shadow[ids].view(np.uint32))
    assert np.array_equal(mk[ids], o.masks.reshape(WH, per)[ids])
    return o


def test_batch64_workload(fd, wl, oracle):
    frames = wl.batch_instances(64)
    scene = frames[0].scene
    ctx = fd.Context()
    ctx.set_backface_cull(1)
    ctx.set_scenes_esp(frames[0].eps)
    verts, idx = fd.DeviceBuffer.from_array(scene.verts), fd.DeviceBuffer.from_array(scene.idx)
    ctx.add_model(verts, scene.nverts, 16, idx, scene.ntris)
    rxbuf, items = {}, []
    for f in frames:
        key = (f.rx.w, f.rx.h)
        if key not in rxbuf:
            rxbuf[key] = (fd.DeviceBuffer.from_array(f.rx.pos), fd.DeviceBuffer.from_array(f.rx.nrm))
        R = f.light.res_w
        items.append(dict(w=f.rx.w, h=f.rx.h, light=f.light, pos=rxbuf[key][0], nrm=rxbuf[key][1],
                          shadow=fd.DeviceBuffer(f.rx.w * f.rx.h * 4),
                          masks=fd.DeviceBuffer(f.rx.w * f.rx.h * fd.mask_words(R) * 4)))
    classes = sorted({(it["w"], it["h"], it["light"].res_w) for it in items})
    print(f"[batch-64] {len(items)} instances, {len(classes)} classes", flush=True)
    assert len({(w, h) for w, h, _ in classes}) == 4 and len({r for _, _, r in classes}) == 3
    # alone: the plain per-frame path, one instance after the other
    ref, ref_stats = [], []
    for it in items:
        ctx.set_receivers(it["pos"], it["nrm"], it["w"], it["h"])
        ctx.set_outputs(it["shadow"], it["masks"])
        ctx.set_area_light(it["light"])
        rc, st = ctx.launch()
        assert rc == 0
        ref.append(_digest([it["shadow"], it["masks"]]))
        ref_stats.append((st.exit_code, st.pairs, st.valid_receivers))
    # batched: direct, captured into a graph, replayed
    stream = fd.Stream()
    b = fd.Batch(ctx, items)
    for mode in ("direct", "graph", "replay"):
        for it in items:
            _poison([it["shadow"], it["masks"]])
        if mode == "graph":
            b.capture(stream.handle)
        b.launch(stream.handle)
        stream.sync()
        print(f"[batch-64] {mode}", flush=True)
        for k, it in enumerate(items):
            rc, st = b.stats(k)
            assert rc == 0 and (st.exit_code, st.pairs, st.valid_receivers) == ref_stats[k], (mode, k)
            assert _digest([it["shadow"], it["masks"]]) == ref[k], (mode, k)
    # the oracle on one instance of each receiver size (sampled receivers)
    seen = {}
    for k, it in enumerate(items):
        seen.setdefault((it["w"], it["h"]), k)
    for (w, h), k in sorted(seen.items()):
        print(f"[batch-64] oracle instance {k} {w}x{h} grid {items[k]['light'].res_w}", flush=True)
        _oracle_subset(oracle, frames[k], items[k]["shadow"], items[k]["masks"], 4000, k)
    b.close()
    stream.close()
    ctx.close()


def test_serve_poisson_workload(fd, wl, oracle):
    import time
    scenes = wl.serve_scenes()
    reqs = wl.poisson_requests(48, rate_hz=500.0)
    assert sorted({q["scene"] for q in reqs}) == [0, 1, 2, 3]
    dev = [(fd.DeviceBuffer.from_array(sc.verts), fd.DeviceBuffer.from_array(sc.idx)) for sc in scenes]
    rx = {}
    for q in reqs:
        key = (q["w"], q["h"])
        if key not in rx:
            r = wl.receivers_heightfield(q["w"], q["h"], 8.0, 8.0 * q["h"] / q["w"])
            rx[key] = (r, fd.DeviceBuffer.from_array(r.pos), fd.DeviceBuffer.from_array(r.nrm))
    lights = [wl.LightDesc(1.0, 1.0, q["pose"], res_w=q["grid"], res_h=q["grid"]) for q in reqs]
    outs = [(fd.DeviceBuffer(q["w"] * q["h"] * 4), fd.DeviceBuffer(q["w"] * q["h"] * fd.mask_words(q["grid"]) * 4))
            for q in reqs]
    # alone: one context per resident scene, the plain per-frame path
    ctxs = []
    for (v, i), sc in zip(dev, scenes):
        c = fd.Context()
        c.set_backface_cull(1)
        c.set_scenes_esp(1e-3)
        c.set_pair_capacity(64 << 20)
        c.add_model(v, sc.nverts, 16, i, sc.ntris)
        ctxs.append(c)
    ref, ref_stats = [], []
    for k, q in enumerate(reqs):
        c = ctxs[q["scene"]]
        r, pos, nrm = rx[(q["w"], q["h"])]
        c.set_receivers(pos, nrm, q["w"], q["h"])
        c.set_outputs(*outs[k])
        c.set_area_light(lights[k])
        rc, st = c.launch()
        assert rc == 0
        ref.append(_digest(outs[k]))
        ref_stats.append((st.exit_code, st.pairs))
    for c in ctxs:
        c.close()
    print(f"[serve-poisson] {len(reqs)} requests alone done", flush=True)
    # served at the Poisson arrival times
    for o in outs:
        _poison(o)
    srv = fd.Server(budget_bytes=48 << 30, streams=4, max_batch=16, eps=1e-3, max_pairs=64 << 20)
    for (v, i), sc in zip(dev, scenes):
        srv.add_scene(v, sc.nverts, 16, i, sc.ntris)
    t0 = srv.now_ms()
    for k, q in enumerate(reqs):
        target = t0 + (q["t"] - reqs[0]["t"]) * 1e3
        while srv.now_ms() < target:
            time.sleep(max(0.0, (target - srv.now_ms()) / 1e3 - 0.0002))
        r, pos, nrm = rx[(q["w"], q["h"])]
        srv.submit(q["scene"], pos, nrm, q["w"], q["h"], lights[k], outs[k][0], outs[k][1], tag=k)
    srv.drain(600000)
    done = {c.user_tag: c for c in srv.poll(len(reqs) + 8)}
    srv.close()
    print(f"[serve-poisson] served; max batch {max(c.batch_size for c in done.values())}", flush=True)
    assert sorted(done) == list(range(len(reqs)))
    assert max(c.batch_size for c in done.values()) > 1   # requests were batched as they arrived
    for k in range(len(reqs)):
        c = done[k]
        assert c.status == 0 and (c.exit_code, c.pairs) == ref_stats[k], k
        assert _digest(outs[k]) == ref[k], k
    # the oracle on one request per scene (sampled receivers)
    first = {}
    for k, q in enumerate(reqs):
        first.setdefault(q["scene"], k)
    for sc, k in sorted(first.items()):
        q = reqs[k]
        print(f"[serve-poisson] oracle request {k} scene {sc} {q['w']}x{q['h']} grid {q['grid']}", flush=True)
        frame = wl.Frame("req", scenes[sc], rx[(q["w"], q["h"])][0], lights[k], eps=1e-3, backface_cull=1)
        _oracle_subset(oracle, frame, outs[k][0], outs[k][1], 3000, 50 + k)
