"""Batched form: every instance's result is the same alone (fd_launch) and
in the mix (one launch per stage for all instances, direct and replayed from
a HIP graph), bit for bit (MI355X only)."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (w, h, light res, light centre, bins per side)
MIX = [(96, 96, 8, (0.5, -0.3, 5.0), 0), (160, 120, 16, (-1.0, 0.7, 4.5), 0), (256, 256, 32, (0.2, 0.1, 6.0), 0),
       (200, 150, 33, (1.5, -1.2, 7.0), 0), (128, 128, 16, (-0.5, -1.5, 5.5), 32), (64, 200, 64, (0.0, 0.0, 6.0), 8),
       (256, 256, 32, (1.0, 1.0, 4.2), 0)]


def _setup(fd, wl):
    scene = wl.spheres(80, 7)
    ctx = fd.Context()
    ctx.set_backface_cull(1)
    ctx.set_scenes_esp(1e-3)
    verts = fd.DeviceBuffer.from_array(scene.verts)
    idx = fd.DeviceBuffer.from_array(scene.idx)
    ctx.add_model(verts, scene.nverts, 16, idx, scene.ntris)
    items = []
    for (w, h, R, c, bins) in MIX:
        rx = wl.receivers_heightfield(w, h, 8.0, 8.0 * h / w)
        items.append(dict(w=w, h=h, bins=bins, rx=rx, light=wl.LightDesc(1.0, 1.0, c, res_w=R, res_h=R),
                          pos=fd.DeviceBuffer.from_array(rx.pos), nrm=fd.DeviceBuffer.from_array(rx.nrm),
                          shadow=fd.DeviceBuffer(w * h * 4), masks=fd.DeviceBuffer(w * h * fd.mask_words(R) * 4)))
    return scene, ctx, items, (verts, idx)


def _alone(fd, wl, scene, it):
    fr = wl.Frame("alone", scene, it["rx"], it["light"], eps=1e-3, backface_cull=1,
                  grid_log2=(it["bins"].bit_length() - 1) if it["bins"] else 4)
    run = fd.FrameRunner(fr)
    rc, st = run.run()
    assert rc == 0
    out = run.outputs()
    run.close()
    return out, st


def test_batch_equals_alone(fd, wl):
    scene, ctx, items, keep = _setup(fd, wl)
    alone = [_alone(fd, wl, scene, it) for it in items]
    stream = fd.Stream()
    b = fd.Batch(ctx, items)
    for mode in ["direct", "graph", "graph-replay"]:
        for it in items:   # poison the outputs between runs
            it["shadow"].upload(np.full(it["w"] * it["h"], -7.0, np.float32))
        if mode == "graph":
            b.capture(stream.handle)
        b.launch(stream.handle)
        stream.sync()
        for k, it in enumerate(items):
            (sh0, mk0), st0 = alone[k]
            sh = it["shadow"].download(np.float32)
            mk = it["masks"].download(np.uint32)
            rc, st = b.stats(k)
            assert rc == 0 and st.exit_code == st0.exit_code and st.pairs == st0.pairs, (mode, k)
            assert np.array_equal(sh.view(np.uint32), sh0.view(np.uint32)), (mode, k)
            assert np.array_equal(mk, mk0), (mode, k)
    b.close()
    stream.close()
    ctx.close()


def test_occlusion_waves_give_identical_bits(fd, wl):
    """2, 4 or 8 waves per occlusion work item: the same bits, alone (both
    row widths) and in one batch that mixes them (six occlusion classes)."""
    scene, ctx, items, keep = _setup(fd, wl)
    alone = [_alone(fd, wl, scene, it) for it in items]
    for it in items[:2] + items[3:4]:   # 8, 16 and 33 light samples per side
        for w in (2, 4, 8):
            fr = wl.Frame("w", scene, it["rx"], it["light"], eps=1e-3, backface_cull=1,
                          grid_log2=(it["bins"].bit_length() - 1) if it["bins"] else 4, occ_waves=w)
            run = fd.FrameRunner(fr)
            rc, st = run.run()
            assert rc == 0
            sh, mk = run.outputs()
            run.close()
            (sh0, mk0), _ = alone[items.index(it)]
            assert np.array_equal(sh.view(np.uint32), sh0.view(np.uint32)) and np.array_equal(mk, mk0), w
    # two assignments that together put instances in all six classes (the
    # 33 and 64 light-sample instances are items 3 and 5)
    stream = fd.Stream()
    for waves in [(2, 4, 8, 2, 4, 8, 4), (8, 2, 4, 4, 8, 2, 2)]:
        print(f"[waves] mixed batch {waves}", flush=True)
        for k, it in enumerate(items):
            it["waves"] = waves[k]
            it["shadow"].upload(np.full(it["w"] * it["h"], -7.0, np.float32))
            it["masks"].upload(np.full(it["masks"].nbytes // 4, 0xA5A5A5A5, np.uint32))
        b = fd.Batch(ctx, items)
        b.launch(stream.handle)
        stream.sync()
        for k, it in enumerate(items):
            (sh0, mk0), st0 = alone[k]
            rc, st = b.stats(k)
            assert rc == 0 and st.exit_code == st0.exit_code, (waves, k)
            assert np.array_equal(it["shadow"].download(np.float32).view(np.uint32), sh0.view(np.uint32)), (waves, k)
            assert np.array_equal(it["masks"].download(np.uint32), mk0), (waves, k)
        b.close()
    stream.close()
    ctx.close()


def test_list_split_gives_identical_bits(fd, wl):
    """Any number of stretches per bin list (fd_set_list_split 1..64, or the
    automatic rule): the same bits, alone, per instance in one batch, and per
    class in a server."""
    scene, ctx, items, keep = _setup(fd, wl)
    alone = [_alone(fd, wl, scene, it) for it in items]
    for k in (2, 5):   # 256x256 / 32x32 light, 64x200 / 64x64 light on 8x8 bins
        it = items[k]
        for segs in (1, 3, 16, 64):
            print(f"[split] instance {k} alone, {segs} stretches", flush=True)
            fr = wl.Frame("s", scene, it["rx"], it["light"], eps=1e-3, backface_cull=1,
                          grid_log2=(it["bins"].bit_length() - 1) if it["bins"] else 4, list_split=segs)
            run = fd.FrameRunner(fr)
            rc, st = run.run()
            assert rc == 0
            sh, mk = run.outputs()
            run.close()
            (sh0, mk0), _ = alone[k]
            assert np.array_equal(sh.view(np.uint32), sh0.view(np.uint32)) and np.array_equal(mk, mk0), (k, segs)
    stream = fd.Stream()
    splits = (1, 2, 5, 16, 64, 7, 0)
    print(f"[split] mixed batch {splits}", flush=True)
    for k, it in enumerate(items):
        it["split"] = splits[k]
        it["shadow"].upload(np.full(it["w"] * it["h"], -7.0, np.float32))
        it["masks"].upload(np.full(it["masks"].nbytes // 4, 0xA5A5A5A5, np.uint32))
    b = fd.Batch(ctx, items)
    b.launch(stream.handle)
    stream.sync()
    for k, it in enumerate(items):
        (sh0, mk0), st0 = alone[k]
        rc, st = b.stats(k)
        assert rc == 0 and st.exit_code == st0.exit_code, k
        assert np.array_equal(it["shadow"].download(np.float32).view(np.uint32), sh0.view(np.uint32)), k
        assert np.array_equal(it["masks"].download(np.uint32), mk0), k
    b.close()
    stream.close()
    ctx.close()
    # per class in a server (bins as each item asks; 16 where it leaves it to the default)
    print("[split] server with per-class splits", flush=True)
    tune = [dict(scene=0, w=it["w"], h=it["h"], res=it["light"].res_w, bins=it["bins"] or 16, tile=8,
                 split=(4, 1, 9, 2, 32, 3, 0)[k], ms=1.0) for k, it in enumerate(items)]
    srv = fd.Server(budget_bytes=1 << 30, streams=2, max_batch=3, eps=1e-3, tune=tune)
    srv.add_scene(keep[0], scene.nverts, 16, keep[1], scene.ntris)
    for k, it in enumerate(items):
        it["masks"].upload(np.full(it["masks"].nbytes, 0x5A, np.uint8))
        srv.submit(0, it["pos"], it["nrm"], it["w"], it["h"], it["light"], it["shadow"], it["masks"], tag=k)
    srv.drain(120000)
    done = srv.poll(len(items) + 2)
    srv.close()
    assert len(done) == len(items) and all(c.status == 0 for c in done)
    for k, it in enumerate(items):
        (sh0, mk0), _ = alone[k]
        assert np.array_equal(it["shadow"].download(np.float32).view(np.uint32), sh0.view(np.uint32)), k
        assert np.array_equal(it["masks"].download(np.uint32), mk0), k


def test_batch_instance_against_oracle(fd, wl, oracle):
    scene, ctx, items, keep = _setup(fd, wl)
    b = fd.Batch(ctx, items)
    b.launch()
    fd.lib().fd_device_sync()
    for k in (0, 3, 5):
        it = items[k]
        fr = wl.Frame("o", scene, it["rx"], it["light"], eps=1e-3, backface_cull=1,
                      grid_log2=(it["bins"].bit_length() - 1) if it["bins"] else 4)
        o = oracle.run(fr)
        assert np.array_equal(it["shadow"].download(np.float32).view(np.uint32), o.shadow.view(np.uint32)), k
        assert np.array_equal(it["masks"].download(np.uint32), o.masks), k
    b.close()
    ctx.close()


def test_tile_widths_and_autotune(fd, wl, oracle):
    """Every receiver tile shape gives the oracle's bits; the autotuner
    reports a time for every bit-exact candidate and picks one of them."""
    frame = wl.plane_256(ntri=3000, res=96, light_res=12)
    o = oracle.run(frame)
    for tw in (8, 16, 32, 64):
        frame.tile_width = tw
        print(f"[tile] width {tw}", flush=True)
        fr = fd.FrameRunner(frame)
        rc, st = fr.run()
        sh, mk = fr.outputs()
        assert rc == 0 and st.exit_code == o.exit
        assert np.array_equal(sh.view(np.uint32), np.asarray(o.shadow, np.float32).view(np.uint32))
        assert np.array_equal(mk, o.masks.reshape(mk.shape))
        fr.close()
    frame.tile_width = 8
    fr = fd.FrameRunner(frame)
    item = dict(pos=fr.pos, nrm=fr.nrm, w=frame.rx.w, h=frame.rx.h, light=frame.light, shadow=fr.shadow)
    cands = [(8, 8), (16, 8), (16, 16), (32, 32), (16, 64), (16, 8, 2), (32, 16, 2), (16, 8, 0, 1), (16, 8, 4, 2),
             (16, 8, 0, 0, 1), (16, 8, 0, 0, 7)]
    best, ms = fr.ctx.autotune(item, cands, reps=2)
    assert 0 <= best < len(cands) and ms[best] > 0
    assert ms[1] > 0   # the reference layout itself is always exact
    assert all(m in (-1, -2) or m > 0 for m in ms)
    assert ms[best] == min(m for m in ms if m > 0)
    assert ms[7] > 0 and ms[8] > 0   # the walk orders change no bit
    assert ms[9] > 0 and ms[10] > 0   # nor does the list split
    at = importlib.import_module("4dzb_amd.autotune")
    e = at.tune_one(fr.ctx, item, [(16, 8, 4), (8, 8, 4)], reps=2, order_items=[item, item])
    assert e["order"] in (1, 2, 3) and e["ms"] > 0
    tot = {o: e["order_ms"][n] for o, n in ((1, "list"), (2, "area"), (3, "mesh"))}
    assert e["order_ms"]["instances"] == 2 and min(tot.values()) > 0
    assert tot[e["order"]] == min(tot.values())
    assert at.tune_one(fr.ctx, item, [(16, 8, 4)], reps=1)["order"] == 0   # no order choice asked: automatic
    fr.close()
