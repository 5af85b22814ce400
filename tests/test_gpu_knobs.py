"""Schedule knobs together (MI355X only): seeded random small frames (a third
of them under a tilted light), each with its own bin grid, receiver tile width, waves per occlusion work item,
walk order of the bins' lists and list split, all compared bit for bit with
the oracle (which runs the same bin grid; the other knobs are not part of
its model).  No knob may change a stage output the oracle restates, a mask
bit or a shadow value.
"""
import numpy as np
import pytest

from parity import compare, failures

pytestmark = pytest.mark.gpu

SCENES = ("uniform", "clusters", "spheres")
N_CASES = 48     # small frames: split lists under the automatic rule
N_LARGE = 8      # up to 512x512 receivers and 100k triangles


def _case(wl, k, large=False):
    r = np.random.default_rng((2000 if large else 1000) + k)
    scene = SCENES[k % len(SCENES)]
    seed = int(r.integers(1, 1 << 20))
    f = 10 if large else 1
    if scene == "uniform":
        sc = wl.tris_uniform(int(r.integers(500, 6000)) * f, seed)
    elif scene == "clusters":
        sc = wl.tris_clusters(int(r.integers(2000, 10000)) * f, seed, box=(-4, 4, -4, 4, 0.5, 3.5))
    else:
        sc = wl.spheres(int(r.integers(5, 25)) * f, seed, box=(-4, 4, -4, 4, 0.5, 3))
    lo, hi = (256, 513) if large else (24, 200)
    w, h = int(r.integers(lo, hi)), int(r.integers(lo, hi))
    rx = wl.receivers_heightfield(w, h, 4.0, 4.0 * h / w) if r.random() < 0.5 else wl.receivers_plane(w, h, 4.0)
    res = int(r.choice([1, 5, 8, 16, 31, 32, 33, 40, 64]))
    pos = (float(r.uniform(-1.5, 1.5)), float(r.uniform(-1.5, 1.5)), float(r.uniform(4.0, 7.0)))
    light = wl.LightDesc(float(r.uniform(0.3, 1.2)), float(r.uniform(0.3, 1.2)), pos, res_w=res, res_h=res)
    if k % 3 == 2:   # a tilted light: the non-axial plane coefficients
        light.view_dir = (float(r.uniform(-0.35, 0.35)), float(r.uniform(-0.35, 0.35)), -1.0)
        light.up_dir = (float(r.uniform(-0.3, 0.3)), 1.0, float(r.uniform(-0.2, 0.2)))
    fr = wl.Frame(f"knobs-{k}", sc, rx, light, eps=1e-3, backface_cull=int(r.random() < 0.8),
                  grid_log2=int(r.integers(3, 7)), tile_width=int(r.choice([8, 16, 32, 64])),
                  occ_waves=int(r.choice([0, 2, 4, 8])), list_split=int(r.choice([0, 0, 1, 2, 5, 16, 64])))
    order = int(r.choice([0, 1, 2, 3]))
    return fr, order


@pytest.mark.parametrize("k,large", [(k, False) for k in range(N_CASES)] + [(k, True) for k in range(N_LARGE)])
def test_knob_combination_matches_oracle(fd, wl, oracle, k, large):
    frame, order = _case(wl, k, large)
    print(f"[knobs] case {k}: {frame.rx.w}x{frame.rx.h}, light {frame.light.res_w}, bins {1 << frame.grid_log2}, "
          f"tile {frame.tile_width}, waves {frame.occ_waves}, split {frame.list_split}, order {order}, "
          f"{frame.scene.ntris} triangles", flush=True)
    fr = fd.FrameRunner(frame)
    fr.ctx.set_triangle_order(order)
    rep, fr, o = compare(fd, oracle, frame, fr=fr)
    bad = failures(rep)
    st = fd.read_state(fr.ctx)
    fr.close()
    assert not bad, (k, bad)
    if order:
        assert int(st["triOrder"]) == order or o.exit != 0


def test_batch_of_random_knobs_matches_oracle(fd, wl, oracle):
    """One scene, 12 instances with their own receivers, light, bin grid,
    tile width, waves, walk order and list split in one batch; the batch is
    captured as a HIP graph and replayed twice (outputs poisoned between the
    replays); every replay's outputs equal the oracle's, instance by
    instance."""
    r = np.random.default_rng(77)
    scene = wl.tris_clusters(15000, 5, box=(-4, 4, -4, 4, 0.5, 3.5))
    ctx = fd.Context()
    ctx.set_backface_cull(1)
    ctx.set_scenes_esp(1e-3)
    verts, idx = fd.DeviceBuffer.from_array(scene.verts), fd.DeviceBuffer.from_array(scene.idx)
    ctx.add_model(verts, scene.nverts, 16, idx, scene.ntris)
    items, refs = [], []
    for k in range(12):
        w, h = int(r.integers(32, 300)), int(r.integers(32, 300))
        rx = wl.receivers_heightfield(w, h, 4.0, 4.0 * h / w)
        res = int(r.choice([1, 7, 16, 32, 33, 64]))
        light = wl.LightDesc(float(r.uniform(0.3, 1.0)), float(r.uniform(0.3, 1.0)),
                             (float(r.uniform(-1, 1)), float(r.uniform(-1, 1)), float(r.uniform(4, 7))), res_w=res, res_h=res)
        bins = int(r.choice([8, 16, 32, 64]))
        it = dict(w=w, h=h, light=light, bins=bins, tile=int(r.choice([8, 16, 32, 64])),
                  waves=int(r.choice([0, 2, 4, 8])), order=int(r.choice([0, 1, 2, 3])),
                  split=int(r.choice([0, 1, 3, 16])), pos=fd.DeviceBuffer.from_array(rx.pos),
                  nrm=fd.DeviceBuffer.from_array(rx.nrm), shadow=fd.DeviceBuffer(w * h * 4),
                  masks=fd.DeviceBuffer(w * h * fd.mask_words(res) * 4))
        items.append(it)
        o = oracle.run(wl.Frame(f"b{k}", scene, rx, light, eps=1e-3, backface_cull=1,
                                grid_log2=bins.bit_length() - 1))
        refs.append(o)
    stream = fd.Stream()
    b = fd.Batch(ctx, items)
    b.capture(stream.handle)
    for rep in range(2):
        print(f"[knobs] batch replay {rep}", flush=True)
        for it in items:
            it["shadow"].upload(np.full(it["w"] * it["h"], -3.0, np.float32))
            it["masks"].upload(np.full(it["masks"].nbytes // 4, 0x5A5A5A5A, np.uint32))
        b.launch(stream.handle)
        stream.sync()
        for k, (it, o) in enumerate(zip(items, refs)):
            rc, st = b.stats(k)
            assert rc == 0 and st.exit_code == o.exit, (rep, k)
            sh = it["shadow"].download(np.float32)
            mk = it["masks"].download(np.uint32)
            assert np.array_equal(sh.view(np.uint32), np.asarray(o.shadow, np.float32).view(np.uint32)), (rep, k)
            assert np.array_equal(mk, o.masks.reshape(mk.shape)), (rep, k)
    b.close()
    stream.close()
    ctx.close()


def test_server_with_random_tuning_matches_oracle(fd, wl, oracle):
    """Two resident scenes, 10 requests of distinct classes, a tuning table
    with a random layout, waves, walk order and list split per class; every
    request's outputs equal the oracle's on the tuned bin grid."""
    r = np.random.default_rng(91)
    scenes = [wl.tris_clusters(12000, 8, box=(-4, 4, -4, 4, 0.5, 3.5)), wl.spheres(20, 9, box=(-4, 4, -4, 4, 0.5, 3))]
    dev = [(fd.DeviceBuffer.from_array(sc.verts), fd.DeviceBuffer.from_array(sc.idx)) for sc in scenes]
    reqs, tune = [], []
    for k in range(10):
        sc = k % 2
        w, h = 40 + 17 * k, 64 + 9 * k   # distinct sizes: one class per request
        res = int(r.choice([4, 16, 32, 40, 64]))
        bins = int(r.choice([8, 16, 32]))
        rx = wl.receivers_heightfield(w, h, 4.0, 4.0 * h / w)
        light = wl.LightDesc(float(r.uniform(0.3, 1.0)), float(r.uniform(0.3, 1.0)),
                             (float(r.uniform(-1, 1)), float(r.uniform(-1, 1)), float(r.uniform(4, 7))), res_w=res, res_h=res)
        tune.append(dict(scene=sc, w=w, h=h, res=res, bins=bins, tile=int(r.choice([8, 16, 32])),
                         waves=int(r.choice([2, 4, 8])), order=int(r.choice([0, 1, 2, 3])),
                         split=int(r.choice([0, 1, 4, 16])), ms=1.0))
        o = oracle.run(wl.Frame(f"s{k}", scenes[sc], rx, light, eps=1e-3, backface_cull=1,
                                grid_log2=bins.bit_length() - 1))
        reqs.append(dict(scene=sc, w=w, h=h, light=light, o=o, pos=fd.DeviceBuffer.from_array(rx.pos),
                         nrm=fd.DeviceBuffer.from_array(rx.nrm), shadow=fd.DeviceBuffer(w * h * 4),
                         masks=fd.DeviceBuffer(w * h * fd.mask_words(res) * 4)))
    srv = fd.Server(budget_bytes=1 << 30, streams=2, max_batch=3, eps=1e-3, tune=tune)
    for (v, i), sc in zip(dev, scenes):
        srv.add_scene(v, sc.nverts, 16, i, sc.ntris)
    for k, q in enumerate(reqs):
        q["masks"].upload(np.full(q["masks"].nbytes, 0xA5, np.uint8))
        srv.submit(q["scene"], q["pos"], q["nrm"], q["w"], q["h"], q["light"], q["shadow"], q["masks"], tag=k)
    print("[knobs] server drain", flush=True)
    srv.drain(120000)
    done = sorted(srv.poll(len(reqs) + 2), key=lambda c: c.user_tag)
    srv.close()
    assert len(done) == len(reqs) and all(c.status == 0 for c in done)
    for k, q in enumerate(reqs):
        o = q["o"]
        sh = q["shadow"].download(np.float32)
        mk = q["masks"].download(np.uint32)
        assert np.array_equal(sh.view(np.uint32), np.asarray(o.shadow, np.float32).view(np.uint32)), k
        assert np.array_equal(mk, o.masks.reshape(mk.shape)), k
