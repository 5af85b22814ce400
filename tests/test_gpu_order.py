"""Walk order of the bins' triangle lists in the occlusion (MI355X only).

fd_set_triangle_order picks the order in which the occlusion kernel walks
each bin's list: the reference's (ascending triangle id), the largest
triangles on the light plane first (16 power-of-two area classes), the mesh
order (the same, except that the triangles covering 1/512 of the light
plane or more come last, in list order), or per frame on the device (the
mesh order for connected meshes, the list order for dense ones, the area
order for triangle soups).  A
receiver's mask is the OR of its triangles' masks, so the order must not
change a bit of any output; it only changes how soon receivers fill up.

Checked here: the automatic choice, oracle bits under every order, the
area- and mesh-ordered lists as stable sorts of the reference's lists by
the triangles' area classes (restated here in float32, and checked against
their definition), and the batched and served forms with per-instance and
per-class orders.
"""
import numpy as np
import pytest
import xxhash

from parity import compare

pytestmark = pytest.mark.gpu

LIST, AREA, MESH, AUTO = 1, 2, 3, 0


def _soup(wl, light_res=64):
    # boxes-1080p's generator at 480x270 receivers and 150k triangles
    # (8 Gaussian clusters of independent triangles)
    return wl.boxes_1080p(w=480, h=270, ntri=150_000, light_res=light_res)


def _mesh(wl):
    return wl.heightfield_1k(res=256, nsph=100)


def _digest(sh, mk):
    h = xxhash.xxh3_128()
    h.update(np.ascontiguousarray(sh).tobytes())
    h.update(np.ascontiguousarray(mk).tobytes())
    return h.hexdigest()


def _run(fd, frame, order):
    fr = fd.FrameRunner(frame)
    fr.ctx.set_triangle_order(order)
    rc, st = fr.run()
    assert rc == 0 and st.exit_code == 0
    return fr


def _dense(wl):
    # 1M triangles of 400-triangle spheres under a close light: about 5M
    # pairs, lists of ~20k entries in 16x16 bins
    scene = wl.serve_scenes(sizes=(1_000_000,), seeds=(3,))[0]
    rx = wl.receivers_heightfield(256, 256, 8.0, 8.0)
    return wl.Frame("dense", scene, rx, wl.LightDesc(1.0, 1.0, (-1.33, -1.05, 4.17), res_w=32, res_h=32),
                    eps=1e-3, backface_cull=1)


def _soup_small(wl):
    # few enough work items that the frame splits its bins' lists into
    # stretches (merged by atomic OR), walked in area order
    return wl.boxes_1080p(w=160, h=90, ntri=60_000, light_res=40)


@pytest.mark.parametrize("which,expect", [("soup", AREA), ("soup32", AREA), ("soup-small", AREA), ("mesh", MESH),
                                          ("dense", LIST)])
def test_auto_choice_and_oracle(fd, wl, oracle, which, expect):
    frame = {"soup": lambda: _soup(wl), "soup32": lambda: _soup(wl, 32), "soup-small": lambda: _soup_small(wl),
             "mesh": lambda: _mesh(wl), "dense": lambda: _dense(wl)}[which]()
    print(f"[order] auto {which}", flush=True)
    fr = _run(fd, frame, AUTO)
    st = fd.read_state(fr.ctx)
    assert int(st["triOrder"]) == expect
    # the rule: meshes up to 7168 pairs per 256 bins take the mesh order, denser ones the list order
    if expect in (MESH, LIST):
        assert (int(st["pairs"]) > 7168 * 256) == (expect == LIST)
    rep, fr, o = compare(fd, oracle, frame, fr=fr)
    bad = {k: v for k, v in rep.items() if v}
    assert not bad, bad
    fr.close()


@pytest.mark.parametrize("which", ["soup", "mesh"])
def test_orders_give_identical_bits(fd, wl, which):
    frame = _soup(wl) if which == "soup" else _mesh(wl)
    print(f"[order] forced orders {which}", flush=True)
    digests, pair_lists = {}, {}
    for order in (LIST, AREA, MESH, AUTO):
        fr = _run(fd, frame, order)
        chosen = int(fd.read_state(fr.ctx)["triOrder"])
        assert chosen == (order if order else chosen) and chosen in (LIST, AREA, MESH)
        sh, mk = fr.outputs()
        digests[order] = _digest(sh, mk)
        pair_lists[order] = fr.ctx.debug_read("pair_tri", np.int32)   # always the reference's order
        walk = fr.ctx.debug_read("pair_tri_walk", np.int32)
        if chosen == LIST:
            assert np.array_equal(walk, pair_lists[order])
        fr.close()
    assert len(set(digests.values())) == 1, digests
    assert all(np.array_equal(pair_lists[LIST], pair_lists[o]) for o in (AREA, MESH, AUTO))


def _area_classes(tris, light, s):
    """Area class of each triangle record, restated in float32 with the
    kernel's expressions: the three makeRectangle points, twice the area
    |(x1 - x0)(y2 - y0) - (y1 - y0)(x2 - x0)|, halved and scaled to bin
    cells, 7 - floor(log2) from the exponent bits, clamped to 0..15."""
    f32 = np.float32
    m = light["m"].astype(f32)

    def rect(v):
        x, y, z = v[:, 0], v[:, 1], v[:, 2]
        rz = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * f32(1)
        rx = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * f32(1)
        ry = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * f32(1)
        return rx / rz * f32(0.5) + f32(0.5), ry / rz * f32(0.5) + f32(0.5)

    with np.errstate(all="ignore"):
        p = [rect(tris[k].astype(f32)) for k in ("v0", "v1", "v2")]
        twice = np.abs((p[1][0] - p[0][0]) * (p[2][1] - p[0][1]) - (p[1][1] - p[0][1]) * (p[2][0] - p[0][0]))
        cells = ((twice * f32(0.5)) * f32(s["factor"][0])) * f32(s["factor"][1])
        e = (cells.astype(f32).view(np.uint32) >> 23).astype(np.int64) - 127
        cls = np.clip(7 - e, 0, 15)
        return np.where(cells > 0, cls, 15), cells.astype(np.float64)


def _walk_keys(cls, order, log2G):
    """Walk key per class: the class (area order); for the mesh order the
    classes above big = 16 - 2 log2G (1/512 of the light plane) first, in
    class order, and the classes up to big last, as one key."""
    if order == AREA:
        return cls
    big = 16 - 2 * log2G
    return np.where(cls > big, cls - big - 1, 15 - big)


@pytest.mark.parametrize("order,which", [(AREA, "soup"), (MESH, "soup"), (MESH, "mesh")])
def test_walked_lists_are_stable_class_sorts(fd, wl, order, which):
    frame = _soup(wl) if which == "soup" else _mesh(wl)
    print(f"[order] walked lists, order {order}, {which}", flush=True)
    fr = _run(fd, frame, order)
    ctx = fr.ctx
    s = fd.read_state(ctx)
    assert int(s["triOrder"]) == order
    cls = ctx.debug_read("tri_class", np.uint8)
    tris = fd.read_tris(ctx)
    light = fd.read_light(ctx)
    pairs = ctx.debug_read("pair_tri", np.int32)
    walk = ctx.debug_read("pair_tri_walk", np.int32)
    ts = ctx.debug_read("bin_tri_start", np.int32)
    te = ctx.debug_read("bin_tri_end", np.int32)
    fr.close()
    assert len(pairs) == int(s["pairs"]) and len(walk) == len(pairs)
    assert cls.max() < 16
    used = np.unique(pairs)
    # the device's classes equal the float32 restatement ...
    exp, cells = _area_classes(tris[used], light, s)
    assert np.array_equal(cls[used].astype(np.int64), exp)
    # ... and obey their definition: class k in 1..14 holds triangles of
    # 2^(7-k) <= area < 2^(8-k) bin cells, class 0 every area >= 128 cells,
    # class 15 every area < 2^-7 cells (area from the float32 points; within
    # 1e-5 of a class boundary either side)
    c = exp
    lo = np.where(c == 15, 0.0, 2.0 ** (7 - c))
    hi = np.where(c == 0, np.inf, 2.0 ** (8 - c))
    ok = (cells >= lo * (1 - 1e-5)) & (cells < hi * (1 + 1e-5))
    assert ok.all(), (int((~ok).sum()), cells[~ok][:5], c[~ok][:5])
    assert len(np.unique(c)) >= 4   # the frame spans several classes
    # each bin's walked list: the reference's list stably sorted by walk key
    keys = _walk_keys(cls.astype(np.int64), order, int(s["log2G"]))
    if order == MESH:
        # the last key holds exactly the triangles covering 1/512 of the
        # light plane or more (G x G bin cells; within 1e-5 of the boundary
        # either side)
        frac = cells / float(int(s["G"]) ** 2)
        last = keys[used] == keys.max()
        assert (last | (frac < (1 + 1e-5) / 512)).all() and (~last | (frac >= (1 - 1e-5) / 512)).all()
        if which == "soup":
            assert last.any() and not last.all()   # the soup has such triangles, and others
    nonempty = 0
    for b in range(len(ts)):
        if te[b] <= ts[b]:
            continue
        nonempty += 1
        ref = pairs[ts[b]:te[b]]
        expw = ref[np.argsort(keys[ref], kind="stable")]
        assert np.array_equal(walk[ts[b]:te[b]], expw), b
    assert nonempty > 0


def test_batch_over_soup(fd, wl):
    # the batched form with the default (automatic) order over a soup scene,
    # instance by instance the same bits as the plain path in list order
    scene = wl.tris_clusters(60_000, 11)
    ctx = fd.Context()
    ctx.set_backface_cull(1)
    ctx.set_scenes_esp(1e-3)
    verts, idx = fd.DeviceBuffer.from_array(scene.verts), fd.DeviceBuffer.from_array(scene.idx)
    ctx.add_model(verts, scene.nverts, 16, idx, scene.ntris)
    items = []
    for (w, h, R, c) in [(160, 90, 16, (0.0, 0.0, 6.0)), (320, 180, 33, (1.0, -0.5, 7.0)), (256, 256, 64, (-1.0, 1.0, 6.5))]:
        rx = wl.receivers_boxes(w, h, 20, 3)
        items.append(dict(w=w, h=h, bins=0, rx=rx, light=wl.LightDesc(1.5, 0.5, c, res_w=R, res_h=R),
                          pos=fd.DeviceBuffer.from_array(rx.pos), nrm=fd.DeviceBuffer.from_array(rx.nrm),
                          shadow=fd.DeviceBuffer(w * h * 4), masks=fd.DeviceBuffer(w * h * fd.mask_words(R) * 4)))
    alone, areas = [], 0
    for it in items:
        fr = wl.Frame("alone", scene, it["rx"], it["light"], eps=1e-3, backface_cull=1)
        run = fd.FrameRunner(fr)
        run.ctx.set_triangle_order(LIST)
        rc, st = run.run()
        assert rc == 0 and st.exit_code == 0
        alone.append(_digest(*run.outputs()))
        run.ctx.set_triangle_order(AUTO)
        rc, st = run.run()
        areas += int(fd.read_state(run.ctx)["triOrder"]) == AREA
        assert rc == 0 and _digest(*run.outputs()) == alone[-1]   # split lists in area order: the same bits
        run.close()
    assert areas == len(items)   # the automatic choice takes the area order on every instance
    print("[order] batch over soup", flush=True)
    stream = fd.Stream()
    b = fd.Batch(ctx, items)
    b.launch(stream.handle)
    stream.sync()
    for k, it in enumerate(items):
        rc, st = b.stats(k)
        assert rc == 0 and st.exit_code == 0, k
        sh = it["shadow"].download(np.float32)
        mk = it["masks"].download(np.uint32)
        assert _digest(sh, mk) == alone[k], k
    b.close()
    # per-instance orders in one batch: the same bits
    print("[order] batch with per-instance orders", flush=True)
    for it, order in zip(items, (LIST, AREA, MESH)):
        it["order"] = order
        it["shadow"].upload(np.zeros(it["shadow"].nbytes, np.uint8))
        it["masks"].upload(np.full(it["masks"].nbytes, 0xA5, np.uint8))
    b = fd.Batch(ctx, items)
    b.launch(stream.handle)
    stream.sync()
    for k, it in enumerate(items):
        rc, st = b.stats(k)
        assert rc == 0 and st.exit_code == 0, k
        assert _digest(it["shadow"].download(np.float32), it["masks"].download(np.uint32)) == alone[k], k
    b.close()
    # per-class orders from a server's tuning table: the same bits
    print("[order] server with per-class orders", flush=True)
    tune = [dict(scene=0, w=it["w"], h=it["h"], res=it["light"].res_w, bins=16, tile=8, order=order, ms=1.0)
            for it, order in zip(items, (AREA, LIST, MESH))]
    srv = fd.Server(budget_bytes=1 << 30, streams=2, max_batch=2, eps=1e-3, tune=tune)
    srv.add_scene(verts, scene.nverts, 16, idx, scene.ntris)
    for k, it in enumerate(items):
        it["masks"].upload(np.full(it["masks"].nbytes, 0x5A, np.uint8))
        srv.submit(0, it["pos"], it["nrm"], it["w"], it["h"], it["light"], it["shadow"], it["masks"], tag=k)
    srv.drain(120000)
    done = srv.poll(len(items) + 2)
    srv.close()
    assert len(done) == len(items) and all(c.status == 0 for c in done)
    for k, it in enumerate(items):
        assert _digest(it["shadow"].download(np.float32), it["masks"].download(np.uint32)) == alone[k], k
    stream.close()
    ctx.close()
