"""Failure paths and remaining exits on the device (MI355X only).

  pair-arena overflow   the reference writes past its fixed pair buffer
                        (FDc/Step4_RasterizeTriangleToBin.inl:146-155,
                        FD/constants.h:15); here FD_ERR_CAPACITY with defined
                        outputs, and the context works again once the arena
                        is raised (plain, batched and served forms)
  exit 4 (empty AABB)   FDc/Step2_3_defineFDbuffer.inl:45-47
  exit 7 (no pairs)     FD/api.cpp:522
  unsupported grid      Step5_2_shadowCalPerSample.inl:251-254
  counting kernel       the profiling variant of the occlusion kernel gives
                        the timed kernel's and the oracle's bits, and its work
                        counters equal the oracle's per-receiver loop counts
"""
import numpy as np
import pytest

from parity import compare, failures

pytestmark = pytest.mark.gpu


def _check(fd, oracle, frame, expect_exit):
    print(f"[failure] {frame.name} {frame.rx.w}x{frame.rx.h} expect exit {expect_exit}", flush=True)
    rep, fr, o = compare(fd, oracle, frame)
    fr.close()
    assert not failures(rep), failures(rep)
    assert o.exit == expect_exit


def test_exit_empty_aabb(fd, oracle, wl):
    fr = wl.plane_256(ntri=500, res=64)
    fr.rx = wl.receivers_plane(64, 64, 0.001)
    fr.light.right_radius = -0.5
    _check(fd, oracle, fr, 4)


def test_exit_no_pairs(fd, oracle, wl):
    fr = wl.plane_256(ntri=500, res=64)
    fr.scene.verts[:, 2] -= 3.6
    _check(fd, oracle, fr, 7)


def _backfacing_frame(wl):
    fr = wl.plane_256(ntri=3000, res=128)
    fr.rx.nrm[:20, :, 2] = -1.0   # 20 rows of back-facing receivers
    return fr


def test_pair_capacity_plain(fd, oracle, wl):
    frame = _backfacing_frame(wl)
    o = oracle.run(frame)
    pairs = int(o.scalars["pairs"])
    assert o.exit == 0 and pairs > 1000
    run = fd.FrameRunner(frame, pair_capacity=pairs - 1)
    print(f"[failure] capacity {pairs - 1} < {pairs} pairs", flush=True)
    rc, st = run.run()
    assert rc == fd.FD_ERR_CAPACITY and st.exit_code == fd.EXIT_CAPACITY
    assert st.pairs == pairs
    sh, mk = run.outputs()
    back = o.backface.astype(bool)
    # defined outputs: the early-exit fill (back-facing receivers occluded)
    assert np.all(sh[back] == 1.0) and np.all(sh[~back] == 0.0)
    per = mk.reshape(len(sh), -1)
    assert np.all(per[back] == 0xFFFF) and np.all(per[~back] == 0)
    # exactly enough: the frame runs, bit-exact
    run.ctx.set_pair_capacity(pairs)
    rep, _, _ = compare(fd, oracle, frame, fr=run, o=o)
    assert not failures(rep), failures(rep)
    # too small again, then raised: the context recovers each time
    run.ctx.set_pair_capacity(max(1, pairs // 3))
    rc, st = run.run()
    assert rc == fd.FD_ERR_CAPACITY
    run.ctx.set_pair_capacity(64 << 20)
    rep, _, _ = compare(fd, oracle, frame, fr=run, o=o)
    assert not failures(rep), failures(rep)
    run.close()


def test_pair_capacity_batch_and_server(fd, oracle, wl):
    """A batch instance whose arena is too small reports FD_ERR_CAPACITY
    without disturbing the other instances; a served request that overflows
    completes with FD_ERR_CAPACITY, the others with their oracle bits."""
    scene = wl.spheres(80, 7)
    ctx = fd.Context()
    ctx.set_backface_cull(1)
    ctx.set_scenes_esp(1e-3)
    verts, idx = fd.DeviceBuffer.from_array(scene.verts), fd.DeviceBuffer.from_array(scene.idx)
    ctx.add_model(verts, scene.nverts, 16, idx, scene.ntris)
    specs = [(128, 128, 16, (0.2, 0.1, 5.5)), (160, 120, 32, (-0.4, 0.3, 6.0)), (96, 96, 8, (0.0, 0.0, 5.0))]
    items, oracles = [], []
    for (w, h, R, c) in specs:
        rx = wl.receivers_heightfield(w, h, 8.0, 8.0 * h / w)
        light = wl.LightDesc(1.0, 1.0, c, res_w=R, res_h=R)
        o = oracle.run(wl.Frame("o", scene, rx, light, eps=1e-3, backface_cull=1))
        assert o.exit == 0
        oracles.append(o)
        items.append(dict(w=w, h=h, rx=rx, light=light, pos=fd.DeviceBuffer.from_array(rx.pos),
                          nrm=fd.DeviceBuffer.from_array(rx.nrm), shadow=fd.DeviceBuffer(w * h * 4),
                          masks=fd.DeviceBuffer(w * h * fd.mask_words(R) * 4), max_pairs=0))
    items[1]["max_pairs"] = int(oracles[1].scalars["pairs"]) // 2
    b = fd.Batch(ctx, items)
    b.launch()
    fd.lib().fd_device_sync()
    for k, it in enumerate(items):
        rc, st = b.stats(k)
        if k == 1:
            assert rc == fd.FD_ERR_CAPACITY and st.exit_code == fd.EXIT_CAPACITY
            assert not it["shadow"].download(np.float32).any() and not it["masks"].download(np.uint32).any()
        else:
            assert rc == 0
            assert np.array_equal(it["shadow"].download(np.float32).view(np.uint32), oracles[k].shadow.view(np.uint32))
            assert np.array_equal(it["masks"].download(np.uint32), oracles[k].masks)
    b.close()
    ctx.close()
    # server: one arena size for every request; the request with the most
    # pairs overflows it
    pmax = max(int(o.scalars["pairs"]) for o in oracles)
    cap = sorted(int(o.scalars["pairs"]) for o in oracles)[-2]
    assert cap < pmax
    srv = fd.Server(budget_bytes=1 << 30, streams=2, max_batch=3, eps=1e-3, max_pairs=cap)
    srv.add_scene(verts, scene.nverts, 16, idx, scene.ntris)
    for k, it in enumerate(items):
        it["shadow"].upload(np.full(it["w"] * it["h"], -2.0, np.float32))
        srv.submit(0, it["pos"], it["nrm"], it["w"], it["h"], it["light"], it["shadow"], it["masks"], tag=k)
    srv.drain(60000)
    done = {c.user_tag: c for c in srv.poll()}
    assert sorted(done) == [0, 1, 2]
    for k, it in enumerate(items):
        if int(oracles[k].scalars["pairs"]) > cap:
            assert done[k].status == fd.FD_ERR_CAPACITY and done[k].exit_code == fd.EXIT_CAPACITY
            assert not it["shadow"].download(np.float32).any()
        else:
            assert done[k].status == 0
            assert np.array_equal(it["shadow"].download(np.float32).view(np.uint32), oracles[k].shadow.view(np.uint32))
            assert np.array_equal(it["masks"].download(np.uint32), oracles[k].masks)
    srv.close()


def test_unsupported_light_grid(fd, oracle, wl):
    frame = wl.plane_256(ntri=1000, res=64)
    run = fd.FrameRunner(frame)
    rc, st = run.run()
    assert rc == 0
    lib = fd.lib()
    for (w, h) in ((16, 8), (0, 0), (65, 65), (32, 31)):
        L = fd.AreaLight.from_desc(wl.LightDesc(0.5, 0.5, (0.0, 0.0, 4.0), res_w=w, res_h=h))
        assert lib.fd_set_area_light(run.ctx.h, L) == fd.FD_ERR_UNSUPPORTED_LIGHT
        rc, st = run.run()
        assert rc == fd.FD_ERR_UNSUPPORTED_LIGHT, (w, h, rc)
    # a supported light again: the frame is the oracle's
    assert lib.fd_set_area_light(run.ctx.h, fd.AreaLight.from_desc(frame.light)) == 0
    rep, _, _ = compare(fd, oracle, frame, fr=run)
    assert not failures(rep), failures(rep)
    run.close()


@pytest.mark.parametrize("light_res", [32, 40])
def test_counting_kernel(fd, oracle, wl, light_res):
    """The profiling frame (fd_set_profiling count_ops) runs a counting
    variant of the occlusion kernel with one wave per work item.  Its masks
    equal the timed kernel's and the oracle's, and in the reference's walk
    order (ascending triangle id) its counters equal the oracle's loop
    counts: candidate tests, cull passes, planes set up and rows visited by
    the reference.  In the default (area) order the masks stay the same."""
    frame = wl.heightfield_1k(res=384, nsph=80, light_res=light_res)
    o = oracle.run(frame)
    assert o.exit == 0
    run = fd.FrameRunner(frame)
    rc, _ = run.run()
    assert rc == 0
    sh0, mk0 = (x.copy() for x in run.outputs())
    run.ctx.set_profiling(stage_events=True, count_ops=True)
    rc, _ = run.run()
    assert rc == 0
    sh, mk = run.outputs()
    assert np.array_equal(sh.view(np.uint32), sh0.view(np.uint32))
    assert np.array_equal(mk, mk0)
    run.ctx.set_triangle_order(fd.TRI_ORDER_LIST)
    rc, st = run.run()
    assert rc == 0
    assert st.work_items > 2048   # unsplit triangle lists: each receiver meets its whole list
    _, ops = run.ctx.get_profile(ops=True)
    sh, mk = run.outputs()
    assert np.array_equal(sh.view(np.uint32), sh0.view(np.uint32))
    assert np.array_equal(mk, mk0)
    assert np.array_equal(sh.view(np.uint32), o.shadow.view(np.uint32))
    assert np.array_equal(mk, o.masks)
    got = dict(cand=ops[0], passes=ops[1], rows_ref=ops[6], planes=ops[7])
    want = dict(cand=int(o.scalars["cand_tests"]), passes=int(o.scalars["cull_pass"]),
                rows_ref=int(o.scalars["row_visits"]), planes=int(o.scalars["plane_tris"]))
    assert got == want, (got, want)
    assert ops[2] <= ops[6]   # rows the kernel visits never exceed the reference's
    run.ctx.set_profiling(stage_events=False, count_ops=False)
    run.ctx.set_triangle_order(fd.TRI_ORDER_AUTO)
    rc, _ = run.run()
    assert rc == 0
    run.close()


def test_knob_arguments_are_checked(fd, wl):
    """Out-of-range schedule knobs are refused with FD_ERR_ARG (context
    setters, batch instances, the autotuner's candidates) and leave the
    context usable."""
    frame = wl.plane_256(ntri=500, res=64, light_res=8)
    run = fd.FrameRunner(frame)
    lib, h = fd.lib(), run.ctx.h
    for f, bad in ((lib.fd_set_list_split, (-1, 65)), (lib.fd_set_triangle_order, (-1, 4)),
                   (lib.fd_set_occlusion_waves, (1, 3, 16))):
        for v in bad:
            assert f(h, v) == fd.FD_ERR_ARG, (f, v)
    item = dict(pos=run.pos, nrm=run.nrm, w=64, h=64, light=frame.light, shadow=run.shadow, masks=run.masks)
    for key, v in (("split", 65), ("order", 4), ("waves", 3)):
        with pytest.raises(fd.FdError, match="fd_batch_create"):
            fd.Batch(run.ctx, [dict(item, **{key: v})])
    with pytest.raises(fd.FdError, match="fd_autotune"):
        run.ctx.autotune(item, [(16, 8, 0, 0, 65)], reps=1)
    rc, st = run.run()
    assert rc == 0 and st.exit_code == 0
    run.close()
