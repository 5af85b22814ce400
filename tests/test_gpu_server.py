"""Serving form: requests of uneven size over two resident scenes, submitted
while others are on the GPU, complete with results identical to the same
request run alone through fd_launch (MI355X only)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REQS = [(0, 96, 96, 16, (0.3, 0.2, 5.0)), (1, 160, 120, 32, (-0.8, 0.4, 6.0)), (0, 256, 256, 32, (0.0, 0.0, 6.0)),
        (1, 128, 128, 8, (1.0, -1.0, 4.5)), (0, 200, 150, 33, (0.5, 0.5, 7.0)), (1, 256, 256, 16, (-1.5, 1.2, 5.2)),
        (0, 64, 64, 64, (0.1, -0.4, 6.5)), (1, 96, 160, 32, (0.9, 0.9, 4.1)), (0, 128, 96, 16, (-0.2, 1.7, 5.9))]


def test_server_results_equal_alone(fd, wl):
    scenes = [wl.spheres(60, 11), wl.spheres(90, 12)]
    srv = fd.Server(budget_bytes=2 << 30, streams=3, max_batch=4, eps=1e-3)
    keep = []
    for sc in scenes:
        v, i = fd.DeviceBuffer.from_array(sc.verts), fd.DeviceBuffer.from_array(sc.idx)
        keep += [v, i]
        srv.add_scene(v, sc.nverts, 16, i, sc.ntris)
    jobs = []
    for k, (sc, w, h, R, c) in enumerate(REQS):
        rx = wl.receivers_heightfield(w, h, 8.0, 8.0 * h / w)
        light = wl.LightDesc(1.0, 1.0, c, res_w=R, res_h=R)
        job = dict(sc=sc, rx=rx, light=light, pos=fd.DeviceBuffer.from_array(rx.pos),
                   nrm=fd.DeviceBuffer.from_array(rx.nrm), shadow=fd.DeviceBuffer(w * h * 4),
                   masks=fd.DeviceBuffer(w * h * fd.mask_words(R) * 4))
        jobs.append(job)
    # staggered submissions: later requests arrive while earlier ones run
    for k, j in enumerate(jobs):
        srv.submit(j["sc"], j["pos"], j["nrm"], j["rx"].w, j["rx"].h, j["light"], j["shadow"], j["masks"], tag=k)
        time.sleep(0.001 * (k % 3))
    srv.drain(60000)
    done = srv.poll()
    assert sorted(c.user_tag for c in done) == list(range(len(jobs)))
    assert all(c.status == 0 for c in done)
    assert max(c.batch_size for c in done) >= 1
    for k, j in enumerate(jobs):
        fr = wl.Frame("alone", scenes[j["sc"]], j["rx"], j["light"], eps=1e-3, backface_cull=1)
        run = fd.FrameRunner(fr)
        rc, st = run.run()
        assert rc == 0
        sh0, mk0 = run.outputs()
        run.close()
        assert np.array_equal(j["shadow"].download(np.float32).view(np.uint32), sh0.view(np.uint32)), k
        assert np.array_equal(j["masks"].download(np.uint32), mk0), k
    srv.close()


def _rot_z_translate(deg, t):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = c, -s, s, c
    m[:3, 3] = t
    return m.T.ravel()   # column-major


def test_server_scene_moves_are_ordered(fd, wl):
    """FD::moveModel on a resident scene through the server: requests
    submitted before a move see the old vertices, later ones the moved
    vertices (cumulative, in place); other scenes are unaffected."""
    scenes = [wl.spheres(60, 11), wl.spheres(90, 12)]
    srv = fd.Server(budget_bytes=1 << 30, streams=2, max_batch=2, eps=1e-3)
    keep = []
    for sc in scenes:
        v, i = fd.DeviceBuffer.from_array(sc.verts), fd.DeviceBuffer.from_array(sc.idx)
        keep += [v, i]
        srv.add_scene(v, sc.nverts, 16, i, sc.ntris)
    rx = wl.receivers_heightfield(160, 120, 8.0, 6.0)
    light = wl.LightDesc(1.0, 1.0, (0.2, -0.3, 5.5), res_w=32, res_h=32)
    pos, nrm = fd.DeviceBuffer.from_array(rx.pos), fd.DeviceBuffer.from_array(rx.nrm)
    m1, m2 = _rot_z_translate(10.0, (0.3, -0.2, 0.1)), _rot_z_translate(-25.0, (-0.5, 0.4, 0.0))
    plan = ["A", "A", "C", "M1", "B", "C", "B", "M2", "D", "C"]
    outs = []
    for k, step in enumerate(plan):
        if step.startswith("M"):
            srv.move_scene(0, m1 if step == "M1" else m2, tag=1000 + k)
            outs.append(None)
            continue
        sh, mk = fd.DeviceBuffer(160 * 120 * 4), fd.DeviceBuffer(160 * 120 * fd.mask_words(32) * 4)
        srv.submit(1 if step == "C" else 0, pos, nrm, 160, 120, light, sh, mk, tag=k)
        outs.append((sh, mk))
    srv.drain(60000)
    done = {c.user_tag: c for c in srv.poll()}
    assert sorted(done) == sorted([k for k, s in enumerate(plan) if not s.startswith("M")] +
                                  [1000 + k for k, s in enumerate(plan) if s.startswith("M")])
    assert all(c.status == 0 for c in done.values())
    assert done[1003].batch_size == 0 and done[1007].batch_size == 0
    # expected: scene 0 alone at each geometry version (the same device transform)
    exp = {}
    run = fd.FrameRunner(wl.Frame("alone", scenes[0], rx, light, eps=1e-3, backface_cull=1))
    for version, m in (("A", None), ("B", m1), ("D", m2)):
        if m is not None:
            run.ctx.move_model(m)
        rc, _ = run.run()
        assert rc == 0
        exp[version] = tuple(x.copy() for x in run.outputs())
    run.close()
    runc = fd.FrameRunner(wl.Frame("alone", scenes[1], rx, light, eps=1e-3, backface_cull=1))
    rc, _ = runc.run()
    exp["C"] = tuple(x.copy() for x in runc.outputs())
    runc.close()
    assert not np.array_equal(exp["A"][1], exp["B"][1]) and not np.array_equal(exp["B"][1], exp["D"][1])
    for k, step in enumerate(plan):
        if step.startswith("M"):
            continue
        sh, mk = outs[k]
        assert np.array_equal(sh.download(np.float32).view(np.uint32), exp[step][0].view(np.uint32)), (k, step)
        assert np.array_equal(mk.download(np.uint32), exp[step][1]), (k, step)
    srv.close()


def test_server_tight_budget(fd, wl, oracle):
    """A 64 MiB workspace pool: a request that can never fit is refused at
    submission (FD_ERR_CAPACITY) and the others wait for workspace to free
    up; every admitted request completes with the oracle's bits."""
    scene = wl.tris_clusters(8000, 21, box=(-6, 6, -6, 6, 0.5, 4))
    srv = fd.Server(budget_bytes=64 << 20, streams=2, max_batch=4, eps=1e-3, max_pairs=1 << 20)
    v, i = fd.DeviceBuffer.from_array(scene.verts), fd.DeviceBuffer.from_array(scene.idx)
    srv.add_scene(v, scene.nverts, 16, i, scene.ntris)
    # 2M receivers need more workspace than the whole pool
    big = wl.receivers_heightfield(1920, 1080, 8.0, 4.5)
    bp, bn = fd.DeviceBuffer.from_array(big.pos), fd.DeviceBuffer.from_array(big.nrm)
    bs = fd.DeviceBuffer(1920 * 1080 * 4)
    with pytest.raises(fd.FdError, match=str(fd.FD_ERR_CAPACITY)):
        srv.submit(0, bp, bn, 1920, 1080, wl.LightDesc(1.0, 1.0, (0, 0, 6), res_w=16, res_h=16), bs, None, tag=99)
    jobs = []
    for k in range(8):
        w, h = 160 + 24 * k, 256 - 12 * k
        R = (8, 16, 32, 33)[k % 4]
        rx = wl.receivers_heightfield(w, h, 8.0, 8.0 * h / w)
        light = wl.LightDesc(1.0, 1.0, (0.2 * k - 0.7, 0.3, 5.0 + 0.2 * k), res_w=R, res_h=R)
        jobs.append(dict(rx=rx, light=light, pos=fd.DeviceBuffer.from_array(rx.pos), nrm=fd.DeviceBuffer.from_array(rx.nrm),
                         shadow=fd.DeviceBuffer(w * h * 4), masks=fd.DeviceBuffer(w * h * fd.mask_words(R) * 4),
                         o=oracle.run(wl.Frame(f"t{k}", scene, rx, light, eps=1e-3, backface_cull=1))))
    print("[server] tight budget: 8 requests", flush=True)
    for k, j in enumerate(jobs):
        srv.submit(0, j["pos"], j["nrm"], j["rx"].w, j["rx"].h, j["light"], j["shadow"], j["masks"], tag=k)
    srv.drain(60000)
    done = sorted(srv.poll(16), key=lambda c: c.user_tag)
    srv.close()
    assert [c.user_tag for c in done] == list(range(8)) and all(c.status == 0 for c in done)
    for k, j in enumerate(jobs):
        o = j["o"]
        assert np.array_equal(j["shadow"].download(np.float32).view(np.uint32),
                              np.asarray(o.shadow, np.float32).view(np.uint32)), k
        assert np.array_equal(j["masks"].download(np.uint32), o.masks.ravel()), k
