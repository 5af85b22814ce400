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
