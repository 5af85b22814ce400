"""Tilted area lights: the general light pose of FD::FDsetAreaLight
(FD/api.cpp:328-336), as the reference application builds it from a
rotation matrix (Render/Render.cpp:446-449), bit-exact against the oracle
stage by stage (MI355X only).

With view (0, 0, -1) and up (0, 1, 0) the light axes are x and y and the
occlusion kernel forms the edge-plane coefficients from one product each;
any other pose takes its general dot-product path.  Both row widths (32-bit
and 64-bit rows) and the batched and serving forms are covered.
"""
import numpy as np
import pytest

from parity import compare, failures

pytestmark = pytest.mark.gpu

POSES = {
    "tilt-a": ((0.3, -0.2, -1.0), (0.2, 1.0, 0.1)),
    "tilt-b": ((-0.25, 0.4, -1.0), (1.0, 0.3, 0.0)),
}


def _rot(ax, ay, az):
    """Rotation matrix R = Rz Ry Rx (radians), as a light pose source."""
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _pose_from_rotation(R):
    """view = R (0, 0, -1), up = R (0, 1, 0) (Render.cpp:446-449 rotates the
    light's default frame)."""
    return tuple(float(x) for x in R @ [0, 0, -1]), tuple(float(x) for x in R @ [0, 1, 0])


def _tilt(frame, pose):
    frame.light.view_dir, frame.light.up_dir = pose
    return frame


def _check(fd, oracle, frame, subset=None):
    L = frame.light
    print(f"[tilted] {frame.name} {frame.rx.w}x{frame.rx.h} light {L.res_w} view {L.view_dir} up {L.up_dir}",
          flush=True)
    rep, fr, o = compare(fd, oracle, frame, subset=subset)
    bad = failures(rep)
    fr.close()
    assert not bad, f"mismatches {bad} (oracle exit {o.exit})"
    assert o.exit == 0 and o.scalars["pairs"] > 0
    return o


@pytest.mark.parametrize("light_res", [16, 32, 64])
@pytest.mark.parametrize("pose", sorted(POSES))
def test_plane256_tilted(fd, oracle, wl, pose, light_res):
    fr = _tilt(wl.plane_256(ntri=6000, res=160, light_res=light_res), POSES[pose])
    o = _check(fd, oracle, fr)
    assert float(o.shadow.mean()) > 0.001


@pytest.mark.parametrize("light_res", [32, 64])
def test_heightfield_192_tilted(fd, oracle, wl, light_res):
    fr = _tilt(wl.heightfield_1k(res=192, nsph=60, light_res=light_res), POSES["tilt-a"])
    _check(fd, oracle, fr)


def test_rotation_matrix_pose(fd, oracle, wl):
    pose = _pose_from_rotation(_rot(0.35, -0.2, 0.6))
    fr = _tilt(wl.heightfield_1k(res=160, nsph=50, light_res=24), pose)
    fr.light.position = (0.8, -0.5, 5.5)
    _check(fd, oracle, fr)


def test_tilted_light_in_batch(fd, oracle, wl):
    """One tilted instance among axial ones in a batched launch (direct and
    replayed from a HIP graph): its bits equal the oracle's."""
    scene = wl.spheres(80, 7)
    ctx = fd.Context()
    ctx.set_backface_cull(1)
    ctx.set_scenes_esp(1e-3)
    verts, idx = fd.DeviceBuffer.from_array(scene.verts), fd.DeviceBuffer.from_array(scene.idx)
    ctx.add_model(verts, scene.nverts, 16, idx, scene.ntris)
    specs = [(128, 96, 32, (0.4, -0.2, 5.5), ((0, 0, -1), (0, 1, 0))),
             (160, 160, 32, (0.1, 0.3, 6.0), POSES["tilt-a"]),
             (96, 128, 64, (-0.3, 0.2, 6.5), POSES["tilt-b"]),
             (64, 64, 16, (0.0, 0.0, 5.0), ((0, 0, -1), (0, 1, 0)))]
    items = []
    for (w, h, R, c, (view, up)) in specs:
        rx = wl.receivers_heightfield(w, h, 8.0, 8.0 * h / w)
        light = wl.LightDesc(1.0, 1.0, c, view, up, R, R)
        items.append(dict(w=w, h=h, rx=rx, light=light, pos=fd.DeviceBuffer.from_array(rx.pos),
                          nrm=fd.DeviceBuffer.from_array(rx.nrm), shadow=fd.DeviceBuffer(w * h * 4),
                          masks=fd.DeviceBuffer(w * h * fd.mask_words(R) * 4)))
    stream = fd.Stream()
    b = fd.Batch(ctx, items)
    expect = []
    for it in items:
        o = oracle.run(wl.Frame("o", scene, it["rx"], it["light"], eps=1e-3, backface_cull=1))
        assert o.exit == 0
        expect.append((o.shadow.view(np.uint32).copy(), o.masks.copy()))
    for mode in ("direct", "graph", "replay"):
        for it in items:
            it["shadow"].upload(np.full(it["w"] * it["h"], -3.0, np.float32))
        if mode == "graph":
            b.capture(stream.handle)
        b.launch(stream.handle)
        stream.sync()
        for k, it in enumerate(items):
            assert np.array_equal(it["shadow"].download(np.float32).view(np.uint32), expect[k][0]), (mode, k)
            assert np.array_equal(it["masks"].download(np.uint32), expect[k][1]), (mode, k)
    b.close()
    stream.close()
    ctx.close()


def test_tilted_light_in_server(fd, oracle, wl):
    """Tilted requests through the serving loop, mixed with axial ones, over
    two resident scenes: each result equals the oracle's."""
    scenes = [wl.spheres(60, 11), wl.spheres(90, 12)]
    srv = fd.Server(budget_bytes=1 << 30, streams=2, max_batch=3, eps=1e-3)
    keep = []
    for sc in scenes:
        v, i = fd.DeviceBuffer.from_array(sc.verts), fd.DeviceBuffer.from_array(sc.idx)
        keep += [v, i]
        srv.add_scene(v, sc.nverts, 16, i, sc.ntris)
    reqs = [(0, 128, 128, 32, (0.2, 0.1, 5.5), POSES["tilt-a"]), (1, 96, 160, 16, (-0.5, 0.3, 6.0), POSES["tilt-b"]),
            (0, 160, 120, 64, (0.0, -0.4, 6.2), POSES["tilt-b"]), (1, 128, 96, 32, (0.3, 0.3, 5.0), ((0, 0, -1), (0, 1, 0)))]
    jobs = []
    for k, (sc, w, h, R, c, (view, up)) in enumerate(reqs):
        rx = wl.receivers_heightfield(w, h, 8.0, 8.0 * h / w)
        light = wl.LightDesc(1.0, 1.0, c, view, up, R, R)
        j = dict(sc=sc, rx=rx, light=light, pos=fd.DeviceBuffer.from_array(rx.pos),
                 nrm=fd.DeviceBuffer.from_array(rx.nrm), shadow=fd.DeviceBuffer(w * h * 4),
                 masks=fd.DeviceBuffer(w * h * fd.mask_words(R) * 4))
        jobs.append(j)
        srv.submit(sc, j["pos"], j["nrm"], w, h, light, j["shadow"], j["masks"], tag=k)
    srv.drain(60000)
    done = srv.poll()
    assert sorted(c.user_tag for c in done) == list(range(len(jobs)))
    assert all(c.status == 0 for c in done)
    for k, j in enumerate(jobs):
        o = oracle.run(wl.Frame("o", scenes[j["sc"]], j["rx"], j["light"], eps=1e-3, backface_cull=1))
        assert o.exit == 0
        assert np.array_equal(j["shadow"].download(np.float32).view(np.uint32), o.shadow.view(np.uint32)), k
        assert np.array_equal(j["masks"].download(np.uint32), o.masks), k
    srv.close()
