"""The application stages on the GPU (MI355X only): the G-buffer producer is
bit-identical to the host ray cast; the shading pass agrees with the
oracle's restatement of phoneShading.shf within float tolerance (powf
differs by an ulp between the device and host libraries; every other
operation is IEEE); and the device G-buffer drives the visibility path to
the oracle's bits end to end."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHADE_RTOL, SHADE_ATOL = 2e-5, 1e-6


def _produce(fd, sc, targets=4, materials=True):
    n = sc.width * sc.height
    g = [fd.DeviceBuffer(n * 16) for _ in range(targets)]
    strm = fd.Stream()
    fd.gbuffer_boxes(sc.width, sc.height, sc.boxes, sc.materials if materials else None, sc.eye, sc.at, sc.up,
                     sc.fovy_deg, sc.ground_half, g[0], g[1], g[2] if targets > 2 else None,
                     g[3] if targets > 3 else None, stream=strm)
    strm.sync()
    strm.close()
    return g


@pytest.mark.parametrize("w,h,nbox", [(1920, 1080, 20), (333, 187, 7), (64, 48, 0)])
def test_gbuffer_bit_exact(fd, wl, w, h, nbox):
    print(f"[gbuffer] {w}x{h} boxes {nbox}", flush=True)
    sc = wl.gbuffer_scene_boxes(w, h, nbox, 3)
    ref = wl.gbuffer_boxes_host(sc)
    g = _produce(fd, sc)
    for k in range(4):
        got = g[k].download(np.float32).reshape(h, w, 4)
        assert np.array_equal(got.view(np.uint32), ref[k].view(np.uint32)), f"target g{k}"


@pytest.mark.parametrize("k", range(6))
def test_gbuffer_random_cameras_bit_exact(fd, wl, k):
    """Seeded random boxes scenes and cameras (eye, target, up, field of
    view, ground size): all four targets bit-identical to the host ray cast."""
    r = np.random.default_rng(300 + k)
    w, h, nbox = int(r.integers(32, 400)), int(r.integers(32, 300)), int(r.integers(0, 31))
    sc = wl.gbuffer_scene_boxes(w, h, nbox, int(r.integers(1, 1000)))
    ang = float(r.uniform(0, 2 * np.pi))
    dist, height = float(r.uniform(8, 18)), float(r.uniform(3, 14))
    sc.eye = (dist * np.cos(ang), dist * np.sin(ang), height)
    sc.at = (float(r.uniform(-2, 2)), float(r.uniform(-2, 2)), float(r.uniform(0, 1)))
    sc.up = (float(r.uniform(-0.2, 0.2)), float(r.uniform(-0.2, 0.2)), 1.0)
    sc.fovy_deg = float(r.uniform(30, 80))
    sc.ground_half = float(r.uniform(6, 14))
    print(f"[gbuffer] random {k}: {w}x{h} boxes {nbox} eye {sc.eye}", flush=True)
    ref = wl.gbuffer_boxes_host(sc)
    g = _produce(fd, sc)
    for t in range(4):
        got = g[t].download(np.float32).reshape(h, w, 4)
        assert np.array_equal(got.view(np.uint32), ref[t].view(np.uint32)), (k, f"target g{t}")


def test_gbuffer_position_normal_only(fd, wl):
    sc = wl.gbuffer_scene_boxes(200, 120, 20, 3)
    g = _produce(fd, sc, targets=2, materials=False)
    rx = wl.receivers_boxes(200, 120, 20, 3)
    assert np.array_equal(g[0].download(np.float32).reshape(120, 200, 4).view(np.uint32), rx.pos.view(np.uint32))
    assert np.array_equal(g[1].download(np.float32).reshape(120, 200, 4).view(np.uint32), rx.nrm.view(np.uint32))


def _light_corner2(L):
    """Render.cpp:451-466: every bound sample is the light's corner 2
    (position - right - top, with the light's axes)."""
    view = np.asarray(L.view_dir, np.float64)
    up = np.asarray(L.up_dir, np.float64)
    view /= np.linalg.norm(view)
    right = np.cross(view, up)
    right /= np.linalg.norm(right)
    top = np.cross(right, view)
    return np.asarray(L.position) - right * L.right_radius - top * L.top_radius


def test_pipeline_gbuffer_visibility_shade(fd, wl, oracle):
    """Device G-buffer -> visibility (receivers read straight from g0 / g1)
    -> shading, against the oracle on the host G-buffer."""
    w, h = 320, 180
    print(f"[pipeline] {w}x{h}", flush=True)
    sc = wl.gbuffer_scene_boxes(w, h, 20, 3)
    g = _produce(fd, sc)
    host = wl.gbuffer_boxes_host(sc)
    frame = wl.boxes_1080p(w=w, h=h, ntri=20000, light_res=16)
    frame.rx = wl.Receivers(host[0], host[1])
    frame.grid_log2 = 4
    o = oracle.run(frame)
    assert o.exit == 0
    # visibility on the device buffers
    ctx = fd.Context()
    ctx.set_backface_cull(frame.backface_cull)
    ctx.set_scenes_esp(frame.eps)
    verts = fd.DeviceBuffer.from_array(frame.scene.verts)
    idx = fd.DeviceBuffer.from_array(frame.scene.idx)
    ctx.add_model(verts, frame.scene.nverts, 16, idx, frame.scene.ntris)
    R = frame.light.res_w
    shadow = fd.DeviceBuffer(w * h * 4)
    masks = fd.DeviceBuffer(w * h * fd.mask_words(R) * 4)
    ctx.set_receivers(g[0], g[1], w, h)
    ctx.set_outputs(shadow, masks)
    ctx.set_area_light(frame.light)
    strm = fd.Stream()
    rc, st = ctx.launch(strm.handle)
    assert rc == 0 and st.exit_code == 0
    sh = shadow.download(np.float32)
    assert np.array_equal(sh.view(np.uint32), np.asarray(o.shadow, np.float32).view(np.uint32))
    assert np.array_equal(masks.download(np.uint32), np.asarray(o.masks, np.uint32).ravel())
    # shading: the reference's single bound sample, then four corner samples
    rgba = fd.DeviceBuffer(w * h * 16)
    cam, La, Ld, Ls = sc.eye, (0.2, 0.2, 0.2), (0.9, 0.85, 0.8), (0.5, 0.5, 0.5)
    c2 = _light_corner2(frame.light)
    for samples in ([c2], [c2, c2 + [0.3, 0, 0], c2 + [0, 0.4, 0], c2 + [0.3, 0.4, 0]]):
        fd.shade(w, h, samples, cam, La, Ld, Ls, shadow, g[0], g[1], g[2], g[3], rgba, stream=strm)
        strm.sync()
        got = rgba.download(np.float32).reshape(-1, 4)
        want = oracle.shade(sh, host[0].reshape(-1, 4), host[1].reshape(-1, 4), host[2].reshape(-1, 4),
                            host[3].reshape(-1, 4), samples, cam, La, Ld, Ls)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.allclose(got[ok], want[ok], rtol=SHADE_RTOL, atol=SHADE_ATOL)
        assert ok.any(axis=1).mean() > 0.3
    strm.close()
    ctx.close()


@pytest.mark.parametrize("k", range(4))
def test_shade_random_inputs(fd, wl, oracle, k):
    """Seeded random G-buffer scene, shadow values, 1-8 light samples, camera
    and light colours: the shading pass against the oracle's restatement of
    phoneShading.shf, within the same tolerance."""
    r = np.random.default_rng(500 + k)
    w, h = int(r.integers(40, 300)), int(r.integers(40, 200))
    sc = wl.gbuffer_scene_boxes(w, h, int(r.integers(0, 25)), int(r.integers(1, 1000)))
    host = wl.gbuffer_boxes_host(sc)
    g = _produce(fd, sc)
    sh = r.uniform(0, 1, w * h).astype(np.float32)
    sh[r.random(w * h) < 0.1] = 0.0
    sh[r.random(w * h) < 0.1] = 1.0
    shadow = fd.DeviceBuffer.from_array(sh)
    samples = [np.array([r.uniform(-5, 5), r.uniform(-5, 5), r.uniform(4, 10)]) for _ in range(int(r.integers(1, 9)))]
    cam = (float(r.uniform(-10, 10)), float(r.uniform(-15, -5)), float(r.uniform(4, 12)))
    La, Ld, Ls = (tuple(float(x) for x in r.uniform(0, 1, 3)) for _ in range(3))
    rgba = fd.DeviceBuffer(w * h * 16)
    strm = fd.Stream()
    print(f"[shade] random {k}: {w}x{h}, {len(samples)} samples", flush=True)
    fd.shade(w, h, samples, cam, La, Ld, Ls, shadow, g[0], g[1], g[2], g[3], rgba, stream=strm)
    strm.sync()
    strm.close()
    got = rgba.download(np.float32).reshape(-1, 4)
    want = oracle.shade(sh, host[0].reshape(-1, 4), host[1].reshape(-1, 4), host[2].reshape(-1, 4),
                        host[3].reshape(-1, 4), samples, cam, La, Ld, Ls)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert ok.any()
    assert np.allclose(got[ok], want[ok], rtol=SHADE_RTOL, atol=SHADE_ATOL)
