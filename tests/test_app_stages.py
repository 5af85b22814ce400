"""CPU pins of the application stages on either side of the visibility path:
the host G-buffer ray cast (the device producer's bit-exact reference) and
the oracle's restatement of the shading pass (phoneShading.shf:23-61)."""
import numpy as np
import pytest


def _rays(sc):
    """Camera rays of every pixel centre in float64 (the generator's formula)."""
    eye, at, up = (np.asarray(v, np.float64) for v in (sc.eye, sc.at, sc.up))
    nz = lambda v: v / np.sqrt((v * v).sum(-1, keepdims=True))
    f = nz(at - eye)
    s = nz(np.cross(f, up))
    u = np.cross(s, f)
    th = np.tan(0.5 * sc.fovy_deg * np.pi / 180.0)
    x = (np.arange(sc.width) + 0.5) / sc.width * 2.0 - 1.0
    y = 1.0 - (np.arange(sc.height) + 0.5) / sc.height * 2.0
    X, Y = np.meshgrid(x * th * sc.width / sc.height, y * th)
    d = f + s * X[..., None] + u * Y[..., None]
    return eye, nz(d)


def _brute_force(sc):
    """Independent slab-test ray cast (numpy, float64): hit distance and
    object id (-1 background, 0 ground, 1 + b box b) per pixel."""
    eye, d = _rays(sc)
    H, W = sc.height, sc.width
    tbest = np.full((H, W), np.inf)
    obj = np.full((H, W), -1)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = -eye[2] / d[..., 2]
        p = eye + d * t[..., None]
        g = (d[..., 2] < 0) & (np.abs(p[..., 0]) <= sc.ground_half) & (np.abs(p[..., 1]) <= sc.ground_half)
        tbest[g] = t[g]
        obj[g] = 0
        for b, bx in enumerate(sc.boxes):
            ta = (bx[:3] - eye) / d
            tb = (bx[3:] - eye) / d
            t0 = np.minimum(ta, tb).max(-1)
            t1 = np.maximum(ta, tb).min(-1)
            hit = (t0 <= t1) & (t0 > 0) & (t0 < tbest)
            tbest[hit] = t0[hit]
            obj[hit] = 1 + b
    return tbest, obj, eye, d


def test_gbuffer_host_matches_brute_force(wl):
    sc = wl.gbuffer_scene_boxes(160, 90, 20, 3)
    g0, g1, g2, g3 = wl.gbuffer_boxes_host(sc)
    tbest, obj, eye, d = _brute_force(sc)
    bg = np.isnan(g0[..., 0])
    # background: NaN in every channel of every target
    assert np.array_equal(bg, obj < 0)
    for g in (g0, g1, g2, g3):
        assert np.isnan(g[bg]).all()
    fg = ~bg
    # positions on the camera ray at the brute-force hit distance
    p = eye + d * tbest[..., None]
    assert np.allclose(g0[fg][:, :3], p[fg], rtol=0, atol=1e-5)
    # normals: ground +z; boxes an axis direction facing the camera
    n = g1[..., :3]
    assert np.array_equal(n[obj == 0], np.tile([0, 0, 1], ((obj == 0).sum(), 1)).astype(np.float32))
    nb = n[obj > 0]
    assert np.array_equal(np.abs(nb).sum(-1), np.ones(len(nb), np.float32))
    assert (np.einsum("ij,ij->i", nb.astype(np.float64), d[obj > 0]) < 0).all()
    # materials follow the object id (gatherData.shf:41-44 packing)
    m = sc.materials[obj[fg]]
    assert np.array_equal(g2[fg][:, :3], m[:, 0:3]) and np.array_equal(g3[fg][:, :3], m[:, 3:6])
    assert np.array_equal(np.stack([g0[fg][:, 3], g1[fg][:, 3], g2[fg][:, 3]], 1), m[:, 6:9])
    assert np.array_equal(g3[fg][:, 3], m[:, 9])
    assert 0.2 < fg.mean() < 0.9 and (obj > 0).sum() > 100


def test_gbuffer_host_receivers_are_the_workload_receivers(wl):
    sc = wl.gbuffer_scene_boxes(96, 54, 20, 3)
    g0, g1, _, _ = wl.gbuffer_boxes_host(sc)
    rx = wl.receivers_boxes(96, 54, 20, 3)
    assert np.array_equal(g0[..., :3].view(np.uint32), rx.pos[..., :3].view(np.uint32))
    assert np.array_equal(g1[..., :3].view(np.uint32), rx.nrm[..., :3].view(np.uint32))


def _texel(P, N, ka, kd, ks, shin):
    g0 = np.array([[*P, ks[0]]], np.float32)
    g1 = np.array([[*N, ks[1]]], np.float32)
    g2 = np.array([[*ka, ks[2]]], np.float32)
    g3 = np.array([[*kd, shin]], np.float32)
    return g0, g1, g2, g3


def test_shade_closed_forms(oracle):
    La, Ld, Ls = [0.2, 0.3, 0.4], [1.0, 0.5, 0.25], [0.5, 0.5, 0.5]
    ka, kd, ks = [0.1, 0.2, 0.3], [0.6, 0.7, 0.8], [0.25, 0.5, 0.75]
    g = _texel([0, 0, 0], [0, 0, 3], ka, kd, ks, 8.0)
    # light and camera along the (unnormalised) normal: L.N = 1, R = N, R.V = 1
    c = oracle.shade([0.0], *g, [[0, 0, 2]], [0, 0, 5], La, Ld, Ls)[0]
    want = np.array(La) * ka + np.array(Ld) * kd + np.array(Ls) * ks
    assert np.allclose(c[:3], want, rtol=1e-6) and c[3] == 1.0
    # fully shadowed: ambient only
    c = oracle.shade([1.0], *g, [[0, 0, 2]], [0, 0, 5], La, Ld, Ls)[0]
    assert np.allclose(c[:3], np.array(La) * ka, rtol=1e-6)
    # light behind the surface: no diffuse; R.V < 0 -> no specular
    c = oracle.shade([0.0], *g, [[0, 0, -2]], [0, 0, 5], La, Ld, Ls)[0]
    assert np.allclose(c[:3], np.array(La) * ka, rtol=1e-6)
    # 60 degrees off the normal: diffuse = cos 60; mirror camera -> R.V = 1
    s, co = np.sin(np.pi / 3), np.cos(np.pi / 3)
    c = oracle.shade([0.25], *g, [[s, 0, co]], [-2 * s, 0, 2 * co], La, Ld, Ls)[0]
    want = np.array(La) * ka + (np.array(Ld) * kd * co + np.array(Ls) * ks) * 0.75
    assert np.allclose(c[:3], want, rtol=2e-6)
    # NaN intensity: the texel keeps the cleared colour
    c = oracle.shade([np.nan], *g, [[0, 0, 2]], [0, 0, 5], La, Ld, Ls)[0]
    assert np.array_equal(c, np.zeros(4, np.float32))


def test_shade_sample_average(oracle):
    rng = np.random.default_rng(5)
    n = 64
    P = rng.uniform(-1, 1, (n, 3))
    N = rng.normal(size=(n, 3))
    g0 = np.concatenate([P, rng.uniform(0, 1, (n, 1))], 1)
    g1 = np.concatenate([N, rng.uniform(0, 1, (n, 1))], 1)
    g2 = rng.uniform(0, 1, (n, 4))
    g3 = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.uniform(1, 32, (n, 1))], 1)
    sh = rng.uniform(0, 1, n)
    S = rng.uniform(-2, 2, (3, 3)) + [0, 0, 4]
    args = ([0, 0, 6], [0.1, 0.1, 0.1], [0.8, 0.8, 0.8], [0.4, 0.4, 0.4])
    both = oracle.shade(sh, g0, g1, g2, g3, S, *args)
    each = [oracle.shade(sh, g0, g1, g2, g3, S[k:k + 1], *args) for k in range(3)]
    assert np.allclose(both, sum(each) / 3, rtol=1e-5, atol=1e-6)
