"""Independent pins of the oracle functions that the end-to-end checks leave
open (runs on the CPU):

  triangle back-face cull   FDc/Step0_triangleSetup.inl:100-116
  receiver back-face test   FDc/Step2_1_samplerRectangleAABB.inl:90-98
  light-plane bounds        FDc/Step2_3_defineFDbuffer.inl:22-63
  FD::moveModel             FDc/moveModel.inl:6-30
  exits 4 and 7             FDc/Step2_3_defineFDbuffer.inl:45-47, FD/api.cpp:522

Each is checked against something the oracle does not restate: a float64
classification with a rounding margin (ambiguous cases skipped), numpy
reductions over the oracle's per-element outputs, exact special cases of a
4x4 transform, or a constructed input whose exit follows from geometry.
"""
import numpy as np
import pytest

F32_MAX = np.float32(3.402823466e+38)


def _corners(o):
    lb = o.light_bits.astype(np.float64)
    return lb[28:40].reshape(4, 3)   # ur, ul, dl, dr


# --------------------------------------------------------------------------
# triangle back-face cull: culled iff n = cross(v0 - v2, v1 - v0) has a
# non-positive dot product with (corner - v_i) for all 4 light corners and
# all 3 vertices (12 tests)
# --------------------------------------------------------------------------
def _cull_scene(rng, corners, n_random=6000, n_through=6000):
    tris = []
    # random orientation: centroids in the receivers' light cone
    c = np.column_stack([rng.uniform(-2.5, 2.5, n_random), rng.uniform(-2.5, 2.5, n_random),
                         rng.uniform(0.5, 3.0, n_random)])
    for k in range(n_random):
        e = rng.normal(size=(3, 3)) * rng.uniform(0.05, 0.4)
        tris.append(c[k] + e)
    # planes through a point of the light rectangle (corners on both sides),
    # including points near one corner (one corner separated from three)
    ur, ul, dl, dr = corners
    for k in range(n_through):
        s, t = rng.uniform(-0.2, 1.2, 2)
        if k % 3 == 0:   # near a corner
            s, t = rng.choice([0.0, 1.0]) + rng.normal(0, 0.05), rng.choice([0.0, 1.0]) + rng.normal(0, 0.05)
        p = ur + s * (ul - ur) + t * (dr - ur)
        d = np.array([rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.8, 3.2)])
        base = p - d                      # below the light, towards the receivers
        u = d / np.linalg.norm(d)
        w = np.cross(u, rng.normal(size=3))
        w /= np.linalg.norm(w)
        a, b = rng.uniform(0.05, 0.5, 2)
        v0 = base
        v1 = base + a * u * (1 if k % 2 else -1)
        v2 = base + b * w
        tris.append(np.array([v0, v1, v2]))
    T = np.asarray(tris, np.float32)
    T[:, :, 2] = np.clip(T[:, :, 2], 0.3, 3.7)   # stay between receivers (z=0) and light (z=4)
    return T


def test_triangle_cull_float64(wl, oracle):
    rng = np.random.default_rng(21)
    fr = wl.plane_256(ntri=10, res=64)
    o0 = oracle.run(fr)
    corners = _corners(o0)
    T = _cull_scene(rng, corners)
    nt = T.shape[0]
    fr.scene = wl.Scene(np.concatenate([T.reshape(-1, 3), np.ones((3 * nt, 1), np.float32)], 1),
                        np.arange(3 * nt, dtype=np.int32).reshape(nt, 3))
    fr.backface_cull = 1
    o = oracle.run(fr)
    assert o.exit in (0, 7), o.exit
    aabb = o.tri_aabb.reshape(nt, 8)
    culled = (aabb[:, 0] == F32_MAX) & (aabb[:, 4] == -F32_MAX)
    V = T.astype(np.float64)
    e0 = V[:, 1] - V[:, 0]
    e1 = V[:, 0] - V[:, 2]
    n = np.cross(e1, e0)
    d = np.einsum("tj,tcvj->tcv", n, corners[None, :, None, :] - V[:, None, :, :])   # (T, 4 corners, 3 verts)
    mag = (np.linalg.norm(e0, axis=1) * np.linalg.norm(e1, axis=1))[:, None, None] * \
        np.linalg.norm(corners[None, :, None, :] - V[:, None, :, :], axis=3)
    margin = 1e-5 * mag
    sure_back = np.all(d < -margin, axis=(1, 2))
    sure_front = np.any(d > margin, axis=(1, 2))
    decided = sure_back | sure_front
    assert decided.sum() >= 10_000
    assert np.array_equal(culled[decided], sure_back[decided]), \
        np.nonzero(culled[decided] != sure_back[decided])[0][:10]
    # the set holds every class a plausible slip would confuse: culled,
    # facing all corners, and planes through the light (corners on both
    # sides, among them a single corner in front)
    front = (d > margin).reshape(nt, 12)
    ncorner_front = np.array([np.any(d[t] > margin[t], axis=1).sum() for t in range(nt)])
    assert sure_back.sum() > 2000 and (ncorner_front == 4).sum() > 2000
    assert ((ncorner_front > 0) & (ncorner_front < 4) & decided).sum() > 1000
    assert ((ncorner_front == 1) & decided).sum() > 100
    assert front.any()
    # with the cull off nothing is culled
    fr.backface_cull = 0
    o2 = oracle.run(fr)
    a2 = o2.tri_aabb.reshape(nt, 8)
    assert not np.any((a2[:, 0] == F32_MAX) & (a2[:, 4] == -F32_MAX))


# --------------------------------------------------------------------------
# receiver back-face test: back-facing iff dot(corner - p, n) <= 0 for all
# four light corners
# --------------------------------------------------------------------------
def test_receiver_backface_float64(wl, oracle):
    rng = np.random.default_rng(22)
    fr = wl.plane_256(ntri=300, res=128)
    o0 = oracle.run(fr)
    corners = _corners(o0)
    P = fr.rx.pos[..., :3].reshape(-1, 3).astype(np.float64)
    n = rng.normal(size=P.shape)
    # a third of the normals nearly perpendicular to the direction of a
    # random point of the light rectangle: corners on both sides
    k = rng.random(len(P)) < 0.35
    s, t = rng.uniform(-0.3, 1.3, (2, len(P)))
    q = corners[0] + s[:, None] * (corners[1] - corners[0]) + t[:, None] * (corners[3] - corners[0])
    dirq = q - P
    dirq /= np.linalg.norm(dirq, axis=1, keepdims=True)
    perp = n - (n * dirq).sum(1, keepdims=True) * dirq
    n[k] = perp[k]
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    fr.rx.nrm[..., :3] = n.reshape(fr.rx.nrm[..., :3].shape).astype(np.float32)
    o = oracle.run(fr)
    assert o.exit in (0, 5, 6, 7)
    back = o.backface.astype(bool)
    N = fr.rx.nrm[..., :3].reshape(-1, 3).astype(np.float64)
    d = np.einsum("ij,icj->ic", N, corners[None] - P[:, None])
    margin = 1e-5 * np.linalg.norm(N, axis=1)[:, None] * np.linalg.norm(corners[None] - P[:, None], axis=2)
    sure_back = np.all(d < -margin, axis=1)
    sure_front = np.any(d > margin, axis=1)
    decided = sure_back | sure_front
    assert decided.mean() > 0.99
    assert np.array_equal(back[decided], sure_back[decided])
    nfront = (d > margin).sum(1)
    assert sure_back.sum() > 1000 and (nfront == 4).sum() > 1000
    assert ((nfront > 0) & (nfront < 4)).sum() > 1000     # mixed: "any" vs "all" would differ here
    assert (nfront == 1).sum() > 50
    # back-facing receivers: shadow 1 and every mask bit set (:97)
    assert np.all(o.shadow[back] == 1.0)


# --------------------------------------------------------------------------
# light-plane bounds: far plane, receiver / vertex rectangle AABBs and their
# intersection, recomputed with numpy from the oracle's per-element
# rectangles (makeRectangle is pinned on its own in test_oracle.py)
# --------------------------------------------------------------------------
def test_light_plane_bounds_numpy(wl, oracle):
    fr = wl.plane_256(ntri=3000, res=96)
    # occluders shifted in +x and stretched in y: the light plane takes its
    # x begin from the vertices, x end and both y bounds from the receivers
    v = fr.scene.verts
    v[:, 0] += 2.5
    v[:, 1] *= 1.8
    rng = np.random.default_rng(5)
    fr.rx.pos[rng.random(fr.rx.pos.shape[:2]) < 0.1, 0] = np.nan   # some empty receivers
    o = oracle.run(fr)
    assert o.exit == 0
    lb = o.light_bits
    mat, F = lb[0:16], lb[40:42].astype(np.float32)
    f, S = np.float32(lb[42]), lb[43:45].astype(np.float32)
    # vertex rectangles
    V = v[:, :3]
    rc = np.array([oracle.make_rectangle(mat, p) for p in V], np.float32)
    rcx, rcy, rcz = rc[:, 0], rc[:, 1], rc[:, 2]
    vmaxx = (rcx + F[0] / rcz) - S[0]
    vmaxy = (rcy + F[1] / rcz) - S[1]
    ma = np.array([rcx.min(), rcy.min(), rcz.min(), vmaxx.max(), vmaxy.max(), rcz.max()], np.float32)
    assert np.array_equal(ma.view(np.uint32), o.model_aabb[[0, 1, 2, 4, 5, 6]].view(np.uint32))
    # receiver rectangles (invalid receivers hold (+MAX, -MAX) and drop out)
    r = o.rect.reshape(-1, 8)
    sa = np.array([r[:, 0].min(), r[:, 1].min(), r[:, 2].min(), r[:, 4].max(), r[:, 5].max(), r[:, 6].max()],
                  np.float32)
    assert np.array_equal(sa.view(np.uint32), o.sample_aabb[[0, 1, 2, 4, 5, 6]].view(np.uint32))
    # far plane: max projected depth of receivers and vertices + 0.1
    P = fr.rx.pos[..., :3].reshape(-1, 3)
    ok = ~np.isnan(P[:, 0])
    rz = np.array([oracle.make_rectangle(mat, p)[2] for p in P[ok][::7]], np.float32)
    zmax = max(np.float32(rz.max()), rcz.max(), np.float32(r[:, 6].max()))
    assert f == np.float32(zmax + np.float32(0.1))
    assert np.array_equal(S, (F / f).astype(np.float32))
    # intersection
    B = np.array([max(sa[0], ma[0]), max(sa[1], ma[1]), ma[2]], np.float32)
    E = np.array([min(sa[3], ma[3]), min(sa[4], ma[4]), sa[5]], np.float32)
    assert np.array_equal(np.concatenate([B, E]).view(np.uint32), lb[45:51].astype(np.float32).view(np.uint32))
    # both sources occur
    assert ma[0] > sa[0] and sa[3] < ma[3] and sa[1] > ma[1] and sa[4] < ma[4]
    # bins: factor = G / (E - B)
    G = np.float32(1 << fr.grid_log2)
    assert np.array_equal((G / (E[:2] - B[:2])).astype(np.float32).view(np.uint32),
                          lb[51:53].astype(np.float32).view(np.uint32))


# --------------------------------------------------------------------------
# FD::moveModel
# --------------------------------------------------------------------------
def _colmajor(M):
    return np.asarray(M, np.float32).T.ravel()


def test_move_model_exact_cases(oracle):
    rng = np.random.default_rng(8)
    V = np.concatenate([rng.uniform(-5, 5, (500, 3)), rng.uniform(-1, 1, (500, 1))], 1).astype(np.float32)
    # identity: unchanged, including the fourth float
    out = oracle.move_model(V, _colmajor(np.eye(4)))
    assert np.array_equal(out.view(np.uint32), V.view(np.uint32))
    # translation by exactly representable steps: column-major m[12..14]
    M = np.eye(4)
    M[:3, 3] = (0.5, -2.0, 0.25)
    out = oracle.move_model(V, _colmajor(M))
    exp = V.copy()
    exp[:, :3] += np.array([0.5, -2.0, 0.25], np.float32)
    assert np.array_equal(out, exp)
    # the matrix is column-major: m[3], m[7], m[11] (the bottom row) do not
    # enter x, y, z; the translation is not read from the bottom row
    Mt = np.zeros(16, np.float32)
    Mt[[0, 5, 10, 15]] = 1.0
    Mt[[3, 7, 11]] = (9.0, 9.0, 9.0)
    out = oracle.move_model(V, Mt)
    assert np.array_equal(out, V)
    # axis permutation and power-of-two scale: exact
    P = np.array([[0, 0, 2, 0], [0.5, 0, 0, 0], [0, -4, 0, 0], [0, 0, 0, 1]])
    out = oracle.move_model(V, _colmajor(P))
    exp = V.copy()
    exp[:, 0], exp[:, 1], exp[:, 2] = 2 * V[:, 2], 0.5 * V[:, 0], -4 * V[:, 1]
    assert np.array_equal(out, exp)
    # only x, y, z change: with an 8-float stride the other five stay
    V8 = rng.normal(size=(200, 8)).astype(np.float32)
    out = oracle.move_model(V8, _colmajor(M))
    assert np.array_equal(out[:, 3:], V8[:, 3:])


def test_move_model_float64(oracle):
    rng = np.random.default_rng(9)
    V = np.concatenate([rng.uniform(-8, 8, (2000, 3)), np.ones((2000, 1))], 1).astype(np.float32)
    for _ in range(20):
        M = rng.normal(size=(4, 4))
        out = oracle.move_model(V, _colmajor(M)).astype(np.float64)
        Mf = M.astype(np.float32).astype(np.float64)
        ref = (np.concatenate([V[:, :3].astype(np.float64), np.ones((len(V), 1))], 1) @ Mf.T)[:, :3]
        scale = np.abs(np.concatenate([V[:, :3].astype(np.float64), np.ones((len(V), 1))], 1)) @ np.abs(Mf.T)[:, :3]
        assert np.all(np.abs(out[:, :3] - ref) <= 4 * 2.0 ** -24 * scale[:, :3] + 1e-30)


# --------------------------------------------------------------------------
# exits 4 (empty AABB) and 7 (no pairs) from constructed geometry
# --------------------------------------------------------------------------
def test_exit_empty_aabb(wl, oracle):
    """A negative right radius makes RectangleFactor.x negative, so every
    rectangle's max x lies below its min x; with the receivers in a tiny patch
    the receiver AABB is empty (Step2_3_defineFDbuffer.inl:45-47)."""
    fr = wl.plane_256(ntri=500, res=64)
    fr.rx = wl.receivers_plane(64, 64, 0.001)
    fr.light.right_radius = -0.5
    o = oracle.run(fr)
    assert o.exit == 4
    assert o.sample_aabb[0] > o.sample_aabb[4]
    assert not o.shadow.any() and not o.masks.any()


def test_exit_no_pairs(wl, oracle):
    """Occluders below the receiver plane face the light and reach the light
    plane, but lie farther from the light than every receiver: no bin passes
    the depth test, zero pairs (FD/api.cpp:522)."""
    fr = wl.plane_256(ntri=500, res=64)
    fr.scene.verts[:, 2] -= 3.6
    o = oracle.run(fr)
    assert o.exit == 7
    assert o.scalars["valid_tris"] > 0 and o.scalars["pairs"] == 0
    assert not o.shadow.any() and not o.masks.any()
