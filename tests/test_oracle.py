"""Pins of the CPU oracle (runs on the CPU).

The reference ships no tests or golden data, so each oracle stage is pinned
here by something the oracle's own code does not restate:

  light set-up       closed forms in float64 (corner = upRight + r*x_delt +
                     c*y_delt, the view/projection matrix from its definition)
  makeRectangle      float64 evaluation of the projection
  row masks          the half-line {c : D + B*c >= 0} it approximates
  occlusion          an exact float64 segment/triangle ray cast per light
                     sample, and an unbinned brute force over all triangles
  fraction           popcount / light samples
  binning / sort     sort invariants, stability, per-bin reductions with numpy,
                     light-plane cell containment of each receiver's centre
  pairing            brute-force (triangle x bin) enumeration
  early exits        constructed inputs for each FDLaunch early return
  regression         digests of the plane-256 outputs (tests/golden)
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "plane256.json")


def _light_frame(wl, light, res=8):
    fr = wl.plane_256(ntri=50, res=res)
    fr.light = light
    return fr


# --------------------------------------------------------------------------
# light set-up (FD/api.cpp:303-400)
# --------------------------------------------------------------------------
def _light_closed_form(L):
    pos = np.array(L.position, np.float64)
    view = np.array(L.view_dir, np.float64)
    up = np.array(L.up_dir, np.float64)
    view /= np.linalg.norm(view)
    up /= np.linalg.norm(up)
    right = np.cross(view, up)
    right /= np.linalg.norm(right)
    top = np.cross(right, view)
    top /= np.linalg.norm(top)
    W, H = L.res_w, L.res_h
    xd = right * (-2.0 * L.right_radius / W)
    yd = top * (-2.0 * L.top_radius / H)
    ur = pos + right * L.right_radius * (W - 1) / W + top * L.top_radius * (H - 1) / H
    # view matrix at ur looking along view (glm lookAtRH), infinite perspective
    f = view
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    MV = np.eye(4)
    MV[0, :3] = s
    MV[1, :3] = u
    MV[2, :3] = -f
    MV[:3, 3] = [-s @ ur, -u @ ur, f @ ur]
    rng = np.tan(np.float64(np.float32(60.0))) * 0.1
    P = np.zeros((4, 4))
    P[0, 0] = 0.2 / (2 * rng)
    P[1, 1] = 0.2 / (2 * rng)
    P[2, 2] = -1
    P[3, 2] = -1
    P[2, 3] = -0.2
    MVP = P @ MV           # row-major (row i, col j)
    mat = MVP.T.ravel()    # column-major storage
    F = np.array([L.right_radius * (W - 1) / W * P[0, 0], L.top_radius * (H - 1) / H * P[1, 1]])
    return dict(xd=xd, yd=yd, ur=ur, mat=mat, F=F)


@pytest.mark.parametrize("pose", [((0, 0, 4), (0, 0, -1), (0, 1, 0), 0.5, 0.5, 16),
                                  ((1, -2, 6), (0.1, 0.2, -1), (0, 1, 0), 1.5, 0.5, 64),
                                  ((-3, 1, 5), (0.3, -0.1, -1), (0.2, 1, 0.1), 1.0, 1.0, 8),
                                  ((0, 0, 7), (0, 0, -1), (1, 0, 0), 0.75, 0.25, 1)])
def test_light_setup_closed_form(wl, oracle, pose):
    pos, view, up, rr, tr, res = pose
    L = wl.LightDesc(rr, tr, pos, view, up, res, res)
    o = oracle.run(_light_frame(wl, L))
    lb = o.light_bits.astype(np.float64)
    ref = _light_closed_form(L)
    np.testing.assert_allclose(lb[0:16], ref["mat"], rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(lb[22:25], ref["xd"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(lb[25:28], ref["yd"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(lb[28:31], ref["ur"], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(lb[40:42], ref["F"], rtol=1e-5)
    # the four corners are light samples: ul = ur + (W-1) xd, dr = ur + (H-1) yd
    ur, ul, dl, dr = lb[28:31], lb[31:34], lb[34:37], lb[37:40]
    xd, yd = lb[22:25], lb[25:28]
    np.testing.assert_allclose(ul, ur + (res - 1) * xd, atol=1e-5)
    np.testing.assert_allclose(dr, ur + (res - 1) * yd, atol=1e-5)
    np.testing.assert_allclose(dl, ur + (res - 1) * (xd + yd), atol=1e-5)


def test_make_rectangle_float64(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        m = rng.normal(size=16).astype(np.float32)
        p = rng.normal(size=3).astype(np.float32)
        got = oracle.make_rectangle(m, p).astype(np.float64)
        md, pd = m.astype(np.float64), p.astype(np.float64)
        z = md[3] * pd[0] + md[7] * pd[1] + md[11] * pd[2] + md[15]
        if abs(z) < 1e-2:
            continue
        x = (md[0] * pd[0] + md[4] * pd[1] + md[8] * pd[2] + md[12]) / z * 0.5 + 0.5
        y = (md[1] * pd[0] + md[5] * pd[1] + md[9] * pd[2] + md[13]) / z * 0.5 + 0.5
        scale = np.abs(md).sum() * (np.abs(pd).sum() + 1) / abs(z) + 1
        np.testing.assert_allclose(got, [x, y, z], atol=1e-5 * scale)


# --------------------------------------------------------------------------
# light-row masks (Step5_2_shadowCalPerSample.inl:14-33)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("W", [1, 5, 16, 32, 33, 64])
def test_rowmask_is_the_half_line(oracle, W):
    rng = np.random.default_rng(W)
    n = W - 1
    bad = 0
    for _ in range(3000):
        D = np.float32(rng.normal() * 10)
        B = np.float32(rng.normal() * rng.choice([0.01, 1, 10]))
        m, clip = oracle.rowmask(D, B, n, W)
        assert m >> W == 0 if W < 64 else True
        for c in range(W):
            v = float(D) + float(B) * c
            if abs(v) < 1e-5 * (abs(float(D)) + abs(float(B)) * c + 1e-30):
                continue
            if ((m >> c) & 1) != (v >= 0):
                bad += 1
    assert bad == 0


def test_rowmask_special_cases(oracle):
    full = (1 << 16) - 1
    assert oracle.rowmask(1.0, 1.0, 15, 16)[0] == full      # both ends non-negative
    assert oracle.rowmask(-1.0, -1.0, 15, 16)[0] == 0       # both negative
    assert oracle.rowmask(0.0, -1.0, 15, 16)[0] == 0        # ceil(0) = 0 samples
    assert oracle.rowmask(2.5, -1.0, 15, 16)[0] == 0b111    # c = 0, 1, 2
    assert oracle.rowmask(-2.5, 1.0, 15, 16)[0] == full & ~0b111
    m, clip = oracle.rowmask(-2.5, 1.0, 15, 16)
    assert clip   # the reference would set bits 16..31 here


# --------------------------------------------------------------------------
# occlusion: exact ray cast (the refCal idea, Step5_2_calRef.inl:31-131)
# --------------------------------------------------------------------------
def _raycast(Q, Ls, tris):
    """Q: (3,), Ls: (S,3) light samples, tris: (T,3,3).  Returns (hit (S,),
    margin (S,)) where margin is the smallest barycentric / segment distance
    of a hit (or near miss) to the triangle boundary."""
    d = Ls - Q                                  # (S,3)
    v0, v1, v2 = tris[:, 0], tris[:, 1], tris[:, 2]
    e1, e2 = v1 - v0, v2 - v0                   # (T,3)
    p = np.cross(d[:, None, :], e2[None])       # (S,T,3)
    det = np.einsum("stk,tk->st", p, e1)
    ok = np.abs(det) > 1e-14
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    s = Q[None, :] - v0                          # (T,3)
    u = np.einsum("stk,tk->st", p, s) * inv
    q = np.cross(s, e1)                          # (T,3)
    v = np.einsum("sk,tk->st", d, q) * inv
    t = np.einsum("tk,tk->t", e2, q)[None, :] * inv
    m = np.minimum.reduce([u, v, 1 - u - v, t, 1 - t])
    m = np.where(ok, m, -1.0)
    hit = (m > 0).any(axis=1)
    margin = np.min(np.abs(m), axis=1)
    return hit, margin


def _samples(lb, R):
    ur, xd, yd = lb[28:31].astype(np.float64), lb[22:25].astype(np.float64), lb[25:28].astype(np.float64)
    r, c = np.meshgrid(np.arange(R), np.arange(R), indexing="ij")
    return ur + r.reshape(-1, 1) * xd + c.reshape(-1, 1) * yd   # row-major: index r*R + c


def _mask_bits(masks_row, R):
    wpr = 1 if R <= 32 else 2
    m = masks_row.reshape(R, wpr).astype(np.uint64)
    words = m[:, 0] | (m[:, 1] << np.uint64(32)) if wpr == 2 else m[:, 0]
    return ((words[:, None] >> np.arange(R, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool).ravel()


@pytest.mark.parametrize("R", [8, 16, 33])
def test_occlusion_matches_raycast(wl, oracle, R):
    rng = np.random.default_rng(100 + R)
    fr = wl.plane_256(ntri=40, res=40, light_res=R)
    # triangles strictly between the receivers (z=0) and the light (z=4)
    fr.scene = wl.tris_uniform(40, 9 + R, box=(-1.5, 1.5, -1.5, 1.5, 1.0, 3.0), e0=0.3, e1=0.9)
    fr.backface_cull = 0
    o = oracle.run(fr)
    assert o.exit == 0
    lb = o.light_bits
    Ls = _samples(lb, R)
    verts = fr.scene.verts[:, :3].astype(np.float64)
    tris = verts[fr.scene.idx]
    pos = fr.rx.pos.reshape(-1, 4)[:, :3].astype(np.float64)
    masks = o.masks.reshape(fr.rx.w * fr.rx.h, -1)
    ids = rng.choice(len(pos), 300, replace=False)
    nbits = nbad = 0
    for i in ids:
        hit, margin = _raycast(pos[i], Ls, tris)
        got = _mask_bits(masks[i], R)
        diff = got != hit
        nbits += got.size
        nbad += int(np.count_nonzero(diff & (margin > 1e-4)))
        # the fraction is the popcount over the grid
        assert o.shadow[i] == np.float32(np.float32(got.sum()) / np.float32(R * R))
    assert nbad == 0, f"{nbad} of {nbits} light samples disagree with the ray cast away from triangle edges"


def test_binned_equals_unbinned(wl, oracle):
    """Masks do not depend on the light-plane bins: the binned result equals
    an occlusion over every light-plane triangle."""
    fr = wl.plane_256(ntri=4000, res=64)
    o = oracle.run(fr)
    ids = np.arange(0, 64 * 64, 3)
    brute = oracle.occlude_all(o, fr, ids)
    got = o.masks.reshape(64 * 64, -1)[ids]
    assert np.array_equal(brute, got)


def _edge_tie_receivers(o):
    """Receivers whose bin statistics touch the light-plane boundary: there the
    LP-clamped triangle AABB of the pair count (Step4_RasterizeTriangleToBin.inl:
    74-77) equals the bin's min range and the strict tests of :94 drop pairs the
    per-sample cull would keep."""
    nb = int(o.scalars["num_bins"])
    mr = o.bin_minrange.reshape(-1, 4)
    B = np.float32([o.scalars["lp_bx"], o.scalars["lp_by"]])
    E = np.float32([o.scalars["lp_ex"], o.scalars["lp_ey"]])
    tie = (mr[:, 0] == B[0]) | (mr[:, 1] == B[1]) | (mr[:, 2] == E[0]) | (mr[:, 3] == E[1])
    k = o.key
    return (k < nb) & tie[np.minimum(k, nb - 1)]


@pytest.mark.parametrize("grid_log2", [3, 5, 6])
def test_masks_independent_of_grid_except_edge_ties(wl, oracle, grid_log2):
    fr = wl.plane_256(ntri=3000, res=64)
    a = oracle.run(fr)
    fr.grid_log2 = grid_log2
    b = oracle.run(fr)
    diff = (a.masks.reshape(64 * 64, -1) != b.masks.reshape(64 * 64, -1)).any(1)
    allowed = _edge_tie_receivers(a) | _edge_tie_receivers(b)
    assert not np.any(diff & ~allowed), f"{int(np.sum(diff & ~allowed))} receivers differ away from edge ties"


# --------------------------------------------------------------------------
# binning / sort / pairing invariants
# --------------------------------------------------------------------------
def test_receiver_binning_invariants(wl, oracle):
    fr = wl.heightfield_1k(res=160, nsph=40)
    o = oracle.run(fr, subset=[])
    nb = int(o.scalars["num_bins"])
    G = int(o.scalars["width_bins"])
    key, sk, sid = o.key, o.sorted_key, o.sorted_id
    assert np.all(np.diff(sk) >= 0)
    assert np.array_equal(np.sort(sid), np.arange(len(sid)))          # a permutation
    assert np.array_equal(key[sid], sk)
    for k in np.unique(sk):                                           # stable
        ids = sid[sk == k]
        assert np.all(np.diff(ids) > 0)
    counts = np.bincount(key, minlength=nb + 1)
    st, en = o.bin_start, o.bin_end
    assert np.array_equal(np.where(counts > 0, en - st, 0), counts)
    assert int((key < nb).sum()) == int(o.scalars["valid_samples"])
    # per-bin reductions
    rect = o.rect.reshape(-1, 8)
    for b in range(nb):
        m = key == b
        if not m.any():
            assert o.bin_maxz[b] == np.float32(-3.402823466e38)
            continue
        assert o.bin_maxz[b] == rect[m, 2].max()
        np.testing.assert_array_equal(o.bin_minrange[4 * b:4 * b + 4],
                                      [rect[m, 0].max(), rect[m, 1].max(), rect[m, 4].min(), rect[m, 5].min()])
    # the key is the light-plane cell holding the rectangle's centre
    B = np.array([o.scalars["lp_bx"], o.scalars["lp_by"]])
    E = np.array([o.scalars["lp_ex"], o.scalars["lp_ey"]])
    valid = key < nb
    c = (rect[valid][:, [0, 1]].astype(np.float64) + rect[valid][:, [4, 5]]) / 2
    cell = np.floor((c - B) / (E - B) * G)
    ok = np.all(np.abs(cell - np.stack([key[valid] & (G - 1), key[valid] >> int(np.log2(G))], 1)) <= 1, axis=1)
    assert ok.all()


def test_pairing_bruteforce(wl, oracle):
    fr = wl.heightfield_1k(res=128, nsph=30)
    o = oracle.run(fr, subset=[])
    G = int(o.scalars["width_bins"])
    nb = G * G
    L2 = int(o.scalars["width_bins_log2"])
    B = np.float32([o.scalars["lp_bx"], o.scalars["lp_by"]])
    E = np.float32([o.scalars["lp_ex"], o.scalars["lp_ey"]])
    a = o.tri_aabb.reshape(-1, 8)
    mn, mx = a[:, 0:4], a[:, 4:8]
    keep = ~((mn[:, 0] > mx[:, 0]) | (mn[:, 1] > mx[:, 1]) | (mn[:, 2] > mx[:, 2]) | (mn[:, 0] > E[0]) |
             (mn[:, 1] > E[1]) | (mx[:, 0] < B[0]) | (mx[:, 1] < B[1]))
    bx = np.maximum(mn[:, 0], B[0])
    by = np.maximum(mn[:, 1], B[1])
    ex = np.minimum(mx[:, 0], E[0])
    ey = np.minimum(mx[:, 1], E[1])
    mz, mr = o.bin_maxz, o.bin_minrange.reshape(-1, 4)
    pairs = set()
    for b in range(nb):
        t = keep & (mz[b] >= mn[:, 2]) & (bx < mr[b, 0]) & (by < mr[b, 1]) & (mr[b, 2] < ex) & (mr[b, 3] < ey)
        for ti in np.nonzero(t)[0]:
            pairs.add((b, int(ti)))
    got = list(zip(o.pair_bin.tolist(), o.pair_tri.tolist()))
    assert got == sorted(pairs)                      # same set, sorted by (bin, triangle)
    assert int(o.scalars["pairs"]) == len(pairs) == int(o.tri_count.sum())
    cnt = np.bincount(o.pair_bin, minlength=nb)
    st, en = o.bin_tri_start, o.bin_tri_end
    assert np.array_equal(np.where(cnt > 0, en - st, 0), cnt)
    del L2


# --------------------------------------------------------------------------
# early exits (FD/api.cpp:495-559) and their outputs
# --------------------------------------------------------------------------
def test_exit_outputs(wl, oracle):
    fr = wl.plane_256(ntri=300, res=32)
    fr.rx.pos[...] = np.nan
    o = oracle.run(fr)
    assert o.exit == 1 and not o.shadow.any() and not o.masks.any()
    fr = wl.plane_256(ntri=300, res=32)
    fr.scene.verts[:, 2] += 10.0
    assert oracle.run(fr).exit == 2
    fr = wl.plane_256(ntri=300, res=32)
    fr.rx.nrm[:16, :, 2] = -1.0          # back-facing rows get shadow 1 (Step2_1:97)
    fr.scene.verts[:, 0] += 60.0         # and the light plane is empty
    o = oracle.run(fr)
    assert o.exit == 5
    sh = o.shadow.reshape(32, 32)
    assert np.all(sh[:16] == 1.0) and np.all(sh[16:] == 0.0)
    full = (1 << 16) - 1
    mk = o.masks.reshape(32, 32, 16)
    assert np.all(mk[:16] == full) and np.all(mk[16:] == 0)


def test_unsupported_light(wl, oracle):
    fr = wl.plane_256(ntri=100, res=16)
    fr.light.res_h = 8
    assert oracle.run(fr).exit == 8


# --------------------------------------------------------------------------
# regression digests
# --------------------------------------------------------------------------
def digests(o):
    def h(a):
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    return {"exit": o.exit, "pairs": int(o.scalars["pairs"]), "valid_samples": int(o.scalars["valid_samples"]),
            "valid_tris": int(o.scalars["valid_tris"]), "shadow": h(o.shadow), "masks": h(o.masks),
            "key": h(o.key), "sorted_id": h(o.sorted_id), "pair_tri": h(o.pair_tri), "pair_bin": h(o.pair_bin),
            "tri_aabb": h(o.tri_aabb), "rect": h(o.rect), "light": h(o.light_bits)}


def test_golden_plane256(wl, oracle):
    with open(GOLDEN) as f:
        gold = json.load(f)
    o = oracle.run(wl.plane_256())
    assert digests(o) == gold
