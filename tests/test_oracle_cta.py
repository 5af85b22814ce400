"""Pins of the oracle's bin-queue (CTA) occlusion variant (CPU).

shadowCal_CTA_kernel (Step5_2_shadowCal.inl:408-1100) is dead code in the
reference; its arithmetic differs from the per-sample kernel's, so it has
its own pins, none of which restates the oracle's code:

  range primitives   calLightRowRange / calLightHeightRange against the
                     half-lines they select, on inputs whose quotient is exact
  occlusion          an exact float64 segment/triangle ray cast per light
                     sample (rows along y_delt, bits along x_delt), for
                     occluders between receiver and light and for occluders
                     that cross the receivers' depth (the triangle plane)
  against per-sample the per-sample kernel's masks, transposed, agree away
                     from edge samples; fractions are popcount / samples
"""
import numpy as np
import pytest

from test_oracle import _raycast


def _bits(m, n):
    return [(m >> j) & 1 for j in range(n)]


@pytest.mark.parametrize("W", [8, 16, 32, 40, 64])
def test_row_range_is_the_half_line(oracle, W):
    """Bit j is set iff B j + D > 0 (B < 0) or B j + D >= 0 (B > 0), for
    |B| a power of two so that -D / B is exact."""
    rng = np.random.default_rng(W)
    full = (1 << W) - 1
    for _ in range(400):
        B = float(rng.choice([-1, 1])) * 2.0 ** int(rng.integers(-3, 4))
        D = float(rng.integers(-4 * W, 4 * W)) / 4.0 * abs(B)
        m = oracle.cta_row_range(B, D, W) & full
        if B < 0 and D <= 0:
            want = 0
        elif D == 0:
            want = full
        else:
            j = np.arange(W, dtype=np.float64)
            sel = (B * j + D > 0) if B < 0 else (B * j + D >= 0)
            want = int(sum(1 << int(k) for k in np.nonzero(sel)[0]))
        assert m == want, (B, D, W, bin(m), bin(want))


def test_row_range_special_cases(oracle):
    full = (1 << 16) - 1
    assert oracle.cta_row_range(0.0, -1.0, 16) == 0
    assert oracle.cta_row_range(0.0, 1.0, 16) & full == full
    assert oracle.cta_row_range(-1.0, 0.0, 16) == 0
    assert oracle.cta_row_range(1.0, 0.0, 16) & full == full
    assert oracle.cta_row_range(-1.0, 2.5, 16) == 0b111
    assert oracle.cta_row_range(1.0, -2.5, 16) & full == full & ~0b111
    assert oracle.cta_row_range(-1.0, 3.0, 16) == 0b111        # strict: j = 3 gives 0
    assert oracle.cta_row_range(1.0, -3.0, 16) & full == full & ~0b111   # j = 3 gives 0: set
    assert oracle.cta_row_range(-1.0, 100.0, 16) == full         # clamped to W
    assert oracle.cta_row_range(-1.0, 100.0, 32) == (1 << 32) - 1   # (T)1 << 32 is 0: all ones
    assert oracle.cta_row_range(float("nan"), 1.0, 16) & full == 0   # NaN B: ~(full), outside the grid


@pytest.mark.parametrize("W", [8, 16, 32])
def test_height_range_is_the_row_half_line(oracle, W):
    """Row i is kept iff A i + D > 0 (A < 0) or >= 0 (A > 0), with
    D = max(B (W-1) + C, C) the edge's largest value over the columns."""
    rng = np.random.default_rng(7 * W)
    full = (1 << W) - 1
    for _ in range(400):
        A = float(rng.choice([-1, 1])) * 2.0 ** int(rng.integers(-2, 3))
        B = float(rng.integers(-8, 9)) / 8.0
        C = float(rng.integers(-8 * W, 8 * W)) / 8.0
        D = max(B * (W - 1) + C, C)
        m = oracle.cta_height_range(A, B, C, W)
        i = np.arange(W, dtype=np.float64)
        if A <= 0 and D <= 0:
            want = 0
        elif D == 0:
            want = full
        else:
            sel = (A * i + D > 0) if A < 0 else (A * i + D >= 0)
            want = int(sum(1 << int(k) for k in np.nonzero(sel)[0]))
        assert m == want, (A, B, C, W, bin(m), bin(want))


def _samples_cta(lb, R):
    """Light sample of bit j in row i: upRight + j x_delt + i y_delt; index i*R + j."""
    ur, xd, yd = lb[28:31].astype(np.float64), lb[22:25].astype(np.float64), lb[25:28].astype(np.float64)
    i, j = np.meshgrid(np.arange(R), np.arange(R), indexing="ij")
    return ur + j.reshape(-1, 1) * xd + i.reshape(-1, 1) * yd


def _bits_of(masks_row, R):
    wpr = 1 if R <= 32 else 2
    m = masks_row.reshape(R, wpr).astype(np.uint64)
    words = m[:, 0] | (m[:, 1] << np.uint64(32)) if wpr == 2 else m[:, 0]
    return ((words[:, None] >> np.arange(R, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool).ravel()


@pytest.mark.parametrize("R,zbox", [(8, (1.0, 3.0)), (16, (1.0, 3.0)), (32, (1.0, 3.0)), (16, (-0.6, 1.2))])
def test_cta_matches_raycast(wl, oracle, R, zbox):
    """Occluders between receiver and light, and occluders that cross the
    receiver plane z = 0 (the triangle plane keeps the part behind the
    receiver from casting a shadow)."""
    rng = np.random.default_rng(300 + R)
    fr = wl.plane_256(ntri=60, res=40, light_res=R)
    fr.scene = wl.tris_uniform(60, 19 + R, box=(-1.5, 1.5, -1.5, 1.5) + zbox, e0=0.3, e1=0.9)
    fr.backface_cull = 0
    o = oracle.run(fr, variant=1)
    assert o.exit == 0
    Ls = _samples_cta(o.light_bits, R)
    tris = fr.scene.verts[:, :3].astype(np.float64)[fr.scene.idx]
    pos = fr.rx.pos.reshape(-1, 4)[:, :3].astype(np.float64)
    masks = o.masks.reshape(fr.rx.w * fr.rx.h, -1)
    ids = rng.choice(len(pos), 300, replace=False)
    nbits = nbad = nset = 0
    for i in ids:
        hit, margin = _raycast(pos[i], Ls, tris)
        got = _bits_of(masks[i], R)
        nbits += got.size
        nset += int(got.sum())
        nbad += int(np.count_nonzero((got != hit) & (margin > 1e-4)))
        assert o.shadow[i] == np.float32(np.float32(got.sum()) / np.float32(R * R))
    assert nset > nbits // 50, "too few occluded samples to pin anything"
    assert nbad == 0, f"{nbad} of {nbits} light samples disagree with the ray cast away from triangle edges"


def test_cta_close_to_per_sample(wl, oracle):
    """Same frame through both kernels: the CTA masks are the per-sample
    masks transposed, up to samples on a shadow edge."""
    fr = wl.plane_256(ntri=2000, res=64, light_res=16)
    a = oracle.run(fr, variant=0)
    b = oracle.run(fr, variant=1)
    assert a.exit == 0 and b.exit == 0
    R = 16
    n = fr.rx.w * fr.rx.h
    ma = a.masks.reshape(n, R).astype(np.uint64)
    mb = b.masks.reshape(n, R).astype(np.uint64)
    ba = ((ma[:, :, None] >> np.arange(R, dtype=np.uint64)) & np.uint64(1)).astype(bool)   # [n, row, bit]
    bb = ((mb[:, :, None] >> np.arange(R, dtype=np.uint64)) & np.uint64(1)).astype(bool)
    diff = ba != np.transpose(bb, (0, 2, 1))
    assert ba.sum() > 0.02 * ba.size
    assert diff.mean() < 2e-3, diff.mean()
    assert np.abs(a.shadow - b.shadow).max() < 0.05
