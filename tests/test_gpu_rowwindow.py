"""The occlusion kernel's row window against the rows it may skip (MI355X only).

The per-sample kernel intersects the reference's row-range mask
(Step5_2_shadowCalPerSample.inl:176-200) with a window of rows derived from
pairs of edge planes; a row outside the window is skipped, so a row that
holds a set light sample must never fall outside it.  fd_selftest_rowwindow
evaluates, for host-given edge-plane triples, every row's three-edge mask by
the general and the fast form, the window and the row-range mask.  The
triples are random and adversarial: planes of real triangles, nearly
parallel edges, coefficients near the window's 2^40 limit and the fast
form's 2^100 limit, zero and denormal pair determinants, infinities, NaN.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _triangle_edges(rng, n, R, spread=1.0):
    """Edge planes of random triangles in (row, col) space, interior >= 0."""
    P = rng.uniform(-0.5 * R * spread, 1.5 * R * spread, (n, 3, 2))
    coef = np.zeros((n, 3, 3))
    for k in range(3):
        a, b, c = P[:, k], P[:, (k + 1) % 3], P[:, (k + 2) % 3]
        A = -(b[:, 1] - a[:, 1])
        B = b[:, 0] - a[:, 0]
        C = -(A * a[:, 0] + B * a[:, 1])
        s = np.sign(A * c[:, 0] + B * c[:, 1] + C)
        s[s == 0] = 1
        coef[:, k] = np.stack([A * s, B * s, C * s], 1)
    return coef


def _cases(rng, R):
    n = 60_000
    out = []
    # triangles overlapping the grid, each edge scaled by its own power of two
    t = _triangle_edges(rng, n, R)
    t *= 2.0 ** rng.integers(-30, 31, (n, 3, 1))
    out.append(t)
    # thin slivers: two nearly parallel edges (pair determinant near zero)
    s = _triangle_edges(rng, n // 2, R)
    lam = rng.uniform(0.5, 2.0, (n // 2, 1))
    d = rng.choice([0.0, 1e-7, 1e-6, 1e-4, 1e-2], (n // 2, 1)) * rng.normal(size=(n // 2, 1))
    s[:, 1, 0] = -s[:, 0, 0] * lam[:, 0] * (1 + d[:, 0])
    s[:, 1, 1] = -s[:, 0, 1] * lam[:, 0]
    s[:, 1, 2] = -s[:, 0, 2] * lam[:, 0] + rng.normal(size=n // 2) * R * 0.05 * np.abs(s[:, 0, 1])
    out.append(s)
    # coefficients near the window's 2^40 limit and near 2^100
    h = _triangle_edges(rng, n // 4, R)
    h /= np.abs(h).max(axis=(1, 2), keepdims=True)
    h *= 2.0 ** rng.choice([38.0, 39.5, 40.0, 40.5, 41.0, 99.0, 100.5], (n // 4, 1, 1))
    out.append(h)
    # tiny coefficients: zero / denormal determinants
    z = _triangle_edges(rng, n // 4, R)
    z *= 2.0 ** rng.choice([-60.0, -70.0, -126.0, -135.0], (n // 4, 1, 1))
    z[::5, :, 0] = 0.0            # A = 0 for every edge: al = 0
    z[1::5, 0, 1] = 0.0           # B = 0 on one edge
    out.append(z)
    # special values sprinkled into otherwise valid triples
    sp = _triangle_edges(rng, 4000, R)
    flat = sp.reshape(-1)
    pos = rng.integers(0, flat.size, 1500)
    flat[pos] = rng.choice([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45], 1500)
    out.append(sp)
    # random triples (mostly empty masks)
    r = rng.normal(size=(n // 4, 3, 3)) * rng.choice([1e-3, 1.0, 1e3], (n // 4, 1, 1))
    r[:, :, 2] *= R
    out.append(r)
    return np.concatenate(out).astype(np.float32).reshape(-1, 9)


def _lowmask(R):
    return np.uint64((1 << R) - 1) if R < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)


@pytest.mark.parametrize("R", [8, 16, 32, 33, 64])
def test_row_window_keeps_every_occluded_row(fd, R):
    rng = np.random.default_rng(100 + R)
    coef = _cases(rng, R)
    print(f"[rowwindow] R={R} triples={len(coef)}", flush=True)
    out = fd.selftest_rowwindow(coef, R)
    gen, fast, win, rng_rows = out[:, 0], out[:, 1], out[:, 2], out[:, 3]
    # the fast row form equals the general one wherever the kernel uses it
    bad_fast = np.nonzero(gen != fast)[0]
    assert len(bad_fast) == 0, coef[bad_fast[:3]].tolist()
    # every visited row (row-range mask) that holds a set sample lies in the window
    lost = gen & rng_rows & ~win
    bad = np.nonzero(lost)[0]
    assert len(bad) == 0, [(coef[i].tolist(), hex(int(gen[i])), hex(int(win[i]))) for i in bad[:3]]
    # stronger: no row with a set sample is outside the window at all
    assert not np.any(gen & ~win)
    # the cases exercise the window: it removes rows from many row ranges
    # that still hold occluded rows
    full = _lowmask(R)
    cuts = (win != full) & ((gen & rng_rows) != 0)
    assert cuts.sum() > 10_000, int(cuts.sum())
    removed = rng_rows & ~win
    assert (removed != 0).sum() > 10_000
