"""HIP path vs the CPU oracle, stage by stage, bit-exact (MI355X only).

Inputs are the deterministic workloads; every array the reference produces
(light block, reductions, receiver rectangles, bin keys, the stable sort
permutation, per-bin statistics, triangle records, pair counts, the sorted
pair list, per-bin ranges, shadow values and light-sample masks) is compared
by value / bit pattern, at the full heightfield-1k (1024^2 receivers, 32^2
light, 200k triangles) and boxes-1080p (1920x1080, 64^2 light, 1M
triangles) sizes as well: every receiver's mask (the oracle runs on all
host threads).
"""
import numpy as np
import pytest

from parity import compare, failures

pytestmark = pytest.mark.gpu


def _check(fd, oracle, frame, subset=None, expect_exit=None):
    print(f"[parity] {frame.name} {frame.rx.w}x{frame.rx.h} light {frame.light.res_w} "
          f"tris {frame.scene.ntris} grid {1 << frame.grid_log2}", flush=True)
    rep, fr, o = compare(fd, oracle, frame, subset=subset)
    bad = failures(rep)
    assert not bad, f"{frame.name}: mismatches {bad} (oracle exit {o.exit})"
    if expect_exit is not None:
        assert o.exit == expect_exit, f"expected exit {expect_exit}, oracle gave {o.exit}"
    fr.close()
    return o


def test_plane256_full(fd, oracle, wl):
    o = _check(fd, oracle, wl.plane_256())
    assert o.exit == 0 and o.scalars["pairs"] > 0


@pytest.mark.parametrize("grid_log2", [3, 5, 6])
def test_plane256_bin_grids(fd, oracle, wl, grid_log2):
    fr = wl.plane_256(ntri=3000)
    fr.grid_log2 = grid_log2
    _check(fd, oracle, fr)


@pytest.mark.parametrize("res", [1, 2, 5, 8, 31, 32, 33, 48, 64])
def test_light_grids(fd, oracle, wl, res):
    fr = wl.plane_256(ntri=2000, res=96, light_res=res)
    _check(fd, oracle, fr)


@pytest.mark.parametrize("res", [40, 64])
def test_wide_rows_one_item_per_chunk(fd, oracle, wl, res):
    """64-bit rows on a frame large enough for unsplit triangle lists: the
    waves of a work item share its 64 receivers' rows in LDS and split its
    triangle list."""
    fr = wl.plane_256(ntri=4000, res=384, light_res=res)
    _check(fd, oracle, fr)


def test_non_pow2_view_and_cull_off(fd, oracle, wl):
    fr = wl.plane_256(ntri=3000)
    fr.rx = wl.receivers_heightfield(200, 120, 4.0, 2.4)
    fr.backface_cull = 0
    _check(fd, oracle, fr)


def test_heightfield_small(fd, oracle, wl):
    fr = wl.heightfield_1k(res=192, nsph=60, light_res=32)
    _check(fd, oracle, fr)


def test_heightfield_1k_full(fd, oracle, wl):
    o = _check(fd, oracle, wl.heightfield_1k())
    assert o.exit == 0 and o.scalars["valid_samples"] == 1024 * 1024


def test_boxes_small_64(fd, oracle, wl):
    fr = wl.boxes_1080p(w=320, h=180, ntri=60_000)
    _check(fd, oracle, fr)


def test_boxes_1080p_full(fd, oracle, wl):
    """The largest single-GPU frame, 64-bit rows with 8 waves per work item:
    all 2,073,600 receivers x 4096 light samples."""
    o = _check(fd, oracle, wl.boxes_1080p())
    assert o.exit == 0 and o.scalars["pairs"] > 0


# --------------------------------------------------------------------------
# early exits of FDLaunch (FD/api.cpp:495-559) and degenerate inputs
# --------------------------------------------------------------------------
def test_exit_no_receivers(fd, oracle, wl):
    fr = wl.plane_256(ntri=500, res=64)
    fr.rx.pos[...] = np.nan
    _check(fd, oracle, fr, expect_exit=1)


def test_exit_tris_behind_light(fd, oracle, wl):
    fr = wl.plane_256(ntri=500, res=64)
    fr.scene.verts[:, 2] += 10.0
    _check(fd, oracle, fr, expect_exit=2)


def test_exit_all_backfacing(fd, oracle, wl):
    fr = wl.plane_256(ntri=500, res=64)
    fr.rx.nrm[..., 2] = -1.0
    _check(fd, oracle, fr, expect_exit=3)


def test_exit_light_plane_empty(fd, oracle, wl):
    fr = wl.plane_256(ntri=500, res=64)
    fr.scene.verts[:, 0] += 60.0   # occluders far to the side
    _check(fd, oracle, fr, expect_exit=5)


def test_exit_no_valid_triangles(fd, oracle, wl):
    fr = wl.plane_256(ntri=500, res=64)
    # flatten every triangle to a horizontal one facing down, away from the light
    v = fr.scene.verts.reshape(-1, 3, 4).copy()
    v[:, :, 2] = v[:, :, 2].mean(axis=1, keepdims=True)
    n = np.cross(v[:, 1, :3] - v[:, 0, :3], v[:, 2, :3] - v[:, 1, :3])
    flip = n[:, 2] > 0
    v[flip, 1], v[flip, 2] = v[flip, 2].copy(), v[flip, 1].copy()
    fr.scene.verts[:] = v.reshape(-1, 4)
    _check(fd, oracle, fr, expect_exit=6)


def test_mixed_nan_and_backfacing(fd, oracle, wl):
    fr = wl.plane_256(ntri=1500, res=128)
    rng = np.random.default_rng(3)
    m = rng.random((128, 128)) < 0.2
    fr.rx.pos[m, 0] = np.nan
    b = rng.random((128, 128)) < 0.2
    fr.rx.nrm[b, 2] = -1.0
    _check(fd, oracle, fr)


def test_repeat_launch_idempotent(fd, oracle, wl):
    """The reference divides its rectangle constant by f in place on every
    FDLaunch (Step2_3_defineFDbuffer.inl:24-25); here launches repeat."""
    fr = wl.plane_256(ntri=2000, res=128)
    rep, run, o = compare(fd, oracle, fr)
    assert not failures(rep)
    rep2, _, _ = compare(fd, oracle, fr, fr=run, o=o)
    assert not failures(rep2), failures(rep2)
    run.close()


@pytest.mark.parametrize("grid_log2", [5, 6])
def test_large_triangles_wave_walk(fd, oracle, wl, grid_log2):
    """Triangles whose bin rectangles exceed 64 bins are walked by the whole
    wave in the pair count and write passes; mixed with small ones so that
    both walks share waves.  Pair lists, counts and ranges stay bit-exact."""
    fr = wl.plane_256(ntri=1500, res=128)
    big = []
    for k, z in enumerate((0.8, 1.2, 1.6, 2.0, 2.4)):
        s = 3.2 - 0.4 * k
        big += [(-s, -s, z), (s, -s + 0.1 * k, z), (0.1 * k, s, z)]      # counter-clockwise from +z
    big += [(-1.5, -0.5, 1.0), (1.5, -0.4, 1.1), (0.2, 1.4, 1.05)]
    bv = np.array([(x, y, zz, 1.0) for x, y, zz in big], np.float32)
    n0 = fr.scene.nverts
    fr.scene.verts = np.concatenate([fr.scene.verts, bv])
    bi = (n0 + np.arange(len(big), dtype=np.int32)).reshape(-1, 3)
    # interleave the large triangles among the small ones (different waves and lanes)
    idx = fr.scene.idx
    fr.scene.idx = np.concatenate([idx[:40], bi[:2], idx[40:700], bi[2:4], idx[700:], bi[4:]]).astype(np.int32)
    fr.grid_log2 = grid_log2
    o = _check(fd, oracle, fr)
    assert o.exit == 0
    tr = o.tri_range
    w = tr["ex"].astype(np.int64) - tr["sx"]
    h = tr["ey"].astype(np.int64) - tr["sy"]
    assert int(((w > 0) & (h > 0) & (w * h > 64)).sum()) >= 3, "the frame must hold triangles with more than 64 bins"
