"""FD::moveModel on the device (fd_move_model, FDc/moveModel.inl:6-30) against
the oracle's restatement, bit for bit (MI355X only): single and cumulative
moves, 16- and 32-byte vertex strides (the floats after x, y, z untouched),
and moves made through the serving loop."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mat(rng, kind):
    if kind == "rigid":
        a, b = rng.uniform(-np.pi, np.pi, 2)
        c, s = np.cos(a), np.sin(a)
        M = np.eye(4)
        M[:3, :3] = [[c, -s, 0], [s, c, 0], [0, 0, 1]]
        cb, sb = np.cos(b), np.sin(b)
        M[:3, :3] = M[:3, :3] @ np.array([[1, 0, 0], [0, cb, -sb], [0, sb, cb]])
        M[:3, 3] = rng.uniform(-2, 2, 3)
    else:
        M = rng.normal(size=(4, 4))
    return np.asarray(M, np.float32).T.ravel()   # column-major


@pytest.mark.parametrize("stride", [4, 8])
def test_move_model_matches_oracle(fd, oracle, wl, stride):
    rng = np.random.default_rng(30 + stride)
    sc = wl.spheres(40, 3)
    V = np.zeros((sc.nverts, stride), np.float32)
    V[:, :3] = sc.verts[:, :3]
    V[:, 3:] = rng.normal(size=(sc.nverts, stride - 3))
    ctx = fd.Context()
    dv = fd.DeviceBuffer.from_array(V)
    di = fd.DeviceBuffer.from_array(sc.idx)
    ctx.add_model(dv, sc.nverts, stride * 4, di, sc.ntris)
    want = V.copy()
    for step, kind in enumerate(("rigid", "general", "rigid")):
        m = _mat(rng, kind)
        print(f"[move] stride {stride} step {step} {kind}", flush=True)
        ctx.move_model(m)
        want = oracle.move_model(want, m)
        got = dv.download(np.float32).reshape(V.shape)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (step, kind)
    assert np.array_equal(want[:, 3:], V[:, 3:])
    ctx.close()


def test_server_moves_match_oracle(fd, oracle, wl):
    """Scene moves through the serving loop transform the resident vertices
    exactly as the oracle does; another resident scene stays as it was."""
    rng = np.random.default_rng(41)
    scenes = [wl.spheres(30, 5), wl.spheres(20, 6)]
    srv = fd.Server(budget_bytes=1 << 30, streams=2, max_batch=2, eps=1e-3)
    bufs = []
    for sc in scenes:
        v, i = fd.DeviceBuffer.from_array(sc.verts), fd.DeviceBuffer.from_array(sc.idx)
        bufs.append((v, i))
        srv.add_scene(v, sc.nverts, 16, i, sc.ntris)
    want = scenes[0].verts.copy()
    for k in range(3):
        m = _mat(rng, "rigid" if k % 2 == 0 else "general")
        srv.move_scene(0, m, tag=k)
        want = oracle.move_model(want, m)
    srv.drain(60000)
    done = srv.poll()
    assert sorted(c.user_tag for c in done) == [0, 1, 2] and all(c.status == 0 for c in done)
    got = bufs[0][0].download(np.float32).reshape(want.shape)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    other = bufs[1][0].download(np.float32).reshape(scenes[1].verts.shape)
    assert np.array_equal(other.view(np.uint32), scenes[1].verts.view(np.uint32))
    srv.close()
