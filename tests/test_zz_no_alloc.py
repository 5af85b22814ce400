"""The allocation-free launch (fd_launch_async_noalloc, workspace in the
library's static device scratch) against the existing entry point
(fd_launch_async with the context's own workspace), bit for bit, through the
C ABI (MI355X only)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(wl):
    a = wl.plane_256(ntri=10_000, res=256, light_res=16)
    b = wl.heightfield_1k(res=160, nsph=40, light_res=40)   # smaller input, 64-bit rows
    b.light.view_dir, b.light.up_dir = (0.2, -0.1, -1.0), (0.1, 1.0, 0.0)
    return a, b


def _stats(fd, ctx):
    st = fd.Stats()
    rc = fd.lib().fd_get_stats(ctx.h, None, ctypes.byref(st))
    return rc, (st.exit_code, st.pairs, st.valid_receivers, st.valid_triangles, st.work_items)


def test_noalloc_equals_launch(fd, wl):
    lib = fd.lib()
    runs = [fd.FrameRunner(f) for f in _frames(wl)]
    want = []
    for r in runs:   # the existing entry point
        lib.fd_launch_async(r.ctx.h, None)
        rc, st = _stats(fd, r.ctx)
        assert rc == 0 and st[0] == 0
        sh, mk = r.outputs()
        want.append((sh.view(np.uint32).copy(), mk.copy(), st))
    for r in runs:   # poison the outputs
        r.shadow.upload(np.full(r.shadow.nbytes // 4, -5.0, np.float32))
        r.masks.upload(np.full(r.masks.nbytes // 4, 0xA5A5A5A5, np.uint32))
    # two allocation-free calls in a row, the second with the smaller input
    for k, r in enumerate(runs):
        print(f"[noalloc] call {k}: {r.frame.rx.w}x{r.frame.rx.h} light {r.frame.light.res_w} "
              f"tris {r.frame.scene.ntris}", flush=True)
        assert lib.fd_launch_async_noalloc(r.ctx.h, None) == 0
        rc, st = _stats(fd, r.ctx)
        assert rc == 0 and st == want[k][2], (k, st, want[k][2])
    for k, r in enumerate(runs):
        sh, mk = r.outputs()
        assert np.array_equal(sh.view(np.uint32), want[k][0]), k
        assert np.array_equal(mk, want[k][1]), k
    for r in runs:
        r.close()


def test_noalloc_rejects_oversized_frame(fd, wl):
    """A frame whose workspace exceeds the scratch is refused before any
    launch (1M triangles: their records alone exceed 128 MiB)."""
    f = wl.plane_256(ntri=10, res=64)
    f.scene = wl.tris_clusters(1_000_000, 3)
    r = fd.FrameRunner(f)
    r.shadow.upload(np.full(64 * 64, -5.0, np.float32))
    assert fd.lib().fd_launch_async_noalloc(r.ctx.h, None) == fd.FD_ERR_SCRATCH
    fd.lib().fd_device_sync()
    assert np.all(r.shadow.download(np.float32) == -5.0)   # nothing ran
    r.close()
