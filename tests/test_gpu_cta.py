"""Bin-queue (CTA) occlusion variant on the GPU against the oracle's
restatement of shadowCal_CTA_kernel (Step5_2_shadowCal.inl:408-1100),
bit-exact (MI355X only).  The variant is selected per context
(fd_set_occlusion_variant); its masks are rows along y_delt."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(wl):
    out = []
    fr = wl.plane_256(ntri=3000, res=96, light_res=16)
    fr.name = "plane-96-16"
    out.append(fr)
    # occluders that cross both the receivers' depth and the light: the
    # groups' triangle-plane decision changes bits here
    fr = wl.plane_256(ntri=60, res=64, light_res=16)
    fr.scene = wl.tris_uniform(60, 35, box=(-1.5, 1.5, -1.5, 1.5, -0.5, 5.0), e0=1.0, e1=4.0)
    fr.backface_cull = 0
    fr.name = "crossing-64-16"
    out.append(fr)
    fr = wl.Frame("hf-160-32", wl.spheres(60, 11), wl.receivers_heightfield(160, 120, 8.0, 6.0),
                  wl.LightDesc(1.0, 1.0, (0.2, -0.3, 5.5), res_w=32, res_h=32), eps=1e-3, backface_cull=1)
    out.append(fr)
    fr = wl.Frame("hf-128-40", wl.spheres(40, 12), wl.receivers_heightfield(128, 128, 8.0, 8.0),
                  wl.LightDesc(1.0, 1.0, (0.0, 0.3, 5.0), res_w=40, res_h=40), eps=1e-3, backface_cull=1)
    out.append(fr)
    fr = wl.Frame("hf-96-64", wl.spheres(40, 13), wl.receivers_heightfield(96, 96, 8.0, 8.0),
                  wl.LightDesc(1.0, 1.0, (0.4, 0.1, 5.0), res_w=64, res_h=64), eps=1e-3, backface_cull=1)
    out.append(fr)
    # large enough for one work item per chunk (no split triangle lists)
    fr = wl.plane_256(ntri=4000, res=384, light_res=8)
    fr.name = "plane-384-8"
    out.append(fr)
    return out


@pytest.mark.parametrize("k", range(6))
def test_bin_queue_matches_oracle(fd, oracle, wl, k):
    fr = _frames(wl)[k]
    print(f"[cta] {fr.name} {fr.rx.w}x{fr.rx.h} light {fr.light.res_w} tris {fr.scene.ntris}", flush=True)
    o = oracle.run(fr, variant=1)
    assert o.exit == 0
    run = fd.FrameRunner(fr)
    run.ctx.set_occlusion_variant(fd.OCC_BIN_QUEUE)
    rc, st = run.run()
    assert rc == 0 and st.exit_code == 0
    sh, mk = run.outputs()
    run.close()
    assert np.array_equal(sh.view(np.uint32), np.asarray(o.shadow, np.float32).view(np.uint32))
    assert np.array_equal(mk, np.asarray(o.masks, np.uint32))
    assert (sh > 0).mean() > 0.02


def test_variant_switch_back(fd, oracle, wl):
    """The same context returns to the per-sample kernel's bits."""
    fr = wl.plane_256(ntri=2000, res=64, light_res=16)
    o0, o1 = oracle.run(fr, variant=0), oracle.run(fr, variant=1)
    run = fd.FrameRunner(fr)
    outs = []
    for v in (fd.OCC_BIN_QUEUE, fd.OCC_PER_SAMPLE, fd.OCC_BIN_QUEUE):
        run.ctx.set_occlusion_variant(v)
        rc, _ = run.run()
        assert rc == 0
        outs.append(run.outputs()[1].copy())
    run.close()
    assert np.array_equal(outs[0], o1.masks) and np.array_equal(outs[2], o1.masks)
    assert np.array_equal(outs[1], o0.masks)
    assert not np.array_equal(o0.masks, o1.masks)
