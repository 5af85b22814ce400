"""The C++ mirror of the reference's FD:: interface (include/fd_api.hpp),
driven in the reference application's call order by the prebuilt
4dzb_amd/fd_api_check program, against the oracle (bit-exact) on two
frames.  The program runs as a child process of the test."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "4dzb_amd", "fd_api_check")


def _write_frame(path, frame):
    sc, rx, L = frame.scene, frame.rx, frame.light
    with open(path, "wb") as f:
        np.array([rx.w, rx.h, sc.nverts, sc.ntris, L.res_w, frame.backface_cull], np.int32).tofile(f)
        np.array([frame.eps, L.right_radius, L.top_radius, *L.position, *L.view_dir, *L.up_dir],
                 np.float32).tofile(f)
        np.ascontiguousarray(sc.verts, np.float32).tofile(f)
        np.ascontiguousarray(sc.idx, np.int32).tofile(f)
        np.ascontiguousarray(rx.pos, np.float32).tofile(f)
        np.ascontiguousarray(rx.nrm, np.float32).tofile(f)


@pytest.mark.parametrize("case", ["plane", "heightfield"])
def test_fd_namespace_mirror_matches_oracle(tmp_path, wl, oracle, case):
    if not os.path.exists(EXE):
        pytest.fail("fd_api_check not built (make -C 4dzb_amd)")
    if case == "plane":
        frame = wl.plane_256(ntri=3000, res=128)
    else:
        frame = wl.heightfield_1k(res=192, nsph=60, light_res=24)
    src, dst = tmp_path / "frame.bin", tmp_path / "out.bin"
    _write_frame(src, frame)
    print(f"[fd_api_check] {case}", flush=True)
    p = subprocess.run([EXE, str(src), str(dst)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    W, H, R = frame.rx.w, frame.rx.h, frame.light.res_w
    words = R * (1 if R <= 32 else 2)
    raw = np.fromfile(dst, np.uint8)
    exit_code = int(raw[:4].view(np.int32)[0])
    shadow = raw[8:8 + 4 * W * H].view(np.float32)
    masks = raw[8 + 4 * W * H:].view(np.uint32)
    o = oracle.run(frame)
    assert exit_code == o.exit == 0
    assert np.array_equal(shadow.view(np.uint32), np.asarray(o.shadow, np.float32).view(np.uint32))
    assert np.array_equal(masks, np.asarray(o.masks, np.uint32).ravel())
    assert masks.size == W * H * words
    assert np.count_nonzero(masks) > 0
