"""Stage-by-stage comparison of the HIP path with the CPU oracle.

Used by the GPU tests and by __graft_entry__.smoke().  Every comparison is
bit-exact: float arrays are compared by bit pattern (NaN payloads included),
integer arrays by value.
"""
from __future__ import annotations

import numpy as np


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _eq_f32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(_bits(a), _bits(b))


def _ndiff_f32(a, b):
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    if a.shape != b.shape:
        return -1
    return int(np.count_nonzero(_bits(a) != _bits(b)))


def run_gpu(fd, frame, pair_capacity=None):
    fr = fd.FrameRunner(frame, pair_capacity=pair_capacity)
    rc, st = fr.run()
    return fr, rc, st


def compare(fd, oracle, frame, subset=None, stages=True, fr=None, o=None):
    """Run frame on the GPU and on the oracle; return (report, fr, o).

    report maps a check name to its number of mismatching elements (0 = ok).
    With `subset`, the oracle evaluates the occlusion only for those dense
    receiver ids and outputs are compared on them."""
    if o is None:
        o = oracle.run(frame, subset=subset)
    if fr is None:
        fr, rc, st = run_gpu(fd, frame)
    else:
        rc, st = fr.run()
    rep = {}
    rep["rc"] = int(rc != 0 and not (o.exit == 9))
    rep["exit"] = int(st.exit_code != o.exit)
    sh, mk = fr.outputs()
    R = frame.light.res_w
    wpr = 1 if R <= 32 else 2
    WH = frame.rx.w * frame.rx.h
    if subset is None:
        rep["shadow"] = _ndiff_f32(sh, o.shadow)
        if mk is not None:
            rep["masks"] = int(np.count_nonzero(mk != o.masks))
    else:
        ids = np.asarray(subset, np.int64)
        rep["shadow"] = _ndiff_f32(sh[ids], o.shadow[ids])
        if mk is not None:
            m1 = mk.reshape(WH, R * wpr)[ids]
            m2 = o.masks.reshape(WH, R * wpr)[ids]
            rep["masks"] = int(np.count_nonzero(m1 != m2))
    if not stages:
        return rep, fr, o
    ctx = fr.ctx
    L = fd.read_light(ctx)
    lb = o.light_bits
    rep["light.mat"] = _ndiff_f32(L["m"], lb[0:16])
    rep["light.xd_yd"] = _ndiff_f32(np.concatenate([L["xd"], L["yd"]]), lb[22:28])
    rep["light.corners"] = _ndiff_f32(np.concatenate([L["ur"], L["ul"], L["dl"], L["dr"]]), lb[28:40])
    rep["light.F"] = _ndiff_f32(L["F"], lb[40:42])
    if o.exit in (1, 2):
        return rep, fr, o
    s = fd.read_state(ctx)
    rep["state.f_S"] = _ndiff_f32([s["f"], s["S"][0], s["S"][1]], lb[42:45])
    rep["state.sAABB"] = _ndiff_f32(s["sAABB"][[0, 1, 2, 4, 5, 6]], o.sample_aabb[[0, 1, 2, 4, 5, 6]])
    rep["state.valid_rect"] = int(s["validRect"] != int(o.scalars["valid_rect"]))
    rect = ctx.debug_read("rect", np.float32)
    rep["rect"] = _ndiff_f32(rect, o.rect)
    if o.exit == 3:
        return rep, fr, o
    rep["state.mAABB"] = _ndiff_f32(s["mAABB"][[0, 1, 2, 4, 5, 6]], o.model_aabb[[0, 1, 2, 4, 5, 6]])
    if o.exit == 4:
        return rep, fr, o
    rep["state.lp"] = _ndiff_f32(np.concatenate([s["B"], s["E"]]), lb[45:51])
    if o.exit == 5:
        return rep, fr, o
    rep["state.factor"] = _ndiff_f32(s["factor"], lb[51:53])
    # triangle setup
    aabb = ctx.debug_read("tri_aabb", np.float32)
    rep["tri_aabb"] = _ndiff_f32(aabb, o.tri_aabb)
    rep["valid_tris"] = int(s["validTris"] != int(o.scalars["valid_tris"]))
    if o.exit == 6:
        return rep, fr, o
    tv = o.tri_valid.astype(bool)
    td = fd.read_tris(ctx)
    od = o.tri_data
    g = td[tv]
    h = od[tv]
    rep["tri_data"] = (_ndiff_f32(g["v0"], h["v0"]) + _ndiff_f32(g["v1"], h["v1"]) + _ndiff_f32(g["v2"], h["v2"]) +
                       _ndiff_f32(g["bminx"], h["aabb_min"][:, 0]) + _ndiff_f32(g["bminy"], h["aabb_min"][:, 1]) +
                       _ndiff_f32(g["bmaxx"], h["aabb_max"][:, 0]) + _ndiff_f32(g["bmaxy"], h["aabb_max"][:, 1]) +
                       _ndiff_f32(g["minz"], h["minz"]) + _ndiff_f32(g["maxz"], h["maxz"]))
    # receiver binning
    key = ctx.debug_read("key", np.int32)
    rep["key"] = int(np.count_nonzero(key != o.key))
    rep["valid_samples"] = int(s["validSamples"] != int(o.scalars["valid_samples"]))
    sid = ctx.debug_read("sorted_id", np.int32)
    rep["sorted_id"] = int(np.count_nonzero(sid != o.sorted_id))
    rep["bin_start"] = int(np.count_nonzero(ctx.debug_read("bin_start", np.int32) != o.bin_start))
    rep["bin_end"] = int(np.count_nonzero(ctx.debug_read("bin_end", np.int32) != o.bin_end))
    rep["bin_maxz"] = _ndiff_f32(ctx.debug_read("bin_maxz", np.float32), o.bin_maxz)
    rep["bin_minrange"] = _ndiff_f32(ctx.debug_read("bin_minrange", np.float32), o.bin_minrange)
    # triangle -> bin pairing
    cnt = ctx.debug_read("tri_count", np.int32)
    rep["tri_count"] = int(np.count_nonzero(cnt != o.tri_count))
    rg = ctx.debug_read("tri_range", np.uint32).reshape(-1, 2)
    orr = o.tri_range
    grg = np.stack([rg[:, 0] & 0xFFFF, rg[:, 0] >> 16, rg[:, 1] & 0xFFFF, rg[:, 1] >> 16], 1).astype(np.int64)
    org = np.stack([orr["sx"], orr["sy"], orr["ex"], orr["ey"]], 1).astype(np.int64)
    rep["tri_range"] = int(np.count_nonzero(grg != org))
    rep["pairs"] = int(int(s["pairs"]) != int(o.scalars["pairs"]))
    if o.exit == 7:
        return rep, fr, o
    pt = ctx.debug_read("pair_tri", np.int32)
    rep["pair_tri"] = int(pt.shape != o.pair_tri.shape or np.count_nonzero(pt != o.pair_tri))
    rep["bin_tri_start"] = int(np.count_nonzero(ctx.debug_read("bin_tri_start", np.int32) != o.bin_tri_start))
    rep["bin_tri_end"] = int(np.count_nonzero(ctx.debug_read("bin_tri_end", np.int32) != o.bin_tri_end))
    return rep, fr, o


def failures(rep):
    return {k: v for k, v in rep.items() if v}
