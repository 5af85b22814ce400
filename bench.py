#!/usr/bin/env python3
"""Soft-shadow visibility benchmark on one MI355X.

A "step" is one frame of the plain per-frame path (FD::FDLaunch's
replacement) over the boxes-1080p workload, the largest frame that one GPU
runs: 1920x1080 receivers from a ray-cast G-buffer (ground plane + 20
boxes), a 3x1 area light at z = 6 sampled 64x64 (4096-bit masks), and 1M
clustered occluder triangles.  Inputs are resident in HBM before the timed
region; each frame writes the per-receiver shadow value and the 4096-bit
light-sample mask of every receiver (1.06 GB).

value = visibility Gbit/s = frames * W*H * Lw*Lh / seconds / 1e9.

The heightfield-1k frame (1024x1024 receivers, 32x32 light, 200k triangles)
is measured the same way in its own object.  Each plain frame's roofline
object describes the dominant kernel (the occlusion kernel): its duration
comes from HIP events recorded around it on the launch stream, its work from
the kernel's own counters (candidate tests, planes set up, light rows)
priced with per-event operation counts of the reference's arithmetic, with
the rows the reference visits (achieved) and the rows this kernel visits
(achieved_kernel_rows).  cpu_baseline times the CPU restatement of the
reference algorithm on this host, single-threaded and on all cores, over a
seeded receiver sample.
"""
import argparse
import copy
import importlib
import json
import os
import sys
import time

import numpy as np
import xxhash

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "visibility Gbit/s + p50/p99 frame ms, batched-served vs plain, bit-exact masks"
# algorithmic operations per counted event of the reference's occlusion
# arithmetic (Step5_2_shadowCalPerSample.inl:110-223): containment/depth
# cull per candidate, planes (93) + row-range masks (18) per set-up pair,
# one light row per row the reference visits.  Rows this kernel skips (no
# set bit possible) still count: the work is the reference algorithm's.
OPS_PER_CAND, OPS_PER_PLANES, OPS_PER_ROW = 5, 93 + 18, 15
FP32_VECTOR_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


PROFILED = {("heightfield-1k", 32): "k_occlude_heightfield-1k.md",
            ("boxes-1080p", 32): "k_occlude_boxes-1080p-32x32.md",
            ("boxes-1080p", 64): "k_occlude_boxes-1080p-64x64.md"}


def profiled_traffic(workload, light_res):
    """HBM bytes per occlusion launch and the counter-derived VALU issue and
    occupancy from the committed rocprofv3 summary of the same workload
    (FETCH_SIZE and WRITE_SIZE passes, calibrated with a 256 MiB copy; SQ
    counters), or None when no summary covers it."""
    name = PROFILED.get((workload, light_res))
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    try:
        blob = open(path).read().split("```json", 1)[1].split("```", 1)[0]
        return dict(json.loads(blob), source=os.path.relpath(path, ROOT))
    except (IndexError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pctl(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q))


def gpu_plain(pkg, frame, steps, warmup, quick=False):
    fd = pkg.fd4d
    strm = fd.Stream()
    run = fd.FrameRunner(frame, stream=strm)
    ctx = run.ctx
    stream = strm.handle
    W, H, R = frame.rx.w, frame.rx.h, frame.light.res_w
    bits = W * H * R * R
    rc, st = run.run()
    if rc != 0:
        raise RuntimeError(f"frame failed: rc {rc}")
    for _ in range(warmup):
        ctx.launch_async(stream)
    strm.sync()
    # throughput: frames enqueued back to back
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.launch_async(stream)
    strm.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / steps
    rc, st = ctx.get_stats(stream)
    if quick:
        # a device-to-device copy of known size: calibrates HBM byte counters
        # (FETCH_SIZE / WRITE_SIZE) in profiler runs
        nb = 256 << 20
        src, dst = fd.DeviceBuffer(nb), fd.DeviceBuffer(nb)
        fd.lib().fd_memcpy_dtod(fd.ctypes.c_void_p(dst.ptr), fd.ctypes.c_void_p(src.ptr), nb, fd.ctypes.c_void_p(stream))
        strm.sync()
        src.close()
        dst.close()
        return dict(run=run, bits=bits, ms_per_step=ms_per_step, lat=[ms_per_step], stage=[0.0] * 5,
                    ops=[0] * 8, ops_ref=[0] * 8, walk_order=0, stats=st.as_dict(), shadow=None, masks=None)
    # latency: one frame at a time, host wall clock around fd_launch
    lat = []
    for _ in range(max(steps, 10)):
        a = time.perf_counter()
        rc, st = ctx.launch(stream)
        lat.append((time.perf_counter() - a) * 1e3)
    # stage times (HIP events around each stage on the launch stream)
    ctx.set_profiling(True, False)
    stage = []
    for _ in range(5):
        ctx.launch_async(stream)
        stage.append(ctx.get_profile()[0])
    stage = np.median(np.asarray(stage), axis=0)
    walk_order = int(fd.read_state(ctx)["triOrder"])
    # occlusion work counters (separate frames: counting adds instructions),
    # in the order the frame walks the bins' lists and in the reference's
    # (ascending triangle id)
    ctx.set_profiling(False, True)
    ctx.launch_async(stream)
    _, ops = ctx.get_profile(ops=True)
    ctx.set_triangle_order(fd.TRI_ORDER_LIST)
    ctx.launch_async(stream)
    _, ops_ref = ctx.get_profile(ops=True)
    ctx.set_triangle_order(fd.TRI_ORDER_AUTO)
    ctx.set_profiling(False, False)
    # the outputs checked below come from the timed kernel, not the counting one
    ctx.launch_async(stream)
    strm.sync()
    shadow, masks = run.outputs()
    return dict(run=run, bits=bits, ms_per_step=ms_per_step, lat=lat, stage=stage.tolist(), ops=ops,
                ops_ref=ops_ref, walk_order=walk_order, stats=st.as_dict(), shadow=shadow, masks=masks)


def make_frame(wl, workload, light_res):
    make = {"heightfield-1k": wl.heightfield_1k, "plane-256": wl.plane_256, "boxes-1080p": wl.boxes_1080p}[workload]
    return make(light_res=light_res) if light_res else make()


def occlusion_roofline(g, occ_ms, workload, light_res):
    """Roofline object of the occlusion kernel: algorithmic operations per
    launch from the kernel's counters (the reference's work: its list order
    and the rows it visits; and, separately, the work this kernel does in
    its walk order with its row window) over the kernel's event-timed
    duration, against the FP32 vector peak."""
    cand, npass, rows, items, wave_it, wave_rows, ref_rows, planes = g["ops"]
    rcand, rpass, _, _, rwave_it, _, rref_rows, rplanes = g["ops_ref"]
    ops_ref = OPS_PER_CAND * rcand + OPS_PER_PLANES * rplanes + OPS_PER_ROW * rref_rows
    ops_run = OPS_PER_CAND * cand + OPS_PER_PLANES * planes + OPS_PER_ROW * rows
    achieved = ops_ref / (occ_ms * 1e-3) / 1e12
    achieved_run = ops_run / (occ_ms * 1e-3) / 1e12
    prof = profiled_traffic(workload, light_res)
    return dict(bound="valu", achieved=round(achieved, 3), peak=FP32_VECTOR_PEAK_TFLOPS, unit="TOP/s",
                frac=round(achieved / FP32_VECTOR_PEAK_TFLOPS, 4),
                traffic=prof["traffic_bytes"] if prof else None,
                achieved_kernel_rows=round(achieved_run, 3),
                frac_kernel_rows=round(achieved_run / FP32_VECTOR_PEAK_TFLOPS, 4),
                work_model=("achieved: 5 ops per candidate test + 111 per pair whose planes are set up + 15 per "
                            "light row the reference visits, all counted in the reference's list order (work this "
                            "kernel skips by its walk order and row window counts); achieved_kernel_rows: the "
                            "same with the tests, planes and rows this kernel does in its walk order"),
                traffic_source=prof["source"] if prof else None,
                valu_issue_profiled=prof.get("valu_issue") if prof else None,
                occupancy_profiled=prof.get("occupancy") if prof else None,
                kernel="k_occlude", kernel_ms=round(occ_ms, 4),
                walk_order={1: "list", 2: "area classes", 3: "mesh (area classes, largest last)"}.get(g["walk_order"], "n/a"),
                work=dict(candidate_tests=int(cand), cull_passes=int(npass), row_visits=int(rows),
                          work_items=int(items), wave_tri_iterations=int(wave_it),
                          wave_row_iterations=int(wave_rows),
                          lane_util_rows=round(rows / max(1, 64 * wave_rows), 4), planes=int(planes),
                          ops_per_event=[OPS_PER_CAND, OPS_PER_PLANES, OPS_PER_ROW]),
                reference_work=dict(candidate_tests=int(rcand), cull_passes=int(rpass), planes=int(rplanes),
                                    row_visits=int(rref_rows), wave_tri_iterations=int(rwave_it)))


def plain_leg(pkg, frame, workload, steps, warmup, quick):
    """One plain frame configuration: throughput (frames back to back),
    latency, stage times and the occlusion roofline."""
    g = gpu_plain(pkg, frame, steps, warmup, quick)
    W, H, R = frame.rx.w, frame.rx.h, frame.light.res_w
    value = g["bits"] / (g["ms_per_step"] * 1e-3) / 1e9
    occ_ms = max(g["stage"][4], 1e-9)
    leg = dict(gbps=round(value, 4), ms_per_step=round(g["ms_per_step"], 4),
               config=dict(workload=workload, receivers=f"{W}x{H}", light_samples=f"{R}x{R}",
                           triangles=frame.scene.ntris, bins=f"{1 << frame.grid_log2}x{1 << frame.grid_log2}",
                           tile_width=frame.tile_width,
                           occ_waves=frame.occ_waves or pkg.autotune.default_waves(frame.light.res_w, frame.rx.w,
                                                                                   frame.rx.h, 1 << frame.grid_log2),
                           path="plain per-frame"),
               p50_ms=round(pctl(g["lat"], 50), 3), p99_ms=round(pctl(g["lat"], 99), 3),
               stage_ms=dict(zip(["light_plane", "receiver_binning", "triangle_pairing", "tables", "occlusion"],
                                 [round(x, 4) for x in g["stage"]])),
               frame_stats=dict(valid_receivers=g["stats"]["valid_receivers"], pairs=g["stats"]["pairs"],
                                valid_triangles=g["stats"]["valid_triangles"]),
               roofline=occlusion_roofline(g, occ_ms, workload, R))
    return g, leg


def oracle_sample_check(frame, g, n):
    """The GPU frame's masks and shadow values equal the oracle's on a seeded
    receiver sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    W, H, R = frame.rx.w, frame.rx.h, frame.light.res_w
    ids = np.sort(np.random.default_rng(99).choice(W * H, n, replace=False))
    o = oracle.run(frame, subset=ids)
    wpr = 1 if R <= 32 else 2
    return bool(np.array_equal(g["masks"].reshape(W * H, R * wpr)[ids], o.masks.reshape(W * H, R * wpr)[ids])
                and np.array_equal(g["shadow"][ids].view(np.uint32), o.shadow[ids].view(np.uint32)))


def gpu_batched(pkg, reps=3):
    """batch-64: 64 instances of uneven resolution / light grid / light pose
    over the heightfield-1k scene.  Plain = each instance through the
    per-frame path (fd_launch, blocking like FDLaunch, and back-to-back
    fd_launch_async); batched = one graph-replayed launch per stage."""
    fd, wl = pkg.fd4d, pkg.workloads
    frames = wl.batch_instances(64)
    scene = frames[0].scene
    strm = fd.Stream()
    ctx = fd.Context()
    ctx.set_backface_cull(1)
    ctx.set_scenes_esp(frames[0].eps)
    verts = fd.DeviceBuffer.from_array(scene.verts)
    idx = fd.DeviceBuffer.from_array(scene.idx)
    ctx.add_model(verts, scene.nverts, 16, idx, scene.ntris)
    rxbuf = {}
    items = []
    bits = 0
    for f in frames:
        key = (f.rx.w, f.rx.h)
        if key not in rxbuf:
            rxbuf[key] = (fd.DeviceBuffer.from_array(f.rx.pos), fd.DeviceBuffer.from_array(f.rx.nrm))
        R = f.light.res_w
        items.append(dict(w=f.rx.w, h=f.rx.h, light=f.light, pos=rxbuf[key][0], nrm=rxbuf[key][1], bins=0,
                          shadow=fd.DeviceBuffer(f.rx.w * f.rx.h * 4),
                          masks=fd.DeviceBuffer(f.rx.w * f.rx.h * fd.mask_words(R) * 4)))
        bits += f.rx.w * f.rx.h * R * R
    # plain per-frame path through one context (sized by a warm-up pass)
    def plain(sync_each):
        for it in items:
            ctx.set_receivers(it["pos"], it["nrm"], it["w"], it["h"])
            ctx.set_outputs(it["shadow"], it["masks"])
            ctx.set_area_light(it["light"])
            if sync_each:
                rc, _ = ctx.launch(strm.handle)
                assert rc == 0
            else:
                ctx.launch_async(strm.handle)
        strm.sync()
    plain(True)
    ref = [(it["shadow"].download(np.float32), it["masks"].download(np.uint32)) for it in items]
    t = []
    for _ in range(reps):
        a = time.perf_counter()
        plain(True)
        t.append(time.perf_counter() - a)
    plain_sync = min(t)
    t = []
    for _ in range(reps):
        a = time.perf_counter()
        plain(False)
        t.append(time.perf_counter() - a)
    plain_async = min(t)
    # autotuner: per class (receiver size, light grid) the fastest bit-exact
    # layout (bins x tile width x waves) and walk order, measured on this
    # device (untimed)
    at = pkg.autotune
    t_tune = time.perf_counter()
    cache = at.tune_classes(ctx, 0, items, orders=True)
    t_tune = time.perf_counter() - t_tune
    for it in items:
        it["bins"], it["tile"], it["waves"] = at.layout(cache, 0, it["w"], it["h"], it["light"].res_w)
        it["order"] = at.order(cache, 0, it["w"], it["h"], it["light"].res_w)
    # batched, graph-captured
    b = fd.Batch(ctx, items)
    b.capture(strm.handle)
    b.launch(strm.handle)
    strm.sync()
    exact = all(np.array_equal(it["shadow"].download(np.float32).view(np.uint32), r[0].view(np.uint32)) and
                np.array_equal(it["masks"].download(np.uint32), r[1]) for it, r in zip(items, ref))
    t = []
    for _ in range(reps):
        a = time.perf_counter()
        b.launch(strm.handle)
        strm.sync()
        t.append(time.perf_counter() - a)
    batched = min(t)
    b.close()
    # per instance class (receiver size, light grid): the class's instances
    # back to back through the plain path vs as one graph-replayed batch
    classes = {}
    for k, it in enumerate(items):
        classes.setdefault((it["w"], it["h"], it["light"].res_w), []).append(k)
    per_class = {}
    for key, ks in sorted(classes.items()):
        sub = [items[k] for k in ks]
        tp = []
        for _ in range(reps):
            a = time.perf_counter()
            for it in sub:
                ctx.set_receivers(it["pos"], it["nrm"], it["w"], it["h"])
                ctx.set_outputs(it["shadow"], it["masks"])
                ctx.set_area_light(it["light"])
                ctx.launch_async(strm.handle)
            strm.sync()
            tp.append(time.perf_counter() - a)
        bc = fd.Batch(ctx, sub)
        bc.capture(strm.handle)
        bc.launch(strm.handle)
        strm.sync()
        tb = []
        for _ in range(reps):
            a = time.perf_counter()
            bc.launch(strm.handle)
            strm.sync()
            tb.append(time.perf_counter() - a)
        bc.close()
        cb = sum(it["w"] * it["h"] * it["light"].res_w ** 2 for it in sub) / 1e9
        per_class[f"{key[0]}x{key[1]}:{key[2]}"] = dict(
            instances=len(sub), plain_ms=round(min(tp) * 1e3, 3), batched_ms=round(min(tb) * 1e3, 3),
            plain_gbps=round(cb / min(tp), 2), batched_gbps=round(cb / min(tb), 2),
            speedup=round(min(tp) / min(tb), 3))
    ctx.close()
    gb = bits / 1e9
    return dict(workload="batch-64 (heightfield-1k scene, 200k triangles; receivers from {256^2, 512^2, 1024^2, "
                "1920x1080}, light grids from {8^2, 16^2, 32^2}, poses U[-2,2]^2 x [4,8])",
                instances=len(items), gbit_per_batch=round(gb, 4),
                plain_sync_ms=round(plain_sync * 1e3, 2), plain_async_ms=round(plain_async * 1e3, 2),
                batched_graph_ms=round(batched * 1e3, 2),
                plain_sync_gbps=round(gb / plain_sync, 3), plain_async_gbps=round(gb / plain_async, 3),
                batched_gbps=round(gb / batched, 3), speedup_vs_plain_sync=round(plain_sync / batched, 3),
                speedup_vs_plain_async=round(plain_async / batched, 3), masks_bit_exact_vs_plain=bool(exact),
                per_class=per_class,
                autotune=dict(seconds=round(t_tune, 2), classes={k: dict(bins=v["bins"], tile=v["tile"],
                                                                        waves=v.get("waves"),
                                                                        order={1: "list", 2: "area", 3: "mesh"}.get(v.get("order"), "auto"),
                                                                        ms=v["ms"], default_ms=v["default_ms"])
                                                               for k, v in sorted(cache.items())}))


def _digest(bufs):
    h = xxhash.xxh3_128()
    for b in bufs:
        h.update(b.download(np.uint8))
    return h.hexdigest()


def gpu_served(pkg, nreq=160, rate=500.0, budget_gb=48):
    """serve-poisson: Poisson arrivals at `rate` req/s over 4 resident sphere
    scenes (100k/300k/1M/2M triangles); resolution from {256^2, 512^2,
    1024^2, 1920x1080}, light grid 16^2 or 32^2, light pose U[-2,2]^2 x [4,8].
    Served = the continuous-batching server; plain = each request alone
    through fd_launch, queued FIFO on the same arrival times (its latency is
    computed from the measured per-request frame times)."""
    fd, wl = pkg.fd4d, pkg.workloads
    scenes = wl.serve_scenes()
    reqs = wl.poisson_requests(nreq, rate_hz=rate)
    dev_scenes = [(fd.DeviceBuffer.from_array(sc.verts), fd.DeviceBuffer.from_array(sc.idx)) for sc in scenes]
    rx = {}
    for q in reqs:
        key = (q["w"], q["h"])
        if key not in rx:
            r = wl.receivers_heightfield(q["w"], q["h"], 8.0, 8.0 * q["h"] / q["w"])
            rx[key] = (r, fd.DeviceBuffer.from_array(r.pos), fd.DeviceBuffer.from_array(r.nrm))
    outs = []
    bits = 0
    for q in reqs:
        R = q["grid"]
        outs.append((fd.DeviceBuffer(q["w"] * q["h"] * 4), fd.DeviceBuffer(q["w"] * q["h"] * fd.mask_words(R) * 4)))
        bits += q["w"] * q["h"] * R * R
    lights = [wl.LightDesc(1.0, 1.0, q["pose"], res_w=q["grid"], res_h=q["grid"]) for q in reqs]
    # plain per-request path: one context per scene, frame times measured alone
    strm = fd.Stream()
    ctxs = []
    for (v, i), sc in zip(dev_scenes, scenes):
        c = fd.Context()
        c.set_backface_cull(1)
        c.set_scenes_esp(1e-3)
        c.add_model(v, sc.nverts, 16, i, sc.ntris)
        ctxs.append(c)
    svc = []
    ref = []
    for k, q in enumerate(reqs):
        c = ctxs[q["scene"]]
        r, pos, nrm = rx[(q["w"], q["h"])]
        c.set_receivers(pos, nrm, q["w"], q["h"])
        c.set_outputs(*outs[k])
        c.set_area_light(lights[k])
        if k == 0:
            c.launch(strm.handle)
        a = time.perf_counter()
        rc, st = c.launch(strm.handle)
        svc.append((time.perf_counter() - a) * 1e3)
        assert rc == 0
        ref.append(_digest(outs[k]))
    # autotuner: per class (scene, size, grid) the layout on its first
    # request, the walk order over up to 8 of its requests, the list split
    at = pkg.autotune
    t_tune = time.perf_counter()
    cache = {}
    per_scene = {}
    for k, q in enumerate(reqs):
        r, pos, nrm = rx[(q["w"], q["h"])]
        per_scene.setdefault(q["scene"], []).append(
            dict(pos=pos, nrm=nrm, w=q["w"], h=q["h"], light=lights[k], shadow=outs[k][0], max_pairs=64 << 20))
    for sc, items in sorted(per_scene.items()):
        at.tune_classes(ctxs[sc], sc, items, cache, reps=2, orders=True, splits=True)
    t_tune = time.perf_counter() - t_tune
    for c in ctxs:
        c.close()
    arr = [q["t"] * 1e3 for q in reqs]
    t_free, lat_plain = 0.0, []
    for a, sv in zip(arr, svc):
        start = max(a, t_free)
        t_free = start + sv
        lat_plain.append(t_free - a)
    plain_span = t_free - arr[0]
    # served: the continuous-batching server at the given arrival times
    def serve(arrivals):
        # same pair arena per request as the plain path's contexts (64M pairs)
        srv = fd.Server(budget_bytes=budget_gb << 30, streams=4, max_batch=16, eps=1e-3, max_pairs=64 << 20,
                        tune=at.server_table(cache))
        for (v, i), sc in zip(dev_scenes, scenes):
            srv.add_scene(v, sc.nverts, 16, i, sc.ntris)
        for o in outs:
            o[0].upload(np.zeros(o[0].nbytes, np.uint8))
            o[1].upload(np.zeros(o[1].nbytes, np.uint8))
        t0 = srv.now_ms()
        for k, q in enumerate(reqs):
            target = t0 + arrivals[k]
            while srv.now_ms() < target:
                dt = (target - srv.now_ms()) / 1e3
                if dt > 0.0005:
                    time.sleep(dt - 0.0003)
            r, pos, nrm = rx[(q["w"], q["h"])]
            srv.submit(q["scene"], pos, nrm, q["w"], q["h"], lights[k], outs[k][0], outs[k][1], tag=k)
        srv.drain(600000)
        done = sorted(srv.poll(len(reqs) + 8), key=lambda c: c.user_tag)
        srv.close()
        lat = [c.t_done_ms - c.t_submit_ms for c in done]
        span = max(c.t_done_ms for c in done) - min(c.t_submit_ms for c in done)
        failed = sum(1 for c in done if c.status != 0)
        mismatched = sum(1 for k in range(len(reqs)) if _digest(outs[k]) != ref[k])
        return dict(lat=lat, span=span, failed=failed, mismatched=mismatched,
                    exact=failed == 0 and mismatched == 0 and len(done) == len(reqs),
                    mean_batch=float(np.mean([c.batch_size for c in done])))

    over = serve([a - arr[0] for a in arr])
    gb = bits / 1e9
    # a sustainable load: 60% of the capacity measured above, same requests
    cap_rps = len(reqs) / (over["span"] * 1e-3)
    rate2 = 0.6 * cap_rps
    arr2 = [(a - arr[0]) * rate / rate2 for a in arr]
    sus = serve(arr2)
    t_free, lat_plain2 = 0.0, []
    for a, sv in zip(arr2, svc):
        start = max(a, t_free)
        t_free = start + sv
        lat_plain2.append(t_free - a)
    return dict(workload=f"serve-poisson ({nreq} requests at {rate:.0f} req/s; scenes 100k/300k/1M/2M triangles; "
                "receivers from {256^2, 512^2, 1024^2, 1920x1080}; light grids 16^2 / 32^2)",
                requests=len(reqs), gbit=round(gb, 3), offered_span_ms=round(arr[-1] - arr[0], 1),
                served_span_ms=round(over["span"], 1), plain_span_ms=round(plain_span, 1),
                served_gbps=round(gb / (over["span"] * 1e-3), 3), plain_gbps=round(gb / (plain_span * 1e-3), 3),
                speedup=round(plain_span / over["span"], 3),
                served_p50_ms=round(pctl(over["lat"], 50), 2), served_p99_ms=round(pctl(over["lat"], 99), 2),
                plain_p50_ms=round(pctl(lat_plain, 50), 2), plain_p99_ms=round(pctl(lat_plain, 99), 2),
                plain_frame_ms_mean=round(float(np.mean(svc)), 2),
                mean_batch=round(over["mean_batch"], 2),
                masks_bit_exact_vs_plain=bool(over["exact"] and sus["exact"]), failed_requests=over["failed"],
                mismatched_requests=over["mismatched"],
                sustainable=dict(note="the same requests at 60% of the capacity measured at 500 req/s",
                                 rate_rps=round(rate2, 1), capacity_rps=round(cap_rps, 1),
                                 served_p50_ms=round(pctl(sus["lat"], 50), 2),
                                 served_p99_ms=round(pctl(sus["lat"], 99), 2),
                                 plain_p50_ms=round(pctl(lat_plain2, 50), 2),
                                 plain_p99_ms=round(pctl(lat_plain2, 99), 2),
                                 served_span_ms=round(sus["span"], 1), mean_batch=round(sus["mean_batch"], 2),
                                 failed_requests=sus["failed"], mismatched_requests=sus["mismatched"]),
                autotune=dict(seconds=round(t_tune, 2), classes=len(cache),
                              non_default=sum(1 for k, v in cache.items()
                                              if (v["bins"], v["tile"], v.get("waves"))
                                              != pkg.autotune.default_layout(int(k.split(":")[2]),
                                                                             *map(int, k.split(":")[1].split("x")))),
                              split_classes=sum(1 for v in cache.values() if v.get("split")),
                              orders={o: sum(1 for v in cache.values() if v.get("order") == i)
                                      for i, o in ((1, "list"), (2, "area"), (3, "mesh"))}))


def gpu_target_config(pkg, steps=10):
    """The latency target configuration: 1920x1080 receivers (boxes G-buffer),
    32x32 light samples, 1M triangles; plain path, one frame at a time."""
    fd, wl = pkg.fd4d, pkg.workloads
    frame = wl.boxes_1080p(light_res=32)
    strm = fd.Stream()
    # layout (bins per side x receiver tile width) from the autotuner: the
    # fastest candidate whose every output bit equals the default layout's
    run = fd.FrameRunner(frame, stream=strm)
    item = dict(pos=run.pos, nrm=run.nrm, w=frame.rx.w, h=frame.rx.h, light=frame.light, shadow=run.shadow,
                max_pairs=64 << 20)
    cands = pkg.autotune.CANDIDATES
    best, _ = run.ctx.autotune(item, cands, 3)
    bins, tile, waves = cands[best] if best >= 0 else pkg.autotune.default_layout(frame.light.res_w, frame.rx.w,
                                                                                   frame.rx.h)
    run.close()
    frame.grid_log2 = bins.bit_length() - 1
    frame.tile_width = tile
    frame.occ_waves = waves
    run = fd.FrameRunner(frame, stream=strm)
    run.run()
    lat = []
    for _ in range(steps):
        a = time.perf_counter()
        rc, st = run.run()
        lat.append((time.perf_counter() - a) * 1e3)
    run.close()
    return dict(workload="boxes-1080p with a 32x32 light (1920x1080 receivers, 1M triangles)",
                layout=dict(bins=bins, tile=tile, waves=waves), frames=steps,
                p50_ms=round(pctl(lat, 50), 2), p99_ms=round(pctl(lat, 99), 2), exit=st.exit_code,
                gbps=round(1920 * 1080 * 1024 / (pctl(lat, 50) * 1e-3) / 1e9, 3))


def gpu_autotuned_layout(pkg, frame, steps=20):
    """The headline frame with the autotuner's layout (bins per side x tile
    width; the reference fixes 16x16 bins): frames enqueued back to back,
    as the plain leg's throughput.  Every candidate layout gives the same
    output bits (the autotuner checks them against the default layout)."""
    fd = pkg.fd4d
    strm = fd.Stream()
    run = fd.FrameRunner(frame, stream=strm)
    item = dict(pos=run.pos, nrm=run.nrm, w=frame.rx.w, h=frame.rx.h, light=frame.light, shadow=run.shadow,
                max_pairs=64 << 20)
    cands = pkg.autotune.CANDIDATES
    best, _ = run.ctx.autotune(item, cands, 3)
    run.close()
    bins, tile, waves = cands[best] if best >= 0 else (1 << frame.grid_log2, frame.tile_width, frame.occ_waves)
    f2 = copy.copy(frame)
    f2.grid_log2 = bins.bit_length() - 1
    f2.tile_width = tile
    f2.occ_waves = waves
    run = fd.FrameRunner(f2, stream=strm)
    run.run()
    for _ in range(3):
        run.ctx.launch_async(strm.handle)
    strm.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.ctx.launch_async(strm.handle)
    strm.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    run.close()
    strm.close()
    W, H, R = frame.rx.w, frame.rx.h, frame.light.res_w
    return dict(layout=dict(bins=bins, tile=tile, waves=waves), ms_per_step=round(ms, 4),
                gbps=round(W * H * R * R / (ms * 1e-3) / 1e9, 3))


def gpu_bin_queue(pkg, g, frame, steps=10):
    """The reference's unused bin-queue occlusion kernel (shadowCal_CTA_kernel)
    on the same frame: frame time next to the per-sample kernel's, and how
    far its different arithmetic moves the shadow values."""
    fd = pkg.fd4d
    run = g["run"]
    ctx = run.ctx
    ctx.set_occlusion_variant(fd.OCC_BIN_QUEUE)
    rc, st = run.run()
    if rc != 0:
        raise RuntimeError(f"bin-queue frame failed: rc {rc}")
    lat = []
    for _ in range(steps):
        a = time.perf_counter()
        ctx.launch(run.stream_handle)
        lat.append((time.perf_counter() - a) * 1e3)
    sh, _ = run.outputs()
    ctx.set_occlusion_variant(fd.OCC_PER_SAMPLE)
    d = np.abs(sh.astype(np.float64) - g["shadow"].astype(np.float64))
    return dict(kernel="shadowCal_CTA_kernel (Step5_2_shadowCal.inl:408-1100)", frame_ms=round(pctl(lat, 50), 3),
                per_sample_frame_ms=round(pctl(g["lat"], 50), 3),
                slowdown=round(pctl(lat, 50) / max(pctl(g["lat"], 50), 1e-9), 2),
                shadow_abs_diff=dict(max=float(d.max()), mean=float(d.mean())),
                receivers_differing=int(np.count_nonzero(d)))


def gpu_shard_tile(pkg, steps=10):
    """One GPU's share of a frame sharded by screen tiles over 8 GPUs: a
    480x540 tile (259,200 receivers, 1/8 of 1920x1080) of the boxes-1080p
    G-buffer with its 64x64 light against all 1M triangles, autotuned layout.
    (A real shard would take the light plane from an all-reduce over the
    tiles; here the tile computes its own.)"""
    fd, wl = pkg.fd4d, pkg.workloads
    frame = wl.boxes_1080p()
    rx = frame.rx
    frame.rx = wl.Receivers(np.ascontiguousarray(rx.pos[540:1080, 720:1200]),
                            np.ascontiguousarray(rx.nrm[540:1080, 720:1200]))
    strm = fd.Stream()
    run = fd.FrameRunner(frame, stream=strm)
    item = dict(pos=run.pos, nrm=run.nrm, w=frame.rx.w, h=frame.rx.h, light=frame.light, shadow=run.shadow,
                max_pairs=64 << 20)
    cands = pkg.autotune.CANDIDATES
    best, _ = run.ctx.autotune(item, cands, 3)
    bins, tile, waves = cands[best] if best >= 0 else pkg.autotune.default_layout(frame.light.res_w, frame.rx.w,
                                                                                   frame.rx.h)
    run.close()
    frame.grid_log2 = bins.bit_length() - 1
    frame.tile_width = tile
    frame.occ_waves = waves
    run = fd.FrameRunner(frame, stream=strm)
    run.run()
    lat = []
    for _ in range(steps):
        a = time.perf_counter()
        rc, st = run.run()
        lat.append((time.perf_counter() - a) * 1e3)
    run.close()
    strm.close()
    n = frame.rx.w * frame.rx.h
    R = frame.light.res_w
    return dict(workload="480x540 tile (1/8 of boxes-1080p's 1920x1080) with the 64x64 light, 1M triangles",
                layout=dict(bins=bins, tile=tile, waves=waves), p50_ms=round(pctl(lat, 50), 2), exit=st.exit_code,
                gbps=round(n * R * R / (pctl(lat, 50) * 1e-3) / 1e9, 3))


def gpu_app_pipeline(pkg, frames=30):
    """The application around the visibility path at 1080p: device G-buffer
    producer (20 boxes on a ground plane, the boxes-1080p receivers) ->
    visibility (boxes-1080p occluders, 1M triangles, 32x32 light, 32x32 bins)
    -> deferred shading; each stage timed alone and the whole frame
    end to end.  The device G-buffer is checked bit for bit against the host
    ray cast."""
    fd, wl = pkg.fd4d, pkg.workloads
    W, H = 1920, 1080
    sc = wl.gbuffer_scene_boxes(W, H, 20, 3)
    frame = wl.boxes_1080p(light_res=32)
    frame.grid_log2 = 5
    strm = fd.Stream()
    n = W * H
    g = [fd.DeviceBuffer(n * 16) for _ in range(4)]
    shadow, rgba = fd.DeviceBuffer(n * 4), fd.DeviceBuffer(n * 16)
    masks = fd.DeviceBuffer(n * fd.mask_words(32) * 4)
    ctx = fd.Context()
    ctx.set_backface_cull(frame.backface_cull)
    ctx.set_scenes_esp(frame.eps)
    ctx.set_bin_num(32)
    verts = fd.DeviceBuffer.from_array(frame.scene.verts)
    idx = fd.DeviceBuffer.from_array(frame.scene.idx)
    ctx.add_model(verts, frame.scene.nverts, 16, idx, frame.scene.ntris)
    ctx.set_receivers(g[0], g[1], W, H)
    ctx.set_outputs(shadow, masks)
    ctx.set_area_light(frame.light)
    L = frame.light
    view, up = np.asarray(L.view_dir, np.float64), np.asarray(L.up_dir, np.float64)
    right = np.cross(view / np.linalg.norm(view), up)
    right /= np.linalg.norm(right)
    top = np.cross(right, view / np.linalg.norm(view))
    samples = [np.asarray(L.position) - right * L.right_radius - top * L.top_radius]   # Render.cpp:451-466
    light_terms = (sc.eye, (0.2, 0.2, 0.2), (0.9, 0.85, 0.8), (0.5, 0.5, 0.5))

    def produce():
        fd.gbuffer_boxes(W, H, sc.boxes, sc.materials, sc.eye, sc.at, sc.up, sc.fovy_deg, sc.ground_half,
                         g[0], g[1], g[2], g[3], stream=strm)

    def shade():
        fd.shade(W, H, samples, *light_terms, shadow, g[0], g[1], g[2], g[3], rgba, stream=strm)

    def timed(fn, reps):
        fn()
        strm.sync()
        t = []
        for _ in range(reps):
            a = time.perf_counter()
            fn()
            strm.sync()
            t.append((time.perf_counter() - a) * 1e3)
        return t

    produce()
    strm.sync()
    host = wl.gbuffer_boxes_host(sc)
    exact = all(np.array_equal(g[k].download(np.float32).reshape(H, W, 4).view(np.uint32), host[k].view(np.uint32))
                for k in range(4))
    t_g = timed(produce, 20)
    t_v = timed(lambda: ctx.launch(strm.handle), 10)
    t_s = timed(shade, 20)

    def whole():
        produce()
        rc, _ = ctx.launch(strm.handle)
        assert rc == 0
        shade()
    t_all = timed(whole, frames)
    ctx.close()
    strm.close()
    gb_bytes = 64 * n                      # four float4 targets written
    sh_bytes = (4 * 16 + 4 + 16) * n       # four targets + shadow read, one float4 written
    gms, sms = pctl(t_g, 50), pctl(t_s, 50)
    return dict(workload="1920x1080 boxes G-buffer (20 boxes + ground, device ray cast) -> visibility "
                "(boxes-1080p occluders, 1M triangles, 32x32 light, 32x32 bins) -> deferred shading (1 bound sample)",
                gbuffer_ms=round(gms, 3), visibility_ms=round(pctl(t_v, 50), 3), shade_ms=round(sms, 3),
                frame_p50_ms=round(pctl(t_all, 50), 3), frame_p99_ms=round(pctl(t_all, 99), 3), frames=frames,
                gbuffer_bit_exact_vs_host=bool(exact),
                gbuffer_gigarays_per_s=round(n / (gms * 1e-3) / 1e9, 3),
                gbuffer_hbm=dict(bytes=gb_bytes, gbs=round(gb_bytes / (gms * 1e-3) / 1e9, 1),
                                 frac=round(gb_bytes / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
                shade_hbm=dict(bytes=sh_bytes, gbs=round(sh_bytes / (sms * 1e-3) / 1e9, 1),
                               frac=round(sh_bytes / (sms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
                note="stage times include the host round trip of one call (launch + stream sync)")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_run(oracle, frame, gpu, seconds, nthreads, rng):
    """One timed oracle run over a receiver sample sized to about `seconds`
    of occlusion work; returns the full-frame estimate."""
    W, H, R = frame.rx.w, frame.rx.h, frame.light.res_w
    n = 2048 if nthreads == 1 else 8192
    o = oracle.run(frame, subset=rng.choice(W * H, n, replace=False), nthreads=nthreads)
    per_rx = o.scalars["ms_occ"] / max(1, int(o.computed.sum()))
    n2 = int(min(W * H, max(n, seconds * 1e3 / max(per_rx, 1e-6))))
    sub = rng.choice(W * H, n2, replace=False)
    t0 = time.perf_counter()
    o = oracle.run(frame, subset=sub, nthreads=nthreads)
    wall = time.perf_counter() - t0
    s = o.scalars
    computed = int(o.computed.sum())
    fixed_ms = s["ms_lp"] + s["ms_setup"] + s["ms_bin"] + s["ms_pair"]
    binned = int((o.key < int(s["num_bins"])).sum())
    occ_full_ms = s["ms_occ"] * binned / max(1, computed)
    frame_ms = fixed_ms + occ_full_ms
    ids = sub[o.computed[sub] > 0]
    wpr = 1 if R <= 32 else 2
    match = bool(np.array_equal(gpu["masks"].reshape(W * H, R * wpr)[ids], o.masks.reshape(W * H, R * wpr)[ids])
                 and np.array_equal(gpu["shadow"][ids].view(np.uint32), o.shadow[ids].view(np.uint32)))
    return dict(value=round(W * H * R * R / (frame_ms * 1e-3) / 1e9, 6), frame_ms=round(frame_ms, 2),
                sample=(f"oracle on {W}x{H} receivers / {frame.scene.ntris} triangles, {nthreads} thread(s): all "
                        f"stages in full ({fixed_ms:.0f} ms), occlusion of {computed} sampled receivers "
                        f"({wall:.1f} s wall), scaled to the {binned} binned receivers ({occ_full_ms:.0f} ms)"),
                masks_match_gpu_on_sample=match, sample_receivers=computed)


def cpu_baseline(frame, gpu, seconds):
    """The oracle (CPU restatement of the reference, compiled -O3
    -ffp-contract=off) on this host: one thread and all cores, each over a
    receiver sample; full-frame time = fixed stages + occlusion scaled."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nthreads = oracle.max_threads()
    rng = np.random.default_rng(123)
    multi = _cpu_run(oracle, frame, gpu, seconds, nthreads, rng)
    single = _cpu_run(oracle, frame, gpu, seconds, 1, rng)
    return dict(value=multi["value"], unit="Gbit/s", cores=nthreads, kind="port", sample=multi["sample"],
                cpu_model=cpu_model(), frame_ms=multi["frame_ms"],
                masks_match_gpu_on_sample=multi["masks_match_gpu_on_sample"] and single["masks_match_gpu_on_sample"],
                sample_receivers=multi["sample_receivers"],
                single_thread=dict(value=single["value"], unit="Gbit/s", cores=1, frame_ms=single["frame_ms"],
                                   sample=single["sample"]),
                gpu_vs_cpu=dict(all_cores=round(gpu["gbps"] / max(multi["value"], 1e-12), 1),
                                one_core=round(gpu["gbps"] / max(single["value"], 1e-12), 1)))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="boxes-1080p", choices=["heightfield-1k", "plane-256", "boxes-1080p"])
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="occlusion seconds per CPU baseline run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="timed frames only (for profiler runs)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--light-res", type=int, default=0, help="override the workload's light grid side")
    ap.add_argument("--bins", type=int, default=0, help="bins per light-plane side (default: the workload's, 16)")
    ap.add_argument("--tile", type=int, default=0, help="receiver tile width (default 8)")
    ap.add_argument("--waves", type=int, default=0, choices=[0, 2, 4, 8],
                    help="waves per occlusion work item (default: 8 above 32x32 light samples and for frames of "
                         "up to 8192 work items, else 4)")
    ap.add_argument("--leg", choices=["batched", "served", "target", "app_pipeline", "shard_tile"], default=None,
                    help="run only this leg and print its JSON (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (instances shard across GPUs "
              "without exchange; see the sharding design).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    pkg = importlib.import_module("4dzb_amd")
    if pkg.fd4d.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    wl = pkg.workloads
    if args.leg:
        legs = {"batched": lambda: gpu_batched(pkg), "served": lambda: gpu_served(pkg),
                "target": lambda: gpu_target_config(pkg, steps=args.steps),
                "app_pipeline": lambda: gpu_app_pipeline(pkg, frames=args.steps),
                "shard_tile": lambda: gpu_shard_tile(pkg, steps=args.steps)}
        print(json.dumps({args.leg: legs[args.leg]()}), flush=True)
        return
    log(f"[bench] generating {args.workload}")
    frame = make_frame(wl, args.workload, args.light_res)
    if args.bins:
        frame.grid_log2 = args.bins.bit_length() - 1
    if args.tile:
        frame.tile_width = args.tile
    if args.waves:
        frame.occ_waves = args.waves
    log("[bench] GPU plain path")
    g, leg = plain_leg(pkg, frame, args.workload, args.steps, args.warmup, args.quick)
    out = dict(metric=METRIC, value=leg["gbps"], unit="Gbit/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=leg["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f32", data="synthetic (seeded generators, 4dzb_amd/csrc/scenes.cpp)",
               config=leg["config"], p50_ms=leg["p50_ms"], p99_ms=leg["p99_ms"], stage_ms=leg["stage_ms"],
               frame_stats=leg["frame_stats"], roofline=leg["roofline"])
    if not args.quick and args.workload != "heightfield-1k" and not args.light_res and not args.bins \
            and not args.tile and not args.waves:
        log("[bench] GPU plain path, heightfield-1k")
        f2 = make_frame(wl, "heightfield-1k", 0)
        g2, leg2 = plain_leg(pkg, f2, "heightfield-1k", args.steps, args.warmup, False)
        leg2["masks_match_oracle_on_sample"] = oracle_sample_check(f2, g2, 3000)
        out["plain_heightfield_1k"] = leg2
        log("[bench] bin-queue occlusion variant (heightfield-1k)")
        out["bin_queue_variant"] = dict(gpu_bin_queue(pkg, g2, f2), workload="heightfield-1k")
        g2["run"].close()
    if not args.quick and not args.bins and not args.tile and not args.waves:
        log("[bench] autotuned layout (same frame)")
        out["autotuned_layout"] = gpu_autotuned_layout(pkg, frame)
    if not args.quick and args.workload == "heightfield-1k":
        log("[bench] bin-queue occlusion variant (same frame)")
        out["bin_queue_variant"] = gpu_bin_queue(pkg, g, frame)
    if not args.no_batch and not args.quick:
        log("[bench] batched form (batch-64)")
        out["batched"] = gpu_batched(pkg)
    if not args.no_serve and not args.quick:
        log("[bench] serving form (serve-poisson)")
        out["served"] = gpu_served(pkg)
        log("[bench] latency target configuration")
        out["target_1080p_32x32_1M"] = gpu_target_config(pkg, steps=100)
        log("[bench] application pipeline (G-buffer -> visibility -> shading)")
        out["app_pipeline"] = gpu_app_pipeline(pkg)
        log("[bench] one GPU's tile of an 8-way screen split")
        out["shard_tile_1_of_8"] = gpu_shard_tile(pkg)
    if not args.no_cpu and not args.quick:
        log("[bench] CPU baseline (oracle)")
        g["gbps"] = out["value"]
        out["cpu_baseline"] = cpu_baseline(frame, g, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
