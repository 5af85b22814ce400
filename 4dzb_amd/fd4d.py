"""ctypes binding of libfd4d.so (include/fd4d.h).

Device buffers are raw device pointers.  `DeviceBuffer` and `Stream` wrap
the library's own allocation / stream helpers, so a process that drives the
library needs no other GPU runtime.  The library must be present: there is
no fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfd4d.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "fd4d.h")

FD_OK = 0
FD_ERR_ARG = -1
FD_ERR_STATE = -2
FD_ERR_HIP = -3
FD_ERR_CAPACITY = -4
FD_ERR_UNSUPPORTED_LIGHT = -5
FD_ERR_NO_DEVICE = -6
FD_ERR_SCRATCH = -7

EXIT_OK, EXIT_SAMPLE_MAXZ, EXIT_TRI_MAXZ, EXIT_NO_VALID_SAMPLE = 0, 1, 2, 3
EXIT_EMPTY_AABB, EXIT_EMPTY_LP, EXIT_NO_VALID_TRI, EXIT_NO_PAIRS, EXIT_CAPACITY = 4, 5, 6, 7, 9

OCC_PER_SAMPLE, OCC_BIN_QUEUE = 0, 1   # fd_set_occlusion_variant
TRI_ORDER_AUTO, TRI_ORDER_LIST, TRI_ORDER_AREA, TRI_ORDER_MESH = 0, 1, 2, 3   # fd_set_triangle_order


class AreaLight(ctypes.Structure):
    _fields_ = [("right_radius", ctypes.c_float), ("top_radius", ctypes.c_float),
                ("position", ctypes.c_float * 3), ("view_dir", ctypes.c_float * 3), ("up_dir", ctypes.c_float * 3),
                ("res_w", ctypes.c_int), ("res_h", ctypes.c_int)]

    @classmethod
    def from_desc(cls, d):
        return cls(d.right_radius, d.top_radius, (ctypes.c_float * 3)(*d.position),
                   (ctypes.c_float * 3)(*d.view_dir), (ctypes.c_float * 3)(*d.up_dir), d.res_w, d.res_h)


class Stats(ctypes.Structure):
    _fields_ = [("exit_code", ctypes.c_int), ("valid_receivers", ctypes.c_int), ("front_receivers", ctypes.c_int),
                ("valid_triangles", ctypes.c_int), ("pairs", ctypes.c_ulonglong), ("key_out_of_range", ctypes.c_int),
                ("work_items", ctypes.c_int), ("elapsed_ms", ctypes.c_float),
                ("light_plane_begin", ctypes.c_float * 2), ("light_plane_end", ctypes.c_float * 2),
                ("far_plane", ctypes.c_float)]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if hasattr(getattr(self, k), "__len__") else getattr(self, k))
                for k, _ in self._fields_}


class Instance(ctypes.Structure):
    _fields_ = [("d_position4", ctypes.c_void_p), ("d_normal4", ctypes.c_void_p), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("light", AreaLight), ("d_shadow", ctypes.c_void_p),
                ("d_masks", ctypes.c_void_p), ("bins_per_side", ctypes.c_int), ("tile_width", ctypes.c_int),
                ("max_pairs", ctypes.c_ulonglong), ("occ_waves", ctypes.c_int), ("tri_order", ctypes.c_int),
                ("list_split", ctypes.c_int), ("pad", ctypes.c_int)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("light_res", ctypes.c_int), ("bins_per_side", ctypes.c_int), ("tile_width", ctypes.c_int),
                ("ms", ctypes.c_float), ("occ_waves", ctypes.c_int), ("tri_order", ctypes.c_int),
                ("list_split", ctypes.c_int)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("streams", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("bins_per_side", ctypes.c_int), ("backface_cull", ctypes.c_int), ("eps", ctypes.c_float),
                ("tile_width", ctypes.c_int), ("max_pairs", ctypes.c_ulonglong), ("tune", ctypes.c_void_p),
                ("ntune", ctypes.c_int), ("pad", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int), ("width", ctypes.c_int), ("height", ctypes.c_int), ("pad", ctypes.c_int),
                ("d_position4", ctypes.c_void_p), ("d_normal4", ctypes.c_void_p), ("light", AreaLight),
                ("d_shadow", ctypes.c_void_p), ("d_masks", ctypes.c_void_p), ("user_tag", ctypes.c_ulonglong)]


class Material(ctypes.Structure):
    _fields_ = [("ka", ctypes.c_float * 3), ("kd", ctypes.c_float * 3), ("ks", ctypes.c_float * 3),
                ("shininess", ctypes.c_float)]


class GbufferScene(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("nbox", ctypes.c_int), ("pad", ctypes.c_int),
                ("boxes", ctypes.c_void_p), ("materials", ctypes.c_void_p), ("eye", ctypes.c_double * 3),
                ("at", ctypes.c_double * 3), ("up", ctypes.c_double * 3), ("fovy_deg", ctypes.c_double),
                ("ground_half", ctypes.c_double)]


class ShadeParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("nsamples", ctypes.c_int), ("pad", ctypes.c_int),
                ("samples", ctypes.c_void_p), ("camera", ctypes.c_float * 3), ("La", ctypes.c_float * 3),
                ("Ld", ctypes.c_float * 3), ("Ls", ctypes.c_float * 3)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("user_tag", ctypes.c_ulonglong), ("status", ctypes.c_int),
                ("exit_code", ctypes.c_int), ("batch_size", ctypes.c_int), ("pad", ctypes.c_int),
                ("pairs", ctypes.c_ulonglong), ("t_submit_ms", ctypes.c_double), ("t_start_ms", ctypes.c_double),
                ("t_done_ms", ctypes.c_double)]


_lib = None


def lib():
    """Load libfd4d.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        sig = {
            "fd_init": ([ctypes.POINTER(vp)], ip),
            "fd_destroy": ([vp], ip),
            "fd_set_backface_cull": ([vp, ip], ip),
            "fd_add_model": ([vp, vp, sz, ip, vp, ip], ip),
            "fd_clear_models": ([vp], ip),
            "fd_set_sample_buffers": ([vp, vp, vp], ip),
            "fd_set_output_buffers": ([vp, vp, vp], ip),
            "fd_register_res": ([vp], ip),
            "fd_unregister_res": ([vp], ip),
            "fd_set_bin_num": ([vp, sz], ip),
            "fd_set_tile_width": ([vp, ip], ip),
            "fd_set_occlusion_variant": ([vp, ip], ip),
            "fd_set_triangle_order": ([vp, ip], ip),
            "fd_set_occlusion_waves": ([vp, ip], ip),
            "fd_set_list_split": ([vp, ip], ip),
            "fd_autotune": ([vp, ctypes.c_void_p, ctypes.c_void_p, ip, ip, ctypes.POINTER(ip)], ip),
            "fd_set_view": ([vp, ip, ip], ip),
            "fd_set_scenes_esp": ([vp, ctypes.c_float], ip),
            "fd_set_scene_params": ([vp, sz, sz], ip),
            "fd_set_area_light": ([vp, ctypes.POINTER(AreaLight)], ip),
            "fd_launch": ([vp, vp, ctypes.POINTER(Stats)], ip),
            "fd_launch_async": ([vp, vp], ip),
            "fd_launch_async_noalloc": ([vp, vp], ip),
            "fd_get_stats": ([vp, vp, ctypes.POINTER(Stats)], ip),
            "fd_set_pair_capacity": ([vp, sz], ip),
            "fd_move_model": ([vp, ctypes.POINTER(ctypes.c_float), vp], ip),
            "fd_mask_words": ([ip], ip),
            "fd_gbuffer_boxes": ([ctypes.POINTER(GbufferScene), vp, vp, vp, vp, vp], ip),
            "fd_shade": ([ctypes.POINTER(ShadeParams), vp, vp, vp, vp, vp, vp, vp], ip),
            "fd_debug_read": ([vp, ctypes.c_char_p, vp, sz, ctypes.POINTER(sz)], ip),
            "fd_status_string": ([ip], ctypes.c_char_p),
            "fd_set_profiling": ([vp, ip, ip], ip),
            "fd_selftest_rowmask": ([vp, vp, ip, ip, ip, vp], ip),
            "fd_selftest_rowwindow": ([vp, ip, ip, vp], ip),
            "fd_device_count": ([ctypes.POINTER(ip)], ip),
            "fd_batch_create": ([vp, ctypes.c_void_p, ip, ctypes.POINTER(vp)], ip),
            "fd_batch_capture": ([vp, vp], ip),
            "fd_batch_launch": ([vp, vp], ip),
            "fd_batch_get_stats": ([vp, ip, ctypes.POINTER(Stats)], ip),
            "fd_batch_destroy": ([vp], ip),
            "fd_server_create": ([ctypes.c_void_p, ctypes.POINTER(vp)], ip),
            "fd_server_add_scene": ([vp, vp, sz, ip, vp, ip], ip),
            "fd_server_submit": ([vp, ctypes.c_void_p], ctypes.c_longlong),
            "fd_server_poll": ([vp, ctypes.c_void_p, ip], ip),
            "fd_server_move_scene": ([vp, ip, ctypes.POINTER(ctypes.c_float), ctypes.c_ulonglong], ctypes.c_longlong),
            "fd_server_drain": ([vp, ip], ip),
            "fd_server_now_ms": ([vp], ctypes.c_double),
            "fd_server_destroy": ([vp], ip),
            "fd_dev_alloc": ([sz, ctypes.POINTER(vp)], ip),
            "fd_dev_free": ([vp], ip),
            "fd_memcpy_htod": ([vp, vp, sz], ip),
            "fd_memcpy_dtoh": ([vp, vp, sz], ip),
            "fd_memcpy_dtod": ([vp, vp, sz, vp], ip),
            "fd_stream_create": ([ctypes.POINTER(vp)], ip),
            "fd_stream_destroy": ([vp], ip),
            "fd_stream_sync": ([vp], ip),
            "fd_device_sync": ([], ip),
            "fd_get_profile": ([vp, ctypes.POINTER(ctypes.c_float), ip, ctypes.POINTER(ctypes.c_ulonglong)], ip),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def mask_words(res):
    return lib().fd_mask_words(res)


def selftest_rowmask(D, B, res, mode=1):
    """Device row-mask primitive on host arrays (verification only); mode 0
    is the general form, mode 1 the fast form where it applies."""
    D = np.ascontiguousarray(D, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    out = np.empty(len(D), np.uint64)
    _check(lib().fd_selftest_rowmask(D.ctypes.data, B.ctypes.data, len(D), res, mode, out.ctypes.data),
           "fd_selftest_rowmask")
    return out


def selftest_rowwindow(coef, res):
    """Device row window on host edge-plane triples (n, 9) float32
    (verification only).  Returns (n, 4) uint64: non-empty rows by the
    general form, by the fast form, the window's rows and the row-range
    mask."""
    coef = np.ascontiguousarray(coef, np.float32).reshape(-1, 9)
    out = np.empty((coef.shape[0], 4), np.uint64)
    _check(lib().fd_selftest_rowwindow(coef.ctypes.data, coef.shape[0], res, out.ctypes.data),
           "fd_selftest_rowwindow")
    return out


class FdError(RuntimeError):
    pass


def _check(rc, what):
    if rc != FD_OK:
        raise FdError(f"{what}: {lib().fd_status_string(rc).decode()} ({rc})")


def _ptr(t):
    return ctypes.c_void_p(t.ptr if t is not None else 0)


def device_count():
    n = ctypes.c_int()
    lib().fd_device_count(ctypes.byref(n))
    return n.value


class DeviceBuffer:
    """Device allocation owned by Python (freed on close / garbage collection)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().fd_dev_alloc(int(nbytes), ctypes.byref(p)), "fd_dev_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().fd_memcpy_htod(ctypes.c_void_p(self.ptr), ctypes.c_void_p(a.ctypes.data), a.nbytes),
               "fd_memcpy_htod")

    def download(self, dtype, count=None):
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dt)
        _check(lib().fd_memcpy_dtoh(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr), out.nbytes),
               "fd_memcpy_dtoh")
        return out

    def close(self):
        if self.ptr:
            lib().fd_dev_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().fd_stream_create(ctypes.byref(h)), "fd_stream_create")
        self.handle = h.value

    def sync(self):
        _check(lib().fd_stream_sync(ctypes.c_void_p(self.handle)), "fd_stream_sync")

    def close(self):
        if self.handle:
            lib().fd_stream_destroy(ctypes.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One fd_ctx.  Mirrors the FD:: call sequence of the reference."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().fd_init(ctypes.byref(h)), "fd_init")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib().fd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_backface_cull(self, on):
        _check(lib().fd_set_backface_cull(self.h, int(on)), "fd_set_backface_cull")

    def set_scenes_esp(self, eps):
        _check(lib().fd_set_scenes_esp(self.h, eps), "fd_set_scenes_esp")

    def set_bin_num(self, n):
        rc = lib().fd_set_bin_num(self.h, n)
        if rc != FD_OK:
            raise FdError(f"fd_set_bin_num({n}) -> {rc}")

    def set_tile_width(self, tw):
        _check(lib().fd_set_tile_width(self.h, tw), f"fd_set_tile_width({tw})")

    def set_occlusion_variant(self, variant):
        """OCC_PER_SAMPLE (the reference's FDLaunch kernel, bit-exact contract)
        or OCC_BIN_QUEUE (its unused shadowCal_CTA_kernel; other arithmetic,
        mask rows along y_delt)."""
        _check(lib().fd_set_occlusion_variant(self.h, int(variant)), f"fd_set_occlusion_variant({variant})")

    def set_occlusion_waves(self, waves):
        """Waves per occlusion work item: 2, 4 or 8; 0 = the default (8 for light
        grids above 32x32 and for frames of up to 8192 work items, else 4).
        Changes no result."""
        _check(lib().fd_set_occlusion_waves(self.h, int(waves)), f"fd_set_occlusion_waves({waves})")

    def set_list_split(self, segs):
        """Stretches per bin triangle list: 0 = the automatic rule, 1..64 that
        many (changes no result)."""
        _check(lib().fd_set_list_split(self.h, int(segs)), f"fd_set_list_split({segs})")

    def set_triangle_order(self, order):
        """TRI_ORDER_AUTO / _LIST / _AREA / _MESH: the order in which the
        occlusion walks each bin's triangle list (changes no result)."""
        _check(lib().fd_set_triangle_order(self.h, int(order)), f"fd_set_triangle_order({order})")

    def autotune(self, inst_item, candidates, reps=3):
        """Autotuner (fd_autotune): inst_item as a Batch item (its outputs are
        not written); candidates: list of (bins, tile_width), (bins,
        tile_width, waves per occlusion work item), (bins, tile_width,
        waves, walk order: TRI_ORDER_*) or (bins, tile_width, waves, order,
        list split).  Returns (best index or
        -1, [ms per candidate: -1 = not bit-exact, -2 = its pairs do not fit
        the instance's pair arena])."""
        it = inst_item
        inst = Instance(it["pos"].ptr, it["nrm"].ptr, it["w"], it["h"], AreaLight.from_desc(it["light"]),
                        it["shadow"].ptr, 0, 0, 0, it.get("max_pairs", 0), 0, 0, 0, 0)
        arr = (TuneEntry * len(candidates))()
        for k, c in enumerate(candidates):
            b, t = c[0], c[1]
            w = c[2] if len(c) > 2 else 0
            o = c[3] if len(c) > 3 else 0
            sp = c[4] if len(c) > 4 else 0
            arr[k] = TuneEntry(-1, it["w"], it["h"], it["light"].res_w, b, t, 0.0, w, o, sp)
        best = ctypes.c_int(-1)
        _check(lib().fd_autotune(self.h, ctypes.byref(inst), ctypes.cast(arr, ctypes.c_void_p), len(candidates),
                                 reps, ctypes.byref(best)), "fd_autotune")
        return best.value, [arr[k].ms for k in range(len(candidates))]

    def set_pair_capacity(self, n):
        _check(lib().fd_set_pair_capacity(self.h, n), "fd_set_pair_capacity")

    def add_model(self, verts_dev, nverts, stride_bytes, idx_dev, ntris):
        """verts_dev / idx_dev: DeviceBuffer of float vertices / int3 indices."""
        lib().fd_clear_models(self.h)
        _check(lib().fd_add_model(self.h, _ptr(verts_dev), stride_bytes, nverts, _ptr(idx_dev), ntris),
               "fd_add_model")
        self._keep += [verts_dev, idx_dev]

    def set_receivers(self, pos_dev, nrm_dev, w, h):
        _check(lib().fd_set_view(self.h, w, h), "fd_set_view")
        _check(lib().fd_set_sample_buffers(self.h, _ptr(pos_dev), _ptr(nrm_dev)), "fd_set_sample_buffers")
        self._keep += [pos_dev, nrm_dev]

    def set_outputs(self, shadow_dev, masks_dev=None):
        _check(lib().fd_set_output_buffers(self.h, _ptr(shadow_dev), _ptr(masks_dev)), "fd_set_output_buffers")
        self._keep += [shadow_dev, masks_dev]

    def set_area_light(self, desc):
        L = AreaLight.from_desc(desc)
        _check(lib().fd_set_area_light(self.h, ctypes.byref(L)), "fd_set_area_light")

    def launch(self, stream=None):
        st = Stats()
        rc = lib().fd_launch(self.h, ctypes.c_void_p(stream or 0), ctypes.byref(st))
        return rc, st

    def launch_async(self, stream=None):
        _check(lib().fd_launch_async(self.h, ctypes.c_void_p(stream or 0)), "fd_launch_async")

    def get_stats(self, stream=None):
        st = Stats()
        rc = lib().fd_get_stats(self.h, ctypes.c_void_p(stream or 0), ctypes.byref(st))
        return rc, st

    def move_model(self, m16, stream=None):
        arr = (ctypes.c_float * 16)(*[float(x) for x in m16])
        _check(lib().fd_move_model(self.h, arr, ctypes.c_void_p(stream or 0)), "fd_move_model")

    def set_profiling(self, stage_events=True, count_ops=False):
        _check(lib().fd_set_profiling(self.h, int(stage_events), int(count_ops)), "fd_set_profiling")

    def get_profile(self, ops=False):
        ms = (ctypes.c_float * 5)()
        cnt = (ctypes.c_ulonglong * 8)()
        _check(lib().fd_get_profile(self.h, ms, 5, cnt if ops else None), "fd_get_profile")
        return list(ms), (list(cnt) if ops else None)

    def debug_read(self, name, dtype):
        n = ctypes.c_size_t()
        _check(lib().fd_debug_read(self.h, name.encode(), None, 0, ctypes.byref(n)), f"fd_debug_read({name})")
        out = np.empty(n.value // np.dtype(dtype).itemsize, dtype)
        if n.value:
            _check(lib().fd_debug_read(self.h, name.encode(), ctypes.c_void_p(out.ctypes.data), n.value,
                                       ctypes.byref(n)), f"fd_debug_read({name})")
        return out


class FrameRunner:
    """Uploads a workloads.Frame to the GPU and runs it through one Context."""

    def __init__(self, frame, with_masks=True, ctx=None, pair_capacity=None, stream=None):
        self.frame = frame
        sc, rx = frame.scene, frame.rx
        self.verts = DeviceBuffer.from_array(np.ascontiguousarray(sc.verts, np.float32))
        self.idx = DeviceBuffer.from_array(np.ascontiguousarray(sc.idx, np.int32))
        self.pos = DeviceBuffer.from_array(np.ascontiguousarray(rx.pos, np.float32))
        self.nrm = DeviceBuffer.from_array(np.ascontiguousarray(rx.nrm, np.float32))
        W, H, R = rx.w, rx.h, frame.light.res_w
        self.shadow = DeviceBuffer(H * W * 4)
        self.masks = None
        if with_masks and 1 <= R <= 64 and frame.light.res_h == R:
            self.masks = DeviceBuffer(H * W * mask_words(R) * 4)
        self.stream = stream
        self.ctx = ctx or Context()
        c = self.ctx
        c.set_backface_cull(frame.backface_cull)
        c.set_scenes_esp(frame.eps)
        c.set_bin_num(1 << frame.grid_log2)
        c.set_tile_width(frame.tile_width)
        c.set_occlusion_waves(getattr(frame, "occ_waves", 0))
        c.set_list_split(getattr(frame, "list_split", 0))
        if pair_capacity:
            c.set_pair_capacity(pair_capacity)
        c.add_model(self.verts, sc.nverts, sc.verts.shape[1] * 4, self.idx, sc.ntris)
        c.set_receivers(self.pos, self.nrm, W, H)
        c.set_outputs(self.shadow, self.masks)
        self.light_rc = lib().fd_set_area_light(c.h, ctypes.byref(AreaLight.from_desc(frame.light)))

    @property
    def stream_handle(self):
        return self.stream.handle if self.stream is not None else None

    def run(self):
        return self.ctx.launch(self.stream_handle)

    def outputs(self):
        sh = self.shadow.download(np.float32)
        mk = self.masks.download(np.uint32) if self.masks is not None else None
        return sh, mk

    def close(self):
        self.ctx.close()
        for b in (self.verts, self.idx, self.pos, self.nrm, self.shadow, self.masks):
            if b is not None:
                b.close()


class Batch:
    """Batched form: many (receivers, light, outputs) instances over one
    context's scene, one launch per stage; optional HIP-graph capture."""

    def __init__(self, ctx, items):
        """items: list of dicts with keys pos, nrm (DeviceBuffer), w, h, light
        (LightDesc), shadow, masks (DeviceBuffer or None), bins, tile,
        waves, order and split (0 = the context's setting), max_pairs."""
        arr = (Instance * len(items))()
        for k, it in enumerate(items):
            arr[k] = Instance(it["pos"].ptr, it["nrm"].ptr, it["w"], it["h"], AreaLight.from_desc(it["light"]),
                              it["shadow"].ptr, it["masks"].ptr if it.get("masks") is not None else 0,
                              it.get("bins", 0), it.get("tile", 0), it.get("max_pairs", 0), it.get("waves", 0),
                              it.get("order", 0), it.get("split", 0), 0)
        h = ctypes.c_void_p()
        _check(lib().fd_batch_create(ctx.h, ctypes.cast(arr, ctypes.c_void_p), len(items), ctypes.byref(h)),
               "fd_batch_create")
        self.h = h
        self.n = len(items)
        self._keep = (ctx, items)

    def capture(self, stream):
        _check(lib().fd_batch_capture(self.h, ctypes.c_void_p(stream)), "fd_batch_capture")

    def launch(self, stream=None):
        _check(lib().fd_batch_launch(self.h, ctypes.c_void_p(stream or 0)), "fd_batch_launch")

    def stats(self, i):
        st = Stats()
        rc = lib().fd_batch_get_stats(self.h, i, ctypes.byref(st))
        return rc, st

    def close(self):
        if self.h:
            lib().fd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving form: resident scenes, continuous batching over streams."""

    def __init__(self, budget_bytes, streams=4, max_batch=16, bins=16, cull=1, eps=1e-3, max_pairs=0, tile=8,
                 tune=None):
        """tune: optional list of dicts (scene, w, h, res, bins, tile, and
        optionally waves, order and split) -- the per-class layouts picked by the
        autotuner."""
        tune = tune or []
        self._tune = (TuneEntry * max(1, len(tune)))()
        for k, t in enumerate(tune):
            self._tune[k] = TuneEntry(t.get("scene", -1), t["w"], t["h"], t["res"], t["bins"], t["tile"],
                                      float(t.get("ms", 0.0)), t.get("waves", 0), t.get("order", 0),
                                      t.get("split", 0))
        cfg = ServerConfig(budget_bytes, streams, max_batch, bins, cull, eps, tile, max_pairs,
                           ctypes.cast(self._tune, ctypes.c_void_p).value if tune else None, len(tune), 0)
        h = ctypes.c_void_p()
        _check(lib().fd_server_create(ctypes.cast(ctypes.pointer(cfg), ctypes.c_void_p), ctypes.byref(h)),
               "fd_server_create")
        self.h = h
        self._keep = []

    def add_scene(self, verts_dev, nverts, stride_bytes, idx_dev, ntris):
        sid = lib().fd_server_add_scene(self.h, _ptr(verts_dev), stride_bytes, nverts, _ptr(idx_dev), ntris)
        if sid < 0:
            raise FdError(f"fd_server_add_scene -> {sid}")
        self._keep += [verts_dev, idx_dev]
        return sid

    def submit(self, scene, pos, nrm, w, h, light, shadow, masks=None, tag=0):
        r = Request(scene, w, h, 0, pos.ptr, nrm.ptr, AreaLight.from_desc(light), shadow.ptr,
                    masks.ptr if masks is not None else 0, tag)
        rid = lib().fd_server_submit(self.h, ctypes.cast(ctypes.pointer(r), ctypes.c_void_p))
        if rid < 0:
            raise FdError(f"fd_server_submit -> {rid}")
        return rid

    def move_scene(self, scene, m16, tag=0):
        """fd_server_move_scene: m16 column-major 4x4, ordered with the requests."""
        m = (ctypes.c_float * 16)(*[float(x) for x in m16])
        rid = lib().fd_server_move_scene(self.h, scene, m, tag)
        if rid < 0:
            raise FdError(f"fd_server_move_scene -> {rid}")
        return rid

    def poll(self, max_n=256):
        arr = (Completion * max_n)()
        n = lib().fd_server_poll(self.h, ctypes.cast(arr, ctypes.c_void_p), max_n)
        return [arr[i] for i in range(n)]

    def drain(self, timeout_ms=60000):
        _check(lib().fd_server_drain(self.h, timeout_ms), "fd_server_drain")

    def now_ms(self):
        return lib().fd_server_now_ms(self.h)

    def close(self):
        if self.h:
            lib().fd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# numpy mirrors of fd_types.h records (for stage read-back)
STATE_DTYPE = np.dtype([
    ("rxZmaxAll", "<u4"), ("rxMin", "<u4", 3), ("rxMaxA", "<u4", 2), ("rxMaxZ", "<u4"),
    ("vxZmax", "<u4"), ("vxMin", "<u4", 3), ("vxMaxA", "<u4", 2), ("vxMaxZ", "<u4"),
    ("validRect", "<i4"), ("status", "<i4"), ("validSamples", "<i4"), ("validTris", "<i4"),
    ("keyOutOfRange", "<i4"), ("capacityHit", "<i4"), ("pairs", "<u8"), ("rxTotal", "<u8"),
    ("totalChunks", "<i4"), ("triOrder", "<i4"), ("sampleMaxZ", "<f4"), ("triMaxZ", "<f4"), ("f", "<f4"),
    ("S", "<f4", 2), ("sAABB", "<f4", 8), ("mAABB", "<f4", 8), ("B", "<f4", 3), ("E", "<f4", 3),
    ("factor", "<f4", 2), ("G", "<i4"), ("log2G", "<i4"), ("nb", "<i4"), ("pad1", "<i4")], align=True)

LIGHT_DTYPE = np.dtype([("m", "<f4", 16), ("xd", "<f4", 3), ("yd", "<f4", 3), ("ur", "<f4", 3), ("ul", "<f4", 3),
                        ("dl", "<f4", 3), ("dr", "<f4", 3), ("F", "<f4", 2), ("res", "<i4"), ("resNum", "<i4"),
                        ("wpr", "<i4"), ("pad", "<i4")])

TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("v1", "<f4", 3), ("v2", "<f4", 3), ("bminx", "<f4"), ("bminy", "<f4"),
                      ("bmaxx", "<f4"), ("bmaxy", "<f4"), ("minz", "<f4"), ("maxz", "<f4"), ("pad0", "<f4"),
                      ("n", "<f4", 3), ("e01", "<f4", 3), ("e12", "<f4", 3), ("e20", "<f4", 3), ("pad1", "<f4", 4)])


def gbuffer_boxes(width, height, boxes, materials, eye, at, up, fovy_deg, ground_half, g0, g1, g2=None, g3=None,
                  stream=None):
    """fd_gbuffer_boxes: boxes (nbox, 6) float64, materials (nbox + 1, 10)
    float32 (ka, kd, ks, shininess) or None; g0..g3 DeviceBuffers of
    width*height float4 (g2, g3 optional)."""
    import numpy as np
    bx = np.ascontiguousarray(np.asarray(boxes, np.float64).reshape(-1, 6))
    mt = None if materials is None else np.ascontiguousarray(np.asarray(materials, np.float32).reshape(-1, 10))
    sc = GbufferScene(width, height, bx.shape[0], 0, bx.ctypes.data if bx.size else None,
                      mt.ctypes.data if mt is not None else None, (ctypes.c_double * 3)(*eye),
                      (ctypes.c_double * 3)(*at), (ctypes.c_double * 3)(*up), fovy_deg, ground_half)
    h = stream.handle if isinstance(stream, Stream) else stream
    _check(lib().fd_gbuffer_boxes(ctypes.byref(sc), _ptr(g0), _ptr(g1), _ptr(g2) if g2 is not None else None,
                                  _ptr(g3) if g3 is not None else None, h), "fd_gbuffer_boxes")


def shade(width, height, samples, camera, La, Ld, Ls, shadow, g0, g1, g2, g3, rgba, stream=None):
    """fd_shade: samples (ns, 3) world positions; shadow / g0..g3 / rgba
    DeviceBuffers (float x W*H, float4 x W*H)."""
    import numpy as np
    sp = np.ascontiguousarray(np.asarray(samples, np.float32).reshape(-1, 3))
    f3 = lambda v: (ctypes.c_float * 3)(*[float(x) for x in v])
    p = ShadeParams(width, height, sp.shape[0], 0, sp.ctypes.data, f3(camera), f3(La), f3(Ld), f3(Ls))
    h = stream.handle if isinstance(stream, Stream) else stream
    _check(lib().fd_shade(ctypes.byref(p), _ptr(shadow), _ptr(g0), _ptr(g1), _ptr(g2), _ptr(g3), _ptr(rgba), h),
           "fd_shade")


def read_state(ctx):
    raw = ctx.debug_read("state", np.uint8)
    return np.frombuffer(raw.tobytes()[:STATE_DTYPE.itemsize], STATE_DTYPE)[0]


def read_light(ctx):
    raw = ctx.debug_read("light", np.uint8)
    return np.frombuffer(raw.tobytes()[:LIGHT_DTYPE.itemsize], LIGHT_DTYPE)[0]


def read_tris(ctx):
    raw = ctx.debug_read("tri_data", np.uint8)
    return np.frombuffer(raw.tobytes(), TRI_DTYPE)
