"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or as the timed CPU baseline.  The
product path (the HIP library) never calls it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")

EXIT_NAMES = {0: "ok", 1: "sample_maxz", 2: "tri_maxz", 3: "no_valid_sample", 4: "empty_aabb",
              5: "empty_lp", 6: "no_valid_tri", 7: "no_pairs", 8: "unsupported_light"}

SCALAR_NAMES = ["exit", "sample_maxz", "tri_maxz", "f", "S_x", "S_y",
                "lp_bx", "lp_by", "lp_bz", "lp_ex", "lp_ey", "lp_ez", "factor_x", "factor_y",
                "width_bins", "width_bins_log2", "num_bins", "valid_rect", "valid_samples", "valid_tris",
                "pairs", "ev_col_clip", "ev_row_clip", "ev_key_oor", "ev_valid_invalid_key",
                "cand_tests", "cull_pass", "plane_tris", "row_visits",
                "ms_lp", "ms_setup", "ms_bin", "ms_pair", "ms_occ", "wpr"]


class _Light(ctypes.Structure):
    _fields_ = [("right_radius", ctypes.c_float), ("top_radius", ctypes.c_float),
                ("position", ctypes.c_float * 3), ("view_dir", ctypes.c_float * 3), ("up_dir", ctypes.c_float * 3),
                ("res_w", ctypes.c_int), ("res_h", ctypes.c_int)]


class _Input(ctypes.Structure):
    _fields_ = [("verts", ctypes.c_void_p), ("vstride_floats", ctypes.c_int), ("nverts", ctypes.c_int),
                ("idx", ctypes.c_void_p), ("ntris", ctypes.c_int),
                ("pos4", ctypes.c_void_p), ("nrm4", ctypes.c_void_p), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("light", _Light), ("eps", ctypes.c_float), ("backface_cull", ctypes.c_int),
                ("grid_log2", ctypes.c_int), ("subset", ctypes.c_void_p), ("nsubset", ctypes.c_int),
                ("nthreads", ctypes.c_int), ("variant", ctypes.c_int)]


TRI_DATA_DTYPE = np.dtype([("v0", "<f4", 3), ("v1", "<f4", 3), ("v2", "<f4", 3), ("maxz", "<f4"),
                           ("minz", "<f4"), ("aabb_min", "<f4", 2), ("aabb_max", "<f4", 2)])
BIN_RANGE_DTYPE = np.dtype([("sx", "<i4"), ("sy", "<i4"), ("ex", "<i4"), ("ey", "<i4"),
                            ("minz", "<f4"), ("maxz", "<f4")])

_ARRAYS = {"shadow": np.float32, "masks": np.uint32, "computed": np.uint8, "rect": np.float32,
           "backface": np.int32, "key": np.int32, "sorted_key": np.int32, "sorted_id": np.int32,
           "bin_start": np.int32, "bin_end": np.int32, "bin_maxz": np.float32, "bin_minrange": np.float32,
           "tri_aabb": np.float32, "tri_valid": np.int32, "tri_data": TRI_DATA_DTYPE, "tri_count": np.int32,
           "tri_prefix": np.int64, "tri_range": BIN_RANGE_DTYPE, "pair_bin": np.int32, "pair_tri": np.int32,
           "bin_tri_start": np.int32, "bin_tri_end": np.int32}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            raise RuntimeError(f"{_SO} not built (make -C oracle)")
        L = ctypes.CDLL(_SO)
        L.or_run.argtypes = [ctypes.POINTER(_Input)]
        L.or_run.restype = ctypes.c_void_p
        L.or_free.argtypes = [ctypes.c_void_p]
        L.or_scalars.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.or_scalars.restype = ctypes.c_int
        L.or_light_bits.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.or_light_bits.restype = ctypes.c_int
        L.or_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
        L.or_array.restype = ctypes.c_longlong
        L.or_rowmask.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_uint, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int)]
        L.or_rowmask.restype = ctypes.c_ulonglong
        L.or_cta_range.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.or_cta_range.restype = ctypes.c_ulonglong
        L.or_make_rectangle.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.or_shade.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int] + [ctypes.c_void_p] * 6
        L.or_shade.restype = None
        L.or_max_threads.restype = ctypes.c_int
        L.or_move_model.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.or_move_model.restype = None
        L.or_occlude_all.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Input), ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p]
        L.or_occlude_all.restype = ctypes.c_int
        _lib = L
    return _lib


class OracleResult:
    def __init__(self, handle):
        self._h = handle
        n = lib().or_scalars(handle, None, 0)
        buf = (ctypes.c_double * n)()
        lib().or_scalars(handle, buf, n)
        vals = list(buf)
        self.scalars = dict(zip(SCALAR_NAMES, vals[:len(SCALAR_NAMES)]))
        rest = vals[len(SCALAR_NAMES):]
        self.sample_aabb = np.array(rest[0:8], np.float32)
        self.model_aabb = np.array(rest[8:16], np.float32)
        self.up_right_mat = np.array(rest[16:32], np.float32)
        lb = (ctypes.c_float * 128)()
        k = lib().or_light_bits(handle, lb)
        self.light_bits = np.array(lb[:k], np.float32)
        self.exit = int(self.scalars["exit"])
        self._cache = {}

    def __getattr__(self, name):
        if name.startswith("_") or name not in _ARRAYS:
            raise AttributeError(name)
        if name not in self._cache:
            n = lib().or_array(self._h, name.encode(), None)
            arr = np.empty(max(n, 0), _ARRAYS[name])
            if n > 0:
                lib().or_array(self._h, name.encode(), arr.ctypes.data)
            self._cache[name] = arr
        return self._cache[name]

    def __del__(self):
        try:
            if self._h:
                lib().or_free(self._h)
                self._h = None
        except Exception:
            pass


def _make_input(frame):
    sc, rx, L = frame.scene, frame.rx, frame.light
    keep = dict(verts=np.ascontiguousarray(sc.verts, np.float32), idx=np.ascontiguousarray(sc.idx, np.int32),
                pos=np.ascontiguousarray(rx.pos, np.float32), nrm=np.ascontiguousarray(rx.nrm, np.float32))
    inp = _Input()
    inp.verts = keep["verts"].ctypes.data
    inp.vstride_floats = keep["verts"].shape[1]
    inp.nverts = keep["verts"].shape[0]
    inp.idx = keep["idx"].ctypes.data
    inp.ntris = keep["idx"].shape[0]
    inp.pos4 = keep["pos"].ctypes.data
    inp.nrm4 = keep["nrm"].ctypes.data
    inp.w = rx.w
    inp.h = rx.h
    inp.light = _Light(L.right_radius, L.top_radius, (ctypes.c_float * 3)(*L.position),
                       (ctypes.c_float * 3)(*L.view_dir), (ctypes.c_float * 3)(*L.up_dir), L.res_w, L.res_h)
    inp.eps = frame.eps
    inp.backface_cull = frame.backface_cull
    inp.grid_log2 = frame.grid_log2
    return inp, keep


def occlude_all(result, frame, ids):
    """Masks of receivers `ids` against every light-plane triangle (no bins)."""
    inp, keep = _make_input(frame)
    ids = np.ascontiguousarray(ids, np.int32)
    R = frame.light.res_w
    out = np.zeros((len(ids), R * (1 if R <= 32 else 2)), np.uint32)
    rc = lib().or_occlude_all(result._h, ctypes.byref(inp), ids.ctypes.data, len(ids), out.ctypes.data)
    if rc != 0:
        raise RuntimeError("occlude_all needs a frame that reached the occlusion stage")
    return out


def run(frame, subset=None, nthreads=0, variant=0):
    """Run the oracle on a workloads.Frame.  `subset`: dense receiver ids for
    which the occlusion stage is evaluated (None = all).  `variant`: 0 the
    per-sample occlusion kernel, 1 the bin-queue (CTA) kernel."""
    sc, rx, L = frame.scene, frame.rx, frame.light
    verts = np.ascontiguousarray(sc.verts, np.float32)
    idx = np.ascontiguousarray(sc.idx, np.int32)
    pos = np.ascontiguousarray(rx.pos, np.float32)
    nrm = np.ascontiguousarray(rx.nrm, np.float32)
    inp = _Input()
    inp.verts = verts.ctypes.data
    inp.vstride_floats = verts.shape[1]
    inp.nverts = verts.shape[0]
    inp.idx = idx.ctypes.data
    inp.ntris = idx.shape[0]
    inp.pos4 = pos.ctypes.data
    inp.nrm4 = nrm.ctypes.data
    inp.w = rx.w
    inp.h = rx.h
    inp.light = _Light(L.right_radius, L.top_radius, (ctypes.c_float * 3)(*L.position),
                       (ctypes.c_float * 3)(*L.view_dir), (ctypes.c_float * 3)(*L.up_dir), L.res_w, L.res_h)
    inp.eps = frame.eps
    inp.backface_cull = frame.backface_cull
    inp.grid_log2 = frame.grid_log2
    sub = None
    if subset is not None:
        sub = np.ascontiguousarray(subset, np.int32)
        inp.subset = sub.ctypes.data
        inp.nsubset = sub.shape[0]
    inp.nthreads = nthreads
    inp.variant = variant
    h = lib().or_run(ctypes.byref(inp))
    return OracleResult(h)


def rowmask(D, B, n, light_w):
    clip = ctypes.c_int()
    m = lib().or_rowmask(D, B, n, light_w, ctypes.byref(clip))
    return int(m), bool(clip.value)


def cta_row_range(B, D, light_w):
    """calLightRowRange of the bin-queue kernel (Step5_2_shadowCal.inl:302-357)."""
    return int(lib().or_cta_range(0, B, D, 0.0, light_w))


def cta_height_range(A, B, C, light_w):
    """calLightHeightRange of the bin-queue kernel (Step5_2_shadowCal.inl:225-300), H = W."""
    return int(lib().or_cta_range(1, A, B, C, light_w))


def make_rectangle(mat16, p3):
    m = np.ascontiguousarray(mat16, np.float32)
    p = np.ascontiguousarray(p3, np.float32)
    out = np.empty(3, np.float32)
    lib().or_make_rectangle(m.ctypes.data, p.ctypes.data, out.ctypes.data)
    return out


def max_threads():
    return lib().or_max_threads()


def move_model(verts, m16):
    """FD::moveModel (moveModel.inl:6-30) on a copy of verts ((V, stride)
    float32); m16: the column-major 4x4.  Returns the transformed copy."""
    v = np.array(verts, np.float32, order="C", copy=True)
    m = np.ascontiguousarray(np.asarray(m16, np.float32).ravel())
    assert m.size == 16 and v.ndim == 2 and v.shape[1] >= 3
    lib().or_move_model(v.ctypes.data, v.shape[1], v.shape[0], m.ctypes.data)
    return v


def shade(shadow, g0, g1, g2, g3, samples, cam, La, Ld, Ls):
    """phoneShading.shf:23-61 restated in float (see or_shade); returns float4
    per texel, shape (n, 4)."""
    f32 = lambda a: np.ascontiguousarray(np.asarray(a, np.float32))
    shadow, g0, g1, g2, g3 = f32(shadow).ravel(), f32(g0), f32(g1), f32(g2), f32(g3)
    samples, cam, La, Ld, Ls = f32(samples).ravel(), f32(cam), f32(La), f32(Ld), f32(Ls)
    n = shadow.size
    out = np.zeros((n, 4), np.float32)
    lib().or_shade(n, shadow.ctypes.data, g0.ctypes.data, g1.ctypes.data, g2.ctypes.data, g3.ctypes.data,
                   samples.size // 3, samples.ctypes.data, cam.ctypes.data, La.ctypes.data, Ld.ctypes.data,
                   Ls.ctypes.data, out.ctypes.data)
    return out
