"""Deterministic synthetic workloads for the soft-shadow visibility path.

The geometry is produced by the host C++ generators in ``csrc/scenes.cpp``
(``libfd4d_scenes.so``), so that the bytes are identical wherever this image
runs.  The five workloads follow the reference's use of ``FDZB_Soft_Shadow``:
a receiver G-buffer (positions + normals, NaN = background), one occluder
triangle soup (vertices + int3 indices, as ``FD::FDaddModel`` takes them,
4DZB/FDZB_Soft_Shadow/api.cpp:152-160) and a rectangular area light
(``AreaLightDes``, 4DZB/FDZB_Soft_Shadow/AreaLightDesc.h:3-13).

  plane-256      256x256 receivers on z=0 over [-4,4]^2, 1x1 light at z=4,
                 16x16 light samples, 10k triangles (the CPU-sized case)
  heightfield-1k 1024x1024 receivers on z=0.2 sin x cos y over [-8,8]^2,
                 2x2 light at z=6, 32x32 samples, 500 UV spheres = 200k tris
  boxes-1080p    1920x1080 ray-cast G-buffer (ground + 20 boxes), 3x1 light at
                 z=6, 64x64 samples, 1M clustered log-normal triangles
  batch-64       64 instances of uneven resolution / light grid / light pose
                 over the heightfield-1k scene
  serve-poisson  Poisson request stream over 4 resident sphere scenes
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SCENES_SO = os.path.join(_HERE, "libfd4d_scenes.so")
_lib = None


def _scenes():
    global _lib
    if _lib is None:
        if not os.path.exists(_SCENES_SO):
            raise RuntimeError(f"{_SCENES_SO} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(_SCENES_SO)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.fdgen_receivers_plane.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, fp, fp]
        lib.fdgen_receivers_heightfield.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                                    ctypes.c_float, fp, fp]
        lib.fdgen_receivers_boxes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, fp, fp]
        lib.fdgen_tris_uniform.argtypes = [ctypes.c_long, ctypes.c_uint64, fp, ctypes.c_float, ctypes.c_float, fp, ip]
        lib.fdgen_tris_uniform.restype = ctypes.c_long
        lib.fdgen_tris_clusters.argtypes = [ctypes.c_long, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, fp,
                                            ctypes.c_float, ctypes.c_float, fp, ip]
        lib.fdgen_tris_clusters.restype = ctypes.c_long
        lib.fdgen_sphere_counts.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_long),
                                            ctypes.POINTER(ctypes.c_long)]
        lib.fdgen_spheres.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, fp, ctypes.c_float,
                                      ctypes.c_float, fp, ip]
        lib.fdgen_spheres.restype = ctypes.c_long
        lib.fdgen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_long, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_double)]
        dp = ctypes.POINTER(ctypes.c_double)
        lib.fdgen_boxes_layout.argtypes = [ctypes.c_int, ctypes.c_uint64, dp]
        lib.fdgen_gbuffer_boxes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, ctypes.c_double,
                                            ctypes.c_double, fp, fp, fp, fp, fp]
        _lib = lib
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


@dataclass
class LightDesc:
    """Mirror of the reference's AreaLightDes (half extents, pose, sample grid)."""
    right_radius: float
    top_radius: float
    position: tuple
    view_dir: tuple = (0.0, 0.0, -1.0)
    up_dir: tuple = (0.0, 1.0, 0.0)
    res_w: int = 16
    res_h: int = 16


@dataclass
class Scene:
    verts: np.ndarray   # (V, 4) float32, w = 1 (16-byte vertex stride)
    idx: np.ndarray     # (T, 3) int32

    @property
    def ntris(self):
        return int(self.idx.shape[0])

    @property
    def nverts(self):
        return int(self.verts.shape[0])


@dataclass
class Receivers:
    pos: np.ndarray     # (H, W, 4) float32
    nrm: np.ndarray     # (H, W, 4) float32

    @property
    def w(self):
        return int(self.pos.shape[1])

    @property
    def h(self):
        return int(self.pos.shape[0])


@dataclass
class Frame:
    name: str
    scene: Scene
    rx: Receivers
    light: LightDesc
    eps: float = 1e-3
    backface_cull: int = 1
    grid_log2: int = 4
    tile_width: int = 8
    occ_waves: int = 0       # waves per occlusion work item (0: the library's default)
    list_split: int = 0      # stretches per bin triangle list (0: the library's rule)


def receivers_plane(w, h, half=4.0):
    pos = np.empty((h, w, 4), np.float32)
    nrm = np.empty((h, w, 4), np.float32)
    assert _scenes().fdgen_receivers_plane(w, h, half, _fp(pos), _fp(nrm)) == 0
    return Receivers(pos, nrm)


def receivers_heightfield(w, h, hx=8.0, hy=8.0, amp=0.2):
    pos = np.empty((h, w, 4), np.float32)
    nrm = np.empty((h, w, 4), np.float32)
    assert _scenes().fdgen_receivers_heightfield(w, h, hx, hy, amp, _fp(pos), _fp(nrm)) == 0
    return Receivers(pos, nrm)


def receivers_boxes(w, h, nbox=20, seed=3):
    pos = np.empty((h, w, 4), np.float32)
    nrm = np.empty((h, w, 4), np.float32)
    assert _scenes().fdgen_receivers_boxes(w, h, nbox, seed, _fp(pos), _fp(nrm)) == 0
    return Receivers(pos, nrm)


def tris_uniform(ntri, seed, box=(-3, 3, -3, 3, 0.5, 3), e0=0.05, e1=0.3):
    verts = np.empty((3 * ntri, 4), np.float32)
    idx = np.empty((ntri, 3), np.int32)
    b = np.asarray(box, np.float32)
    assert _scenes().fdgen_tris_uniform(ntri, seed, _fp(b), e0, e1, _fp(verts), _ip(idx)) == ntri
    return Scene(verts, idx)


def tris_clusters(ntri, seed, ncluster=8, csig=1.0, box=(-8, 8, -8, 8, 0.5, 5), mu=-3.0, sigma=1.0):
    verts = np.empty((3 * ntri, 4), np.float32)
    idx = np.empty((ntri, 3), np.int32)
    b = np.asarray(box, np.float32)
    assert _scenes().fdgen_tris_clusters(ntri, seed, ncluster, csig, _fp(b), mu, sigma, _fp(verts), _ip(idx)) == ntri
    return Scene(verts, idx)


def spheres(nsph, seed, slices=20, stacks=11, box=(-6, 6, -6, 6, 1, 4), r0=0.2, r1=1.0):
    nv = ctypes.c_long()
    nt = ctypes.c_long()
    assert _scenes().fdgen_sphere_counts(slices, stacks, ctypes.byref(nv), ctypes.byref(nt)) == 0
    verts = np.empty((nsph * nv.value, 4), np.float32)
    idx = np.empty((nsph * nt.value, 3), np.int32)
    b = np.asarray(box, np.float32)
    got = _scenes().fdgen_spheres(nsph, slices, stacks, seed, _fp(b), r0, r1, _fp(verts), _ip(idx))
    assert got == nsph * nt.value
    return Scene(verts, idx)


def uniform(seed, n, a, b):
    out = np.empty(n, np.float64)
    assert _scenes().fdgen_uniform(seed, n, a, b, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return out


# ---------------------------------------------------------------------------
# The named workloads
# ---------------------------------------------------------------------------
def plane_256(ntri=10_000, res=256, light_res=16, seed=0):
    """256x256 receivers on z=0 over [-4,4]^2; 1x1 light at z=4 facing -z;
    16x16 samples; 10k triangles, centroid U[-3,3]^2 x [0.5,3], edge U[0.05,0.3]."""
    return Frame("plane-256", tris_uniform(ntri, seed), receivers_plane(res, res, 4.0),
                 LightDesc(0.5, 0.5, (0.0, 0.0, 4.0), res_w=light_res, res_h=light_res))


def heightfield_1k(res=1024, nsph=500, light_res=32, seed=2):
    """1024^2 receivers on z=0.2 sin x cos y over [-8,8]^2; 2x2 light at z=6,
    32x32 samples; 500 UV spheres (20x11, 400 tris each), r U[0.2,1],
    centre U[-6,6]^2 x [1,4]."""
    return Frame("heightfield-1k", spheres(nsph, seed), receivers_heightfield(res, res),
                 LightDesc(1.0, 1.0, (0.0, 0.0, 6.0), res_w=light_res, res_h=light_res))


def boxes_1080p(w=1920, h=1080, ntri=1_000_000, light_res=64, seed=3):
    """1920x1080 ray-cast G-buffer (ground + 20 boxes); 3x1 light at z=6,
    64x64 samples; 1M triangles, log-normal edge (mu=-3, sigma=1), 8 Gaussian
    clusters (sigma=1) over [-8,8]^2 x [0.5,5]."""
    return Frame("boxes-1080p", tris_clusters(ntri, seed), receivers_boxes(w, h, 20, seed),
                 LightDesc(1.5, 0.5, (0.0, 0.0, 6.0), res_w=light_res, res_h=light_res))


BATCH_RES = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
BATCH_GRIDS = [8, 16, 32]


def batch_instances(n=64, seed=4, scene=None, res_choices=BATCH_RES, grid_choices=BATCH_GRIDS):
    """n instances over the heightfield-1k scene: resolution drawn from
    res_choices, light grid from grid_choices, light centre U[-2,2]^2 at
    height U[4,8] (2x2 light facing -z)."""
    if scene is None:
        scene = spheres(500, 2)
    u = uniform(seed, 5 * n, 0.0, 1.0).reshape(n, 5)
    frames = []
    rx_cache = {}
    for i in range(n):
        w, h = res_choices[min(int(u[i, 0] * len(res_choices)), len(res_choices) - 1)]
        g = grid_choices[min(int(u[i, 1] * len(grid_choices)), len(grid_choices) - 1)]
        px, py, pz = -2 + 4 * u[i, 2], -2 + 4 * u[i, 3], 4 + 4 * u[i, 4]
        key = (w, h)
        if key not in rx_cache:
            rx_cache[key] = receivers_heightfield(w, h, 8.0, 8.0 * h / w if w != h else 8.0)
        frames.append(Frame(f"inst{i}-{w}x{h}-g{g}", scene, rx_cache[key],
                            LightDesc(1.0, 1.0, (float(px), float(py), float(pz)), res_w=g, res_h=g)))
    return frames


def serve_scenes(sizes=(100_000, 300_000, 1_000_000, 2_000_000), seeds=(1, 2, 3, 4)):
    """Resident sphere scenes of the given triangle counts (400 tris per sphere)."""
    return [spheres(max(1, t // 400), s) for t, s in zip(sizes, seeds)]


def poisson_requests(n, rate_hz=500.0, nscenes=4, seed=5, res_choices=BATCH_RES, grids=(16, 32)):
    """n requests: arrival time (Poisson at rate_hz), scene index, resolution,
    light grid and light pose, all drawn from one seeded stream."""
    u = uniform(seed, 7 * n, 0.0, 1.0).reshape(n, 7)
    gaps = -np.log(np.maximum(1.0 - u[:, 0], 1e-300)) / rate_hz
    t = np.cumsum(gaps)
    reqs = []
    for i in range(n):
        sc = min(int(u[i, 1] * nscenes), nscenes - 1)
        w, h = res_choices[min(int(u[i, 2] * len(res_choices)), len(res_choices) - 1)]
        g = grids[min(int(u[i, 3] * len(grids)), len(grids) - 1)]
        pose = (-2 + 4 * u[i, 4], -2 + 4 * u[i, 5], 4 + 4 * u[i, 6])
        reqs.append(dict(t=float(t[i]), scene=sc, w=w, h=h, grid=g, pose=tuple(float(x) for x in pose)))
    return reqs


@dataclass
class GbufferScene:
    """The boxes G-buffer scene (ground z = 0 + axis-aligned boxes, pinhole
    camera) as fd_gbuffer_boxes takes it."""
    width: int
    height: int
    boxes: np.ndarray       # (nbox, 6) float64: min xyz, max xyz
    materials: np.ndarray   # (nbox + 1, 10) float32: ka, kd, ks, shininess ([0] the ground)
    eye: tuple = (0.0, -13.0, 9.0)
    at: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 0.0, 1.0)
    fovy_deg: float = 50.0
    ground_half: float = 10.0


def boxes_materials(nbox):
    """Ground: grey, weak specular; box b: a hue around the colour wheel."""
    m = np.zeros((nbox + 1, 10), np.float32)
    m[0] = (0.08, 0.08, 0.08, 0.55, 0.55, 0.5, 0.15, 0.15, 0.15, 8.0)
    for b in range(nbox):
        t = 6.283185307179586 * b / max(nbox, 1)
        kd = 0.45 + 0.35 * np.cos(t + np.array([0.0, 2.0943951, 4.1887902]))
        m[1 + b, 0:3] = 0.1 * kd
        m[1 + b, 3:6] = kd
        m[1 + b, 6:9] = 0.35
        m[1 + b, 9] = 24.0 + 8.0 * (b % 4)
    return m


def gbuffer_scene_boxes(w=1920, h=1080, nbox=20, seed=3):
    """The scene of receivers_boxes (same boxes, camera, ground) with materials."""
    boxes = np.zeros((nbox, 6), np.float64)
    assert _scenes().fdgen_boxes_layout(nbox, seed, boxes.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return GbufferScene(w, h, boxes, boxes_materials(nbox))


def gbuffer_boxes_host(sc):
    """Host ray cast of the same scene (the device producer's reference):
    four (H, W, 4) float32 targets g0..g3."""
    H, W = sc.height, sc.width
    g = [np.empty((H, W, 4), np.float32) for _ in range(4)]
    dp = lambda a: np.ascontiguousarray(np.asarray(a, np.float64)).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    bx = np.ascontiguousarray(sc.boxes, np.float64)
    mt = np.ascontiguousarray(sc.materials, np.float32)
    eye, at, up = (np.asarray(v, np.float64) for v in (sc.eye, sc.at, sc.up))
    rc = _scenes().fdgen_gbuffer_boxes(W, H, bx.shape[0], dp(bx), dp(eye), dp(at), dp(up), sc.fovy_deg,
                                       sc.ground_half, _fp(mt), _fp(g[0]), _fp(g[1]), _fp(g[2]), _fp(g[3]))
    assert rc == 0
    return g
