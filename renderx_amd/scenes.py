"""Deterministic synthetic scenes for the path-tracing workloads.

The reference ships no scene data (its ``data/`` folder is git-ignored), so
every workload is generated here from a fixed seed.  Scenes use flat face
normals only (no ``vn`` records), the case in which the reference's shading
normal equals the geometric normal (src/SimpleShape.cpp:393-421).

A scene is a flat description that maps 1:1 onto what the reference builds
through ``SceneObject`` fields and ``Scene::buildKDTree`` (src/Scene.cpp:248-285):
world-space vertices, triangles with global vertex indices, and one record per
object with its contiguous triangle range and material parameters.

Workloads (sizes are the ones the benchmarks use):
  cornell      36 triangles, 256x256, 16 spp, max depth 8
  icosphere    1,003,526 triangles (displaced geodesic sphere + ground + 2 lights),
               1024x1024, 64 spp, max depth 10
  soup_batch   64 triangle-soup scenes of 10k..2M triangles
  media        Cornell box + homogeneous medium + glass sphere + 499,280-triangle
               displaced sphere, 512x512, 256 spp, max depth 16
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# object types (same numbering as include/rxhip.h RX_OBJ_*)
DIFFUSE, GLOSSY, PHONG, REFRACTIVE, EMISSIVE, VPM = range(6)

OBJ_DTYPE = np.dtype([
    ("type", "<i4"), ("tri_begin", "<u4"), ("tri_count", "<u4"), ("pad0", "<u4"),
    ("color", "<f4", 3), ("exponent", "<f4"), ("color2", "<f4", 3), ("ior", "<f4"),
    ("g", "<f4"), ("pad1", "<f4", 3)])
assert OBJ_DTYPE.itemsize == 64

CAM_DTYPE = np.dtype([("position", "<f4", 3), ("focus", "<f4", 3), ("up", "<f4", 3),
                      ("sight_dist", "<u4"), ("width", "<i4"), ("height", "<i4")])
assert CAM_DTYPE.itemsize == 48

SEED_BASE = 0x52454E44


@dataclass
class Scene:
    name: str
    verts: np.ndarray            # (V,3) float32 world positions
    tris: np.ndarray             # (T,3) uint32 global vertex indices
    objects: np.ndarray          # (O,) OBJ_DTYPE
    camera: np.ndarray           # () CAM_DTYPE
    spp: int = 16
    max_depth: int = 8
    rr_start_depth: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_tris(self) -> int:
        return int(self.tris.shape[0])

    def tri_obj(self) -> np.ndarray:
        out = np.empty(self.n_tris, np.int32)
        for i, o in enumerate(self.objects):
            out[o["tri_begin"]:o["tri_begin"] + o["tri_count"]] = i
        return out


class _Builder:
    def __init__(self):
        self.v, self.t, self.o = [], [], []
        self.nv = 0
        self.nt = 0

    def add(self, verts, tris, otype, color=(0, 0, 0), exponent=0.0, color2=(0, 0, 0), ior=1.0, g=0.0):
        verts = np.asarray(verts, np.float64)
        tris = np.asarray(tris, np.int64)
        rec = np.zeros((), OBJ_DTYPE)
        rec["type"] = otype
        rec["tri_begin"] = self.nt
        rec["tri_count"] = len(tris)
        rec["color"] = color
        rec["exponent"] = exponent
        rec["color2"] = color2
        rec["ior"] = ior
        rec["g"] = g
        self.v.append(verts)
        self.t.append(tris + self.nv)
        self.o.append(rec)
        self.nv += len(verts)
        self.nt += len(tris)
        return len(self.o) - 1

    def finish(self):
        v = np.concatenate(self.v).astype(np.float32)
        v = v + np.float32(0.0)  # no negative zeros in the inputs
        t = np.concatenate(self.t).astype(np.uint32)
        return v, t, np.array(self.o, OBJ_DTYPE)


def make_camera(position, focus, up, width, height, fov_deg=40.0):
    cam = np.zeros((), CAM_DTYPE)
    cam["position"] = position
    cam["focus"] = focus
    cam["up"] = up
    cam["width"] = width
    cam["height"] = height
    # integer focal length in pixels (include/Camera.h:16)
    cam["sight_dist"] = int(round((height / 2) / math.tan(math.radians(fov_deg) / 2)))
    return cam


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); face normal = (p1-p0)x(p2-p0)."""
    return np.array([p0, p1, p2, p3], np.float64), np.array([[0, 1, 2], [0, 2, 3]])


def _orient(verts, tris, want_normal):
    """Flip triangle winding so every face normal points along want_normal(centroid)."""
    v = verts[tris]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    c = v.mean(axis=1)
    flip = np.einsum("ij,ij->i", n, want_normal(c)) < 0
    tris = tris.copy()
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris


def _box(center, size, rot_y_deg):
    """Axis box rotated about y; 12 triangles with outward normals."""
    sx, sy, sz = (s / 2 for s in size)
    corners = np.array([[x, y, z] for x in (-sx, sx) for y in (-sy, sy) for z in (-sz, sz)], np.float64)
    a = math.radians(rot_y_deg)
    R = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    v = corners @ R.T + np.asarray(center, np.float64)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a_, b_, c_, d_ in faces:
        tris += [[a_, b_, c_], [a_, c_, d_]]
    tris = np.array(tris)
    ctr = np.asarray(center, np.float64)
    return v, _orient(v, tris, lambda c: c - ctr)


def _axis_box(lo, hi):
    v, t = _box([(a + b) / 2 for a, b in zip(lo, hi)], [b - a for a, b in zip(lo, hi)], 0.0)
    return v, t


def _cornell(b: _Builder, light_y=0.999):
    W = 0.75
    # floor y=0, ceiling y=1, back z=1 (white); normals point into the box
    fl = _quad([0, 0, 0], [0, 0, 1], [1, 0, 1], [1, 0, 0])
    ce = _quad([0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1])
    bk = _quad([0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1])
    wv = np.concatenate([fl[0], ce[0], bk[0]])
    wt = np.concatenate([fl[1], ce[1] + 4, bk[1] + 8])
    wt = _orient(wv, wt, lambda c: np.array([0.5, 0.5, 0.5]) - c)
    b.add(wv, wt, DIFFUSE, (W, W, W))
    lv, lt = _quad([0, 0, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1])
    b.add(lv, _orient(lv, lt, lambda c: np.tile([1.0, 0, 0], (len(c), 1))), DIFFUSE, (0.63, 0.06, 0.05))
    rv, rt = _quad([1, 0, 0], [1, 0, 1], [1, 1, 1], [1, 1, 0])
    b.add(rv, _orient(rv, rt, lambda c: np.tile([-1.0, 0, 0], (len(c), 1))), DIFFUSE, (0.14, 0.45, 0.09))
    sv, st = _box([0.35, 0.15, 0.38], [0.3, 0.3, 0.3], -17.0)
    b.add(sv, st, DIFFUSE, (W, W, W))
    tv, tt = _box([0.66, 0.30, 0.65], [0.3, 0.6, 0.3], 17.0)
    b.add(tv, tt, DIFFUSE, (W, W, W))
    h = 0.125
    qv, qt = _quad([0.5 - h, light_y, 0.5 - h], [0.5 + h, light_y, 0.5 - h],
                   [0.5 + h, light_y, 0.5 + h], [0.5 - h, light_y, 0.5 + h])
    b.add(qv, _orient(qv, qt, lambda c: np.tile([0, -1.0, 0], (len(c), 1))), EMISSIVE, (15, 15, 15))


def cornell(width=256, height=256, spp=16, max_depth=8, rr_start_depth=3) -> Scene:
    b = _Builder()
    _cornell(b)
    v, t, o = b.finish()
    cam = make_camera([0.5, 0.5, -1.374], [0.5, 0.5, 0.5], [0, 1, 0], width, height)
    return Scene("cornell", v, t, o, cam, spp, max_depth, rr_start_depth)


def lit_plane(width=32, height=32, spp=64, max_depth=8, rr_start_depth=0) -> Scene:
    """An open scene with a smooth image: a large diffuse floor lit by a quad
    light beside the view, seen from above.  No edge, emitter or point behind
    the camera is in view, so every strategy of a bidirectional tracer sees the
    same smooth radiance field."""
    b = _Builder()
    fv, ft = _quad([-3, 0, -3], [-3, 0, 3], [3, 0, 3], [3, 0, -3])
    b.add(fv, _orient(fv, ft, lambda c: np.tile([0, 1.0, 0], (len(c), 1))), DIFFUSE, (0.5, 0.5, 0.5))
    h = 0.15
    qv, qt = _quad([1.0 - h, 0.5, -h], [1.0 + h, 0.5, -h], [1.0 + h, 0.5, h], [1.0 - h, 0.5, h])
    b.add(qv, _orient(qv, qt, lambda c: np.tile([0, -1.0, 0], (len(c), 1))), EMISSIVE, (20, 20, 20))
    v, t, o = b.finish()
    cam = make_camera([0, 1.2, 0], [0, 0, 0], [0, 0, 1], width, height)
    return Scene("lit_plane", v, t, o, cam, spp, max_depth, rr_start_depth)


def fog_room(width=16, height=16, spp=64, max_depth=30, rr_start_depth=0, sigma_s=0.3, sigma_a=0.05, g=0.5) -> Scene:
    """The Cornell box (diffuse walls and boxes, quad light) inside a
    homogeneous, index-matched medium that also holds the camera.  Every
    vertex of a path that can still reach the light lies in the medium, so
    every connection of a bidirectional tracer can exist and its image has the
    path tracer's expectation (up to the depth limits: max_depth 30 with
    Russian roulette at every bounce leaves paths that long a negligible
    share).  Diffuse only: a Phong lobe pick would separate the two
    estimators."""
    b = _Builder()
    _cornell(b)
    mv, mt = _axis_box([-0.6, -0.6, -1.8], [1.6, 1.6, 1.6])
    b.add(mv, mt, VPM, (sigma_s, sigma_s, sigma_s), 0.0, (sigma_a, sigma_a, sigma_a), 1.0, g)
    v, t, o = b.finish()
    cam = make_camera([0.5, 0.5, -1.374], [0.5, 0.5, 0.5], [0, 1, 0], width, height)
    return Scene("fog_room", v, t, o, cam, spp, max_depth, rr_start_depth)


def absorbing_slab(width=8, height=8, spp=64, max_depth=8, sigma_a=1.0, thickness=0.5, le=1.0) -> Scene:
    """A purely absorbing, index-matched medium slab (sigma_s = 0, IOR 1) between
    the camera and a large emitter that fills the view.  Every camera ray
    crosses the slab's front and back faces and ends on the emitter, so a
    pixel's expected value is le * exp(-sigma_a * l), l the length of its ray
    inside the slab (meta["expected"], per pixel, from the pixel-centre ray).
    A narrow field of view keeps l within 0.4 % of the thickness."""
    b = _Builder()
    sv, st = _axis_box([-0.8, -0.8, 0.0], [0.8, 0.8, thickness])
    b.add(sv, st, VPM, (0.0, 0.0, 0.0), color2=(sigma_a, sigma_a, sigma_a), ior=1.0, g=0.0)
    z = thickness + 1.5
    ev, et = _quad([-1.5, -1.5, z], [1.5, -1.5, z], [1.5, 1.5, z], [-1.5, 1.5, z])
    b.add(ev, _orient(ev, et, lambda c: np.tile([0, 0, -1.0], (len(c), 1))), EMISSIVE, (le, le, le))
    v, t, o = b.finish()
    cam = make_camera([0.0, 0.0, -3.0], [0.0, 0.0, 0.0], [0, 1, 0], width, height, fov_deg=10.0)
    # pixel-centre ray directions (src/Camera.cpp:38-47 with the jitter at 1/2)
    sd = float(cam["sight_dist"])
    xs = np.arange(width) + 0.5 - width / 2
    ys = height / 2 - (np.arange(height) + 0.5)
    dx, dy = np.meshgrid(xs, ys)
    cos_z = sd / np.sqrt(sd * sd + dx * dx + dy * dy)
    expected = le * np.exp(-sigma_a * thickness / cos_z)
    return Scene("absorbing_slab", v, t, o, cam, spp, max_depth, 0,
                 meta={"expected": expected, "sigma_a": sigma_a, "thickness": thickness})


# ---------------------------------------------------------------- geodesic spheres
def icosphere(freq: int):
    """Class-I geodesic sphere: 20*freq^2 triangles, 10*freq^2+2 shared vertices, unit radius."""
    p = (1 + 5 ** 0.5) / 2
    iv = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p],
                   [0, -1, -p], [0, 1, -p], [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], np.float64)
    iv /= np.linalg.norm(iv, axis=1, keepdims=True)
    faces = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                      [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                      [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    n = freq
    # lattice points (i, j) with i + j <= n on each face
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    m = (ii + jj) <= n
    ii, jj = ii[m], jj[m]
    kk = n - ii - jj
    local_index = -np.ones((n + 1, n + 1), np.int64)
    local_index[ii, jj] = np.arange(len(ii))
    # local triangles
    tl = []
    for i in range(n):
        for j in range(n - i):
            tl.append([local_index[i, j], local_index[i + 1, j], local_index[i, j + 1]])
            if i + j < n - 1:
                tl.append([local_index[i + 1, j], local_index[i + 1, j + 1], local_index[i, j + 1]])
    tl = np.array(tl, np.int64)
    A, B, C = iv[faces[:, 0]], iv[faces[:, 1]], iv[faces[:, 2]]
    pts = (kk[None, :, None] * A[:, None, :] + ii[None, :, None] * B[:, None, :] +
           jj[None, :, None] * C[:, None, :]) / n
    pts = pts.reshape(-1, 3)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    tris = (tl[None, :, :] + (np.arange(20) * len(ii))[:, None, None]).reshape(-1, 3)
    # merge the vertices shared along icosahedron edges
    q = np.round(pts * (1 << 22)).astype(np.int64)
    _, first, inv = np.unique(q, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(len(order))
    verts = pts[first[order]]
    tris = remap[inv[tris]]
    tris = _orient(verts, tris, lambda c: c)
    return verts, tris


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841) ^
         iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed))
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(1 << 24) * 2.0 - 1.0


def value_noise(p, freq, seed):
    """Trilinear value noise with smoothstep weights on an integer lattice."""
    x = p * freq
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    w = f * f * (3 - 2 * f)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                val = _hash3(i0[:, 0] + dx + 1024, i0[:, 1] + dy + 1024, i0[:, 2] + dz + 1024, seed)
                wt = ((w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) *
                      (w[:, 2] if dz else 1 - w[:, 2]))
                out += wt * val
    return out


def displaced_sphere(freq, seed, amplitude=0.05):
    v, t = icosphere(freq)
    nz = (value_noise(v, 4, seed) + 0.5 * value_noise(v, 8, seed + 1) + 0.25 * value_noise(v, 16, seed + 2))
    nz = nz / np.abs(nz).max()
    v = v * (1.0 + amplitude * nz)[:, None]
    t = _orient(v, t, lambda c: c)
    return v, t


def icosphere_scene(width=1024, height=1024, spp=64, max_depth=10, freq=224, seed=None) -> Scene:
    seed = SEED_BASE + 1 if seed is None else seed
    rng = np.random.default_rng(seed)
    b = _Builder()
    v, t = displaced_sphere(freq, seed)
    b.add(v, t, PHONG, (0.5, 0.5, 0.5), 50.0, (0.3, 0.3, 0.3))
    gy = -1.06
    gv, gt = _quad([-10, gy, -10], [-10, gy, 10], [10, gy, 10], [10, gy, -10])
    b.add(gv, _orient(gv, gt, lambda c: np.tile([0, 1.0, 0], (len(c), 1))), DIFFUSE, (0.5, 0.5, 0.5))
    for _ in range(2):
        z = rng.uniform(0.3, 0.9)
        phi = rng.uniform(0, 2 * math.pi)
        r = math.sqrt(1 - z * z)
        c = 4.0 * np.array([r * math.cos(phi), z, r * math.sin(phi)])
        n = -c / np.linalg.norm(c)
        a = np.cross(n, [0, 1.0, 0])
        a /= np.linalg.norm(a)
        bb = np.cross(n, a)
        qv = np.array([c - 0.5 * a - 0.5 * bb, c + 0.5 * a - 0.5 * bb, c + 0.5 * a + 0.5 * bb, c - 0.5 * a + 0.5 * bb])
        qt = np.array([[0, 1, 2], [0, 2, 3]])
        b.add(qv, _orient(qv, qt, lambda cc, n=n: np.tile(n, (len(cc), 1))), EMISSIVE, (10, 10, 10))
    v, t, o = b.finish()
    cam = make_camera([0.0, 1.0, 3.5], [0.0, -0.1, 0.0], [0, 1, 0], width, height)
    return Scene("icosphere", v, t, o, cam, spp, max_depth, 0, {"freq": freq})


def media_scene(width=512, height=512, spp=256, max_depth=16, freq=158, glass_freq=16) -> Scene:
    b = _Builder()
    _cornell(b)
    lo, hi = 0.002, 0.998
    mv, mt = _axis_box([lo, lo, lo], [hi, hi, hi])
    # homogeneous grey medium, boundary index-matched (inrate 1.0), HG g=0.7
    b.add(mv, mt, VPM, (0.5, 0.5, 0.5), 0.0, (0.05, 0.05, 0.05), 1.0, 0.7)
    gv, gt = icosphere(glass_freq)
    gv = gv * 0.15 + np.array([0.35, 0.451, 0.38])
    b.add(gv, gt, REFRACTIVE, (1, 1, 1), 0.0, (1, 1, 1), 1.5)
    dv, dt = displaced_sphere(freq, SEED_BASE + 3)
    dv = dv * 0.2 + np.array([0.76, 0.215, 0.22])
    b.add(dv, dt, PHONG, (0.5, 0.5, 0.5), 50.0, (0.3, 0.3, 0.3))
    v, t, o = b.finish()
    cam = make_camera([0.5, 0.5, -1.374], [0.5, 0.5, 0.5], [0, 1, 0], width, height)
    return Scene("media", v, t, o, cam, spp, max_depth, 0, {"freq": freq})


def glass_scene(width=64, height=64, spp=64, max_depth=10, glass_freq=8, phong_freq=8, phong=True) -> Scene:
    """The Cornell box with a glass sphere (IOR 1.5, refraction both ways,
    caustics) and a Phong sphere (a diffuse one with phong=False), without a
    medium: a surface scene for the bidirectional tracer."""
    b = _Builder()
    _cornell(b)
    gv, gt = icosphere(glass_freq)
    gv = gv * 0.15 + np.array([0.35, 0.451, 0.38])
    b.add(gv, gt, REFRACTIVE, (1, 1, 1), 0.0, (1, 1, 1), 1.5)
    dv, dt = icosphere(phong_freq)
    dv = dv * 0.2 + np.array([0.76, 0.215, 0.22])
    if phong:
        b.add(dv, dt, PHONG, (0.5, 0.5, 0.5), 50.0, (0.3, 0.3, 0.3))
    else:
        b.add(dv, dt, DIFFUSE, (0.5, 0.5, 0.5))
    v, t, o = b.finish()
    cam = make_camera([0.5, 0.5, -1.374], [0.5, 0.5, 0.5], [0, 1, 0], width, height)
    return Scene("glass", v, t, o, cam, spp, max_depth, 0)


def soup_scene(index: int, n_tris: int | None = None) -> Scene:
    """One scene of the 64-scene triangle-soup batch (seeded by its index)."""
    rng = np.random.default_rng(SEED_BASE + 2 + 7919 * index)
    n = int(round(math.exp(rng.uniform(math.log(1e4), math.log(2e6))))) if n_tris is None else n_tris
    res = int(rng.choice([128, 256, 512, 1024]))
    spp = int(rng.choice([4, 16, 64]))
    c = rng.uniform(0, 1, (n, 3))
    L = np.exp(rng.uniform(math.log(0.001), math.log(0.05), n))
    u = rng.normal(size=(n, 3, 3))
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    v = (c[:, None, :] + 0.5 * L[:, None, None] * u).reshape(-1, 3)
    t = np.arange(3 * n).reshape(n, 3)
    b = _Builder()
    half = n // 2
    for k, (lo_, hi_) in enumerate([(0, half), (half, n)]):
        sel_t = t[lo_:hi_] - 3 * lo_
        sel_v = v[3 * lo_:3 * hi_]
        if rng.uniform() < 0.5:
            rho = rng.uniform(0.2, 0.8)
            b.add(sel_v, sel_t, DIFFUSE, (rho, rho, rho))
        else:
            b.add(sel_v, sel_t, PHONG, (0.5, 0.5, 0.5), rng.uniform(10, 100), (0.3, 0.3, 0.3))
    h = 0.25
    qv, qt = _quad([0.5 - h, 1.5, 0.5 - h], [0.5 + h, 1.5, 0.5 - h], [0.5 + h, 1.5, 0.5 + h], [0.5 - h, 1.5, 0.5 + h])
    b.add(qv, _orient(qv, qt, lambda cc: np.tile([0, -1.0, 0], (len(cc), 1))), EMISSIVE, (20, 20, 20))
    vv, tt, oo = b.finish()
    d = rng.normal(size=3)
    d[1] = abs(d[1])
    d /= np.linalg.norm(d)
    cam = make_camera(list(np.array([0.5, 0.5, 0.5]) + 2.5 * d), [0.5, 0.5, 0.5], [0, 1, 0], res, res)
    return Scene(f"soup{index}", vv, tt, oo, cam, spp, 8, 0, {"n_tris": n})


def ray_batch(scene_verts_radius: float, n: int, seed: int):
    """Hit-parity ray batch: half aimed from a radius-3 sphere at points on the
    unit sphere, half with uniform origins in [-2,2]^3 and uniform directions."""
    rng = np.random.default_rng(seed)
    h = n // 2

    def sph(k):
        z = 1 - 2 * rng.uniform(size=k)
        phi = 2 * math.pi * rng.uniform(size=k)
        r = np.sqrt(np.maximum(0, 1 - z * z))
        return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)

    o1 = 3.0 * scene_verts_radius * sph(h)
    tgt = scene_verts_radius * sph(h)
    d1 = tgt - o1
    o2 = rng.uniform(-2, 2, (n - h, 3)) * scene_verts_radius
    d2 = sph(n - h)
    o = np.concatenate([o1, o2]).astype(np.float32)
    d = np.concatenate([d1, d2]).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    return o + np.float32(0), d + np.float32(0)
