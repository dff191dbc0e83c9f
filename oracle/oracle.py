"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product package
renderx_amd never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

RENDER_PARAMS = np.dtype([("spp", "<u4"), ("max_depth", "<u4"), ("rr_start_depth", "<u4"), ("mode", "<u4"),
                          ("seed", "<u8"), ("sample_offset", "<u4"), ("flags", "<u4")])

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_tree_create.restype = P
        L.orc_tree_create.argtypes = [C.c_int, P, C.c_int, P, C.c_int]
        L.orc_tree_free.argtypes = [P]
        L.orc_tree_sizes.argtypes = [P, P, P]
        L.orc_tree_serialize.argtypes = [P] * 7
        L.orc_face_normals.argtypes = [C.c_int, P, C.c_int, P, P]
        L.orc_intersect.argtypes = [P, C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P, P, C.c_int]
        L.orc_barycentrics.argtypes = [C.c_int, P, P, P, P, P, P]
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [C.c_int, P, C.c_int, P, C.c_int, P, P, C.c_int]
        L.orc_scene_free.argtypes = [P]
        L.orc_scene_set_camera.argtypes = [P, P]
        L.orc_scene_tree.restype = P
        L.orc_scene_tree.argtypes = [P]
        L.orc_render.argtypes = [P, P, P, P, C.c_int, P]
        L.orc_mt_nth.restype = C.c_uint32
        L.orc_mt_nth.argtypes = [C.c_uint32]
        L.orc_max_threads.restype = C.c_int
        L.orc_sampler_draws.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, P]
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_bpt_density_stats.argtypes = [P, P, P]
        L.orc_refract_dir.restype = C.c_int
        L.orc_refract_dir.argtypes = [C.c_int, P, P, C.c_float, C.c_float, P, P]
        L.orc_medium_densities.argtypes = [P, C.c_float, P, P]
        L.orc_glossy_eval.argtypes = [P, P, P, P, C.c_float, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Tree:
    """KDTree restatement built from vertices/triangles (KDTree::build(BEST))."""

    def __init__(self, verts, tris, abs_mode=0, _handle=None, _owner=None):
        self._verts = np.ascontiguousarray(verts, np.float32)
        self._tris = np.ascontiguousarray(tris, np.uint32)
        self._owner = _owner
        if _handle is not None:
            self.h = _handle
        else:
            self.h = lib().orc_tree_create(len(self._verts), _p(self._verts), len(self._tris), _p(self._tris),
                                           abs_mode)

    def __del__(self):
        if getattr(self, "h", None) and self._owner is None:
            lib().orc_tree_free(self.h)
            self.h = None

    def serialize(self):
        """KDTree::serializeForGPU arrays: nodes, min, max, v1, v2, v3 (each (n,4) float32)."""
        nn, nl = C.c_int(), C.c_int()
        lib().orc_tree_sizes(self.h, C.byref(nn), C.byref(nl))
        out = [np.zeros((nn.value, 4), np.float32) for _ in range(3)] + \
              [np.zeros((nl.value, 4), np.float32) for _ in range(3)]
        lib().orc_tree_serialize(self.h, *[_p(a) for a in out])
        return out

    def intersect(self, org, dirs, mode=0, tri_obj=None, tri_normal=None, inside=None, contact=None,
                  threads=0, skip=None):
        """mode 0 exhaustive recursion, 1 ordered (canonical counts), 2 brute force.
        skip: per-ray triangle id never hit (IntersectionGPU::Ray::last_tid), -1 = none."""
        org = np.ascontiguousarray(org, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(org)
        dist = np.empty(n, np.float32)
        tid = np.empty(n, np.int32)
        boxes = np.zeros(n, np.uint32)
        tris = np.zeros(n, np.uint32)
        cv = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
        tri_obj, tri_normal = cv(tri_obj, np.int32), cv(tri_normal, np.float32)
        inside, contact, skip = cv(inside, np.int32), cv(contact, np.int32), cv(skip, np.int32)
        lib().orc_intersect(self.h, mode, n, _p(org), _p(dirs), _p(tri_obj), _p(tri_normal), _p(inside),
                            _p(contact), _p(skip), _p(dist), _p(tid), _p(boxes), _p(tris), threads)
        return dist, tid, boxes, tris


def face_normals(verts, tris):
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    out = np.empty((len(tris), 3), np.float32)
    lib().orc_face_normals(len(verts), _p(verts), len(tris), _p(tris), _p(out))
    return out


def barycentrics(org, dirs, tid, verts, tris):
    org = np.ascontiguousarray(org, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    tid = np.ascontiguousarray(tid, np.int32)
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    uv = np.empty((len(org), 2), np.float64)
    lib().orc_barycentrics(len(org), _p(org), _p(dirs), _p(tid), _p(verts), _p(tris), _p(uv))
    return uv


class SceneOracle:
    """Restated PathTracer / PathTracerTest over a scene description."""

    def __init__(self, scene, abs_mode=0):
        self.scene = scene
        self._v = np.ascontiguousarray(scene.verts, np.float32)
        self._t = np.ascontiguousarray(scene.tris, np.uint32)
        self._o = np.ascontiguousarray(scene.objects)
        self._c = np.ascontiguousarray(scene.camera.reshape(1))
        self.h = lib().orc_scene_create(len(self._v), _p(self._v), len(self._t), _p(self._t), len(self._o),
                                        _p(self._o), _p(self._c), abs_mode)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None

    def set_camera(self, camera):
        """Render the same built scene through another camera (e.g. a smaller image)."""
        self._c = np.ascontiguousarray(camera.reshape(1))
        lib().orc_scene_set_camera(self.h, _p(self._c))
        import copy

        self.scene = copy.copy(self.scene)
        self.scene.camera = camera

    def tree(self) -> Tree:
        return Tree(self._v, self._t, _handle=lib().orc_scene_tree(self.h), _owner=self)

    def render(self, spp, max_depth=None, rr_start_depth=None, mode=0, seed=1, sample_offset=0, threads=0,
               stats=None, registered=False, density_check=False):
        """FP64 per-pixel (sum, sum of squares), shape (H, W, 3).  mode 0 is the
        PathTracer, 1 PathTracerTest, 3 the BidirectionalPathTracer; `registered` (mode 3) lands light-tracing contributions on
        the pixel that covers them instead of the reference's half-pixel-shifted
        one.  If `stats` is a dict it receives the number of traversal queries
        under "rays"; with `density_check` (mode 3) also the balance weights'
        density check (bpt_density_stats)."""
        cam = self.scene.camera
        H, W = int(cam["height"]), int(cam["width"])
        p = np.zeros((), RENDER_PARAMS)
        p["spp"] = spp
        p["max_depth"] = self.scene.max_depth if max_depth is None else max_depth
        p["rr_start_depth"] = self.scene.rr_start_depth if rr_start_depth is None else rr_start_depth
        p["mode"] = mode
        p["seed"] = seed
        p["sample_offset"] = sample_offset
        p["flags"] = (1 if registered else 0) | (2 if density_check else 0)
        p = p.reshape(1)
        s = np.zeros((H, W, 3), np.float64)
        q = np.zeros((H, W, 3), np.float64)
        rays = np.zeros(1, np.uint64)
        lib().orc_render(self.h, _p(p), _p(s), _p(q), threads, _p(rays))
        if stats is not None:
            stats["rays"] = int(rays[0])
            if density_check:
                stats.update(bpt_density_stats())
        return s, q


def bpt_density_stats():
    """The last density-checked bidirectional render: connections compared,
    those among them with a DEFINITE vertex ("skipped": that step's densities
    are left out on both sides), and the largest relative
    difference between the densities of the balance weight's sampled
    prefixes and the densities the subpaths were drawn with."""
    c, k, m = C.c_uint64(), C.c_uint64(), C.c_double()
    lib().orc_bpt_density_stats(C.byref(c), C.byref(k), C.byref(m))
    return {"checked": c.value, "skipped": k.value, "max_rel": m.value}


def mt_nth(n: int) -> int:
    return int(lib().orc_mt_nth(n))


def sampler_draws(kind: int, n: int, param: float, param2: float = 0.0, seed: int = 1):
    """kind 0 cosine lobe, 1 HG phase, 2 medium distance, 3 Fresnel er; (n,3) float32."""
    out = np.empty((n, 3), np.float32)
    lib().orc_sampler_draws(kind, n, param, param2, seed, _p(out))
    return out


# ---------------------------------------------------------------- single functions, for their pins
# Planted mistakes (see rx_oracle_scene.h): each must make one pin fail.
(MUT_NONE, MUT_ROTATION_SIGN, MUT_INVERTED_RATIO, MUT_P_SURFACE_COMPLEMENT, MUT_P_MEDIUM_NO_SIGMA, MUT_GLOSSY_AXIS,
 MUT_ORIGIN_PDF_SWAP) = range(7)


class mutation:
    """Context manager: the oracle computes with a planted mistake inside the block."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        lib().orc_set_mutation(self.m)

    def __exit__(self, *exc):
        lib().orc_set_mutation(0)


def refract_dir(d, n, n_i, n_t, kind=0):
    """The oracle's Snell/Fresnel step: kind 0 RefractiveMaterial, 1 SceneVPMObject.
    Returns (ok, direction (3,) float32, er); ok False = total internal reflection."""
    d = np.ascontiguousarray(d, np.float32)
    n = np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    er = np.zeros(1, np.float32)
    ok = lib().orc_refract_dir(kind, _p(d), _p(n), n_i, n_t, _p(out), _p(er))
    return bool(ok), out, float(er[0])


def medium_densities(sigma_t, x):
    """(p_medium(x), P_surface(x)) of a homogeneous medium with extinction sigma_t (3,)."""
    st = np.ascontiguousarray(sigma_t, np.float32)
    a = np.zeros(1, np.float32)
    b = np.zeros(1, np.float32)
    lib().orc_medium_densities(_p(st), x, _p(a), _p(b))
    return float(a[0]), float(b[0])


def glossy_eval(n, d_in, d_out, ks, exponent):
    """GlossyBSDF::evaluate restated: colour (3,) of the lobe for incoming d_in, outgoing d_out."""
    arrs = [np.ascontiguousarray(x, np.float32) for x in (n, d_in, d_out, ks)]
    out = np.zeros(3, np.float32)
    lib().orc_glossy_eval(*[_p(a) for a in arrs], exponent, _p(out))
    return out


# ---------------------------------------------------------------- guiding MLP (FP64)
MLP_IN, MLP_H, MLP_OUT = 96, 64, 48
MLP_SHAPES = [(96, 64), (64, 64), (64, 64), (64, 48)]
MLP_NW = sum(a * b for a, b in MLP_SHAPES)
MLP_NB = 64 * 3 + 48


def _mlp_sig():
    L = lib()
    if not getattr(L, "_mlp_ready", False):
        P = C.c_void_p
        L.orc_mlp_encode.argtypes = [C.c_int, P, P, P, P, P]
        L.orc_mlp_forward.argtypes = [C.c_int, P, P, P, P, P]
        L.orc_mlp_loss.argtypes = [C.c_int, P, P, P, P, P, P, P]
        L.orc_mlp_loss.restype = C.c_double
        L.orc_mlp_backward.argtypes = [C.c_int, P, P, P, P, P, P]
        L.orc_adam.argtypes = [C.c_int, P, P, P, P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L._mlp_ready = True
    return L


def mlp_encode(pos, dirs, lo, hi):
    L = _mlp_sig()
    pos = np.ascontiguousarray(pos, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    out = np.empty((len(pos), MLP_IN), np.float64)
    L.orc_mlp_encode(len(pos), _p(pos), _p(dirs), _p(lo), _p(hi), _p(out))
    return out


def mlp_forward(x, params):
    """params: flat float64 [W1,W2,W3,W4 row-major (in x out), b1..b4]."""
    L = _mlp_sig()
    x = np.ascontiguousarray(x, np.float64)
    W = np.ascontiguousarray(params[:MLP_NW], np.float64)
    b = np.ascontiguousarray(params[MLP_NW:], np.float64)
    z4 = np.empty((len(x), MLP_OUT), np.float64)
    acts = np.empty((len(x), 3, MLP_H), np.float64)
    L.orc_mlp_forward(len(x), _p(x), _p(W), _p(b), _p(z4), _p(acts))
    return z4, acts


def mlp_loss(z4, omega, pb, wt, want_grad=True):
    L = _mlp_sig()
    z4 = np.ascontiguousarray(z4, np.float64)
    omega = np.ascontiguousarray(omega, np.float32)
    pb = np.ascontiguousarray(pb, np.float32)
    wt = np.ascontiguousarray(wt, np.float32)
    g = np.empty_like(z4) if want_grad else None
    q = np.empty(len(z4), np.float64)
    a = np.empty(len(z4), np.float64)
    tot = L.orc_mlp_loss(len(z4), _p(z4), _p(omega), _p(pb), _p(wt), _p(g), _p(q), _p(a))
    return tot, g, q, a


def mlp_backward(x, params, acts, dz4):
    L = _mlp_sig()
    x = np.ascontiguousarray(x, np.float64)
    W = np.ascontiguousarray(params[:MLP_NW], np.float64)
    acts = np.ascontiguousarray(acts, np.float64)
    dz4 = np.ascontiguousarray(dz4, np.float64)
    dW = np.empty(MLP_NW, np.float64)
    db = np.empty(MLP_NB, np.float64)
    L.orc_mlp_backward(len(x), _p(x), _p(W), _p(acts), _p(dz4), _p(dW), _p(db))
    return np.concatenate([dW, db])


def adam(p, m, v, g, lr, b1, b2, eps, t):
    L = _mlp_sig()
    p, m, v = [np.ascontiguousarray(a, np.float64).copy() for a in (p, m, v)]
    g = np.ascontiguousarray(g, np.float64)
    L.orc_adam(len(p), _p(p), _p(m), _p(v), _p(g), lr, b1, b2, eps, t)
    return p, m, v
