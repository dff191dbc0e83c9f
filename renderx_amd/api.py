"""Python mirror of the reference's interface for the hot path.

  KDTree           -- KDTree::build / serializeForGPU (include/KDTree.h:10-59)
  IntersectionGPU  -- IntersectionGPU::init / loadKDTreeToShader / query
                      (include/IntersectionGPU.h:11-45), answered by HIP kernels
  Scene            -- Scene::buildKDTree + objects (include/Scene.h:10-107)
  PathTracerHIP    -- MCRenderer::renderPixels (include/MCRenderer.h:145-148)

All calls go through the C ABI of librxhip.so (include/rxhip.h).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import RxError, check, lib, ptr  # noqa: F401  (RxError re-exported)

ORDERED, EXHAUSTIVE, LEGALITY = 0, 1, 2

RENDER_PARAMS = np.dtype([("spp", "<u4"), ("max_depth", "<u4"), ("rr_start_depth", "<u4"), ("mode", "<u4"),
                          ("seed", "<u8"), ("sample_offset", "<u4"), ("flags", "<u4")])
RENDER_STATS = np.dtype([("rays", "<u8"), ("extend_rays", "<u8"), ("shadow_rays", "<u8"), ("segments", "<u8"),
                         ("samples", "<u8"), ("box_tests", "<u8"), ("tri_tests", "<u8"), ("ms_total", "<f8"),
                         ("ms_extend", "<f8"), ("ms_shade", "<f8"), ("ms_shadow", "<f8"),
                         ("extend_launches", "<u8"), ("guide_rows", "<u8"), ("guide_train_rows", "<u8"),
                         ("ms_extend_bounce", "<f8", (16,)), ("extend_rays_bounce", "<u8", (16,)),
                         ("box_tests_bounce", "<u8", (16,)), ("tri_tests_bounce", "<u8", (16,))])
RENDER_TIMING, RENDER_COUNT_WORK = 1, 2
RENDER_TUNING = np.dtype([("path_slots", "<u4"), ("blocks_per_cu", "<u4"), ("use_graph", "<u4"), ("stack_depth", "<u4"),
                          ("block_size", "<u4"), ("camera_rays", "<u4"), ("tail_bounce", "<u4")])
GUIDE_PARAMS = np.dtype([("train_passes", "<u4"), ("batch_size", "<u4"), ("steps_per_pass", "<u4"), ("learning_rate", "<f4"),
                         ("beta1", "<f4"), ("beta2", "<f4"), ("eps", "<f4"), ("seed", "<u4"), ("flags", "<u4"),
                         ("samples_per_pass", "<u4")])
MODE_PT, MODE_NEE, MODE_GUIDED, MODE_BPT = 0, 1, 2, 3


def _tuning(path_slots, blocks_per_cu, use_graph, stack_depth, block_size, camera_rays=0, tail_bounce=0):
    t = np.zeros(1, RENDER_TUNING)
    t["path_slots"], t["blocks_per_cu"], t["use_graph"] = path_slots, blocks_per_cu, use_graph
    t["stack_depth"], t["block_size"], t["camera_rays"] = stack_depth, block_size, camera_rays
    t["tail_bounce"] = tail_bounce
    return t
GUIDE_LR_DECAY = 1
GUIDE_KEEP_SAMPLES = 2


class KDTree:
    """Host KD-tree with the reference's median split (KDTree::build(BEST))."""

    def __init__(self, vertexPositionList, triangleList, abs_int: bool = False):
        self.verts = np.ascontiguousarray(vertexPositionList, np.float32).reshape(-1, 3)
        self.tris = np.ascontiguousarray(triangleList, np.uint32).reshape(-1, 3)
        h = C.c_void_p()
        check(None, lib().rx_tree_build(len(self.verts), ptr(self.verts), len(self.tris), ptr(self.tris),
                                        1 if abs_int else 0, C.byref(h)), "rx_tree_build")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().rx_tree_free(self.h)
            self.h = None

    def info(self):
        n, m, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(None, lib().rx_tree_info(self.h, C.byref(n), C.byref(m), C.byref(d)), "rx_tree_info")
        return n.value, m.value, d.value

    def serializeForGPU(self):
        n, m, _ = self.info()
        out = [np.zeros((n, 4), np.float32) for _ in range(3)] + [np.zeros((m, 4), np.float32) for _ in range(3)]
        check(None, lib().rx_tree_serialize(self.h, *[ptr(a) for a in out]), "rx_tree_serialize")
        return out


class Context:
    """One device + stream (rx_create)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(None, lib().rx_create(device, C.byref(h)), "rx_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().rx_destroy(self.h)
            self.h = None

    def check(self, code, what):
        check(self.h, code, what)


class IntersectionGPU:
    """Drop-in for the reference's GL intersection seam, on HIP."""

    def __init__(self, device: int = 0, ctx: Context | None = None):
        self.ctx = ctx or Context(device)

    def loadKDTreeToShader(self, tree_or_arrays):
        arrs = tree_or_arrays.serializeForGPU() if isinstance(tree_or_arrays, KDTree) else tree_or_arrays
        nodes, mn, mx, v1, v2, v3 = [np.ascontiguousarray(a, np.float32) for a in arrs]
        self.ctx.check(lib().rx_load_tree(self.ctx.h, len(nodes), ptr(nodes), ptr(mn), ptr(mx), len(v1), ptr(v1),
                                          ptr(v2), ptr(v3)), "rx_load_tree")

    def load_legality(self, tri_obj, tri_normal=None):
        tri_obj = np.ascontiguousarray(tri_obj, np.int32)
        tn = None if tri_normal is None else np.ascontiguousarray(tri_normal, np.float32)
        self.ctx.check(lib().rx_load_legality(self.ctx.h, len(tri_obj), ptr(tri_obj), ptr(tn)), "rx_load_legality")

    def query(self, org, dirs, inside=None, contact=None, exhaustive=False, bary=False, counts=False,
              last_tid=None):
        """IntersectionGPU::query: nearest hit per ray.  last_tid (per ray, -1 =
        none) is the triangle the ray never hits, IntersectionGPU::Ray::last_tid."""
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(org)
        flags = (EXHAUSTIVE if exhaustive else ORDERED)
        ins = con = None
        if inside is not None:
            flags |= LEGALITY
            ins = np.ascontiguousarray(inside, np.int32)
            con = np.ascontiguousarray(contact, np.int32)
        dist = np.empty(n, np.float32)
        tid = np.empty(n, np.int32)
        b = np.empty((n, 2), np.float32) if bary else None
        c = np.empty((n, 2), np.uint32) if counts else None
        skip = None if last_tid is None else np.ascontiguousarray(last_tid, np.int32)
        self.ctx.check(lib().rx_intersect(self.ctx.h, n, ptr(org), ptr(dirs), ptr(ins), ptr(con), ptr(skip), flags,
                                          ptr(dist), ptr(tid), ptr(b), ptr(c)), "rx_intersect")
        out = {"dist": dist, "tid": tid}
        if bary:
            out["bary"] = b
        if counts:
            out["counts"] = c
        return out

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        self.ctx.check(lib().rx_last_kernel_ms(self.ctx.h, C.byref(ms)), "rx_last_kernel_ms")
        return ms.value


class PathTracerHIP:
    """MCRenderer plugin: renderPixels(camera) on the wavefront HIP renderer."""

    def __init__(self, scene, device: int = 0, ctx: Context | None = None, tree_abs_int: bool = False):
        self.ctx = ctx or Context(device)
        self.scene = scene
        v = np.ascontiguousarray(scene.verts, np.float32)
        t = np.ascontiguousarray(scene.tris, np.uint32)
        o = np.ascontiguousarray(scene.objects)
        self.ctx.check(lib().rx_load_scene(self.ctx.h, len(v), ptr(v), len(t), ptr(t), len(o), ptr(o),
                                           1 if tree_abs_int else 0), "rx_load_scene")

    def guide_init(self, train_passes=16, batch_size=4096, steps_per_pass=16, learning_rate=1e-2, seed=7, flags=0,
                   samples_per_pass=0):
        """rx_guide_init: the network and its online warm-up (flags: GUIDE_LR_DECAY).  samples_per_pass:
        radiance samples each training pass keeps (0: batch_size * steps_per_pass); more steps than
        samples_per_pass / batch_size go over them again."""
        p = np.zeros(1, GUIDE_PARAMS)
        p["flags"] = flags
        p["samples_per_pass"] = samples_per_pass
        p["train_passes"] = train_passes
        p["batch_size"] = batch_size
        p["steps_per_pass"] = steps_per_pass
        p["learning_rate"] = learning_rate
        p["beta1"], p["beta2"], p["eps"] = 0.9, 0.999, 1e-8
        p["seed"] = seed
        self.ctx.check(lib().rx_guide_init(self.ctx.h, ptr(p)), "rx_guide_init")

    def guide_weights(self):
        """rx_guide_get_weights: the network's FP32 master parameters."""
        w = np.zeros(_lib.GUIDE_PARAMS, np.float32)
        self.ctx.check(lib().rx_guide_get_weights(self.ctx.h, ptr(w)), "rx_guide_get_weights")
        return w

    def set_tuning(self, path_slots=0, blocks_per_cu=0, use_graph=0, stack_depth=0, block_size=0, camera_rays=0,
                   tail_bounce=0):
        """Launch shape of renderPixels (rx_render_tuning); images do not depend on it.
        tail_bounce: first bounce of the one-launch tail of RX_MODE_PT (0 the
        default, above max_depth none)."""
        self.ctx.check(lib().rx_set_render_tuning(self.ctx.h, ptr(_tuning(path_slots, blocks_per_cu, use_graph,
                                                                          stack_depth, block_size, camera_rays,
                                                                          tail_bounce))),
                       "rx_set_render_tuning")

    def get_tuning(self):
        t = np.zeros(1, RENDER_TUNING)
        self.ctx.check(lib().rx_get_render_tuning(self.ctx.h, ptr(t)), "rx_get_render_tuning")
        return {k: int(t[k][0]) for k in RENDER_TUNING.names}

    def autotune(self, spp=None, max_depth=None, rr_start_depth=None, mode=MODE_PT, seed=1):
        """rx_autotune_render: time the candidate launch shapes, keep the fastest."""
        cam = np.ascontiguousarray(self.scene.camera.reshape(1))
        p = np.zeros(1, RENDER_PARAMS)
        p["spp"] = self.scene.spp if spp is None else spp
        p["max_depth"] = self.scene.max_depth if max_depth is None else max_depth
        p["rr_start_depth"] = self.scene.rr_start_depth if rr_start_depth is None else rr_start_depth
        p["mode"], p["seed"] = mode, seed
        t = np.zeros(1, RENDER_TUNING)
        ms = np.zeros(1, np.float32)
        self.ctx.check(lib().rx_autotune_render(self.ctx.h, ptr(cam), ptr(p), ptr(t), ptr(ms)), "rx_autotune_render")
        return {k: int(t[k][0]) for k in RENDER_TUNING.names}, float(ms[0])

    def renderPixels(self, camera=None, spp=None, max_depth=None, rr_start_depth=None, mode=MODE_PT, seed=1,
                     sample_offset=0, want_sums=True, flags=0, want_mean=True):
        """Returns dict(mean=(H,W,3) f32, sum, sumsq=(H,W,3) f64, stats).  With want_sums and want_mean
        false the film stays on the device (no host transfer; mean, sum and sumsq are None)."""
        cam = np.ascontiguousarray((self.scene.camera if camera is None else camera).reshape(1))
        H, W = int(cam["height"][0]), int(cam["width"][0])
        p = np.zeros(1, RENDER_PARAMS)
        p["spp"] = self.scene.spp if spp is None else spp
        p["max_depth"] = self.scene.max_depth if max_depth is None else max_depth
        p["rr_start_depth"] = self.scene.rr_start_depth if rr_start_depth is None else rr_start_depth
        p["mode"] = mode
        p["seed"] = seed
        p["sample_offset"] = sample_offset
        p["flags"] = flags
        mean = np.empty((H, W, 3), np.float32) if want_mean else None
        s = np.empty((H, W, 3), np.float64) if want_sums else None
        q = np.empty((H, W, 3), np.float64) if want_sums else None
        st = np.zeros(1, RENDER_STATS)
        self.ctx.check(lib().rx_render(self.ctx.h, ptr(cam), ptr(p), ptr(mean), ptr(s), ptr(q), ptr(st)),
                       "rx_render")
        return {"mean": mean, "sum": s, "sumsq": q, "stats": {k: st[k][0].tolist() if st[k].ndim > 1 else st[k][0].item() for k in RENDER_STATS.names}}


class BidirectionalPathTracerHIP(PathTracerHIP):
    """MCRenderer plugin "BPT" (src/BidirectionalPathTracer.cpp): renderPixels
    renders with RX_MODE_BPT unless another mode is asked for."""

    def renderPixels(self, camera=None, spp=None, max_depth=None, rr_start_depth=None, mode=MODE_BPT, seed=1,
                     sample_offset=0, want_sums=True, flags=0):
        return super().renderPixels(camera, spp, max_depth, rr_start_depth, mode, seed, sample_offset, want_sums,
                                    flags)


class BatchRenderer:
    """Many scenes rendered as one wavefront job (rx_batch_*): every scene's
    image is identical to PathTracerHIP.renderPixels of that scene alone."""

    def __init__(self, device: int = 0, ctx: Context | None = None):
        self.ctx = ctx or Context(device)
        h = C.c_void_p()
        self.ctx.check(lib().rx_batch_create(self.ctx.h, C.byref(h)), "rx_batch_create")
        self.h = h
        self.shapes = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().rx_batch_destroy(self.h)
            self.h = None

    def add(self, scene, spp=None, max_depth=None, rr_start_depth=None, mode=MODE_PT, seed=1, sample_offset=0,
            camera=None, tree_abs_int=False) -> int:
        v = np.ascontiguousarray(scene.verts, np.float32)
        t = np.ascontiguousarray(scene.tris, np.uint32)
        o = np.ascontiguousarray(scene.objects)
        cam = np.ascontiguousarray((scene.camera if camera is None else camera).reshape(1))
        p = np.zeros(1, RENDER_PARAMS)
        p["spp"] = scene.spp if spp is None else spp
        p["max_depth"] = scene.max_depth if max_depth is None else max_depth
        p["rr_start_depth"] = scene.rr_start_depth if rr_start_depth is None else rr_start_depth
        p["mode"], p["seed"], p["sample_offset"] = mode, seed, sample_offset
        idx = C.c_uint32()
        self.ctx.check(lib().rx_batch_add_scene(self.h, len(v), ptr(v), len(t), ptr(t), len(o), ptr(o),
                                                1 if tree_abs_int else 0, ptr(cam), ptr(p), C.byref(idx)),
                       "rx_batch_add_scene")
        self.shapes.append((int(cam["height"][0]), int(cam["width"][0])))
        return int(idx.value)

    def set_tuning(self, path_slots=0, blocks_per_cu=0, use_graph=0, stack_depth=0, block_size=0, camera_rays=0):
        """Launch shape (rx_render_tuning of the context); images do not depend on it.
        use_graph: 0 stream launches, 1 one pass recorded in a hipGraph and
        replayed, 2 the whole batch as one persistent cooperative launch;
        camera_rays: 0 default, 1 one per lane, 2 packets of 64."""
        self.ctx.check(lib().rx_set_render_tuning(self.ctx.h, ptr(_tuning(path_slots, blocks_per_cu, use_graph,
                                                                          stack_depth, block_size, camera_rays))),
                       "rx_set_render_tuning")

    def autotune(self, passes=2):
        """rx_batch_autotune: times the first `passes` passes of the batch under
        each candidate launch shape (path slots, blocks per CU, extend stack
        depth and block size, graph replay or one persistent launch) and keeps
        the fastest as the context's tuning.  Returns (tuning, ms)."""
        t = np.zeros(1, RENDER_TUNING)
        ms = np.zeros(1, np.float32)
        self.ctx.check(lib().rx_batch_autotune(self.h, passes, ptr(t), ptr(ms)), "rx_batch_autotune")
        return {k: int(t[k][0]) for k in RENDER_TUNING.names}, float(ms[0])

    def render(self):
        st = np.zeros(1, RENDER_STATS)
        self.ctx.check(lib().rx_batch_render(self.h, ptr(st)), "rx_batch_render")
        return {k: st[k][0].tolist() if st[k].ndim > 1 else st[k][0].item() for k in RENDER_STATS.names}

    def image(self, k):
        H, W = self.shapes[k]
        mean = np.empty((H, W, 3), np.float32)
        s = np.empty((H, W, 3), np.float64)
        q = np.empty((H, W, 3), np.float64)
        self.ctx.check(lib().rx_batch_get_image(self.h, k, ptr(mean), ptr(s), ptr(q)), "rx_batch_get_image")
        return {"mean": mean, "sum": s, "sumsq": q}


def load_scene_config(config_path: str):
    """ConfigManager::load for this path (rx_scene_load): the Config XML, its
    Scene XML and the OBJ/SHP shapes.  Returns a scenes.Scene (spp, max depth
    from the file) with meta {"mode", "save_path", "runtime_s"}."""
    import os

    from . import scenes

    h = C.c_void_p()
    err = C.create_string_buffer(1024)
    if lib().rx_scene_load(os.fsencode(config_path), C.byref(h), err, len(err)) != 0:
        raise RxError(f"rx_scene_load: {err.value.decode(errors='replace')}")
    try:
        nv, nt, no = C.c_uint32(), C.c_uint32(), C.c_uint32()
        lib().rx_scene_file_sizes(h, C.byref(nv), C.byref(nt), C.byref(no))
        verts = np.empty((nv.value, 3), np.float32)
        tris = np.empty((nt.value, 3), np.uint32)
        objs = np.zeros(no.value, scenes.OBJ_DTYPE)
        cam = np.zeros((), scenes.CAM_DTYPE)
        prm = np.zeros(1, RENDER_PARAMS)
        save = C.create_string_buffer(4096)
        rt = C.c_double()
        check(None, lib().rx_scene_file_get(h, ptr(verts), ptr(tris), ptr(objs), cam.ctypes.data_as(C.c_void_p),
                                            ptr(prm), save, len(save), C.byref(rt)), "rx_scene_file_get")
    finally:
        lib().rx_scene_file_free(h)
    name = os.path.splitext(os.path.basename(config_path))[0]
    return scenes.Scene(name, verts, tris, objs, cam, int(prm["spp"][0]), int(prm["max_depth"][0]), 0,
                        {"mode": int(prm["mode"][0]), "save_path": save.value.decode(errors="replace"),
                         "runtime_s": rt.value})
