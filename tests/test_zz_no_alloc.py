"""rx_intersect_dev_noalloc (the traversal entry point that allocates no
device memory: its scratch is a fixed device array of the library) against
rx_intersect_dev on the same device buffers, through the C ABI: identical
hit ids, distance bits, barycentrics and visit counts, on two calls in a row
with the second smaller than the first, with the legality condition and the
last_tid skip, and an overflow-free rx_synchronize afterwards."""
import ctypes as C

import numpy as np
import pytest

from renderx_amd import scenes

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers from the HIP runtime librxhip.so itself links (the test
    allocates its inputs and outputs; the entry point under test does not)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.bufs.append(p)
        return p

    def empty(self, shape, dtype):
        return self.put(np.zeros(shape, dtype))

    def get(self, p, shape, dtype):
        a = np.empty(shape, dtype)
        assert self.h.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _call(lib, ctx, name, hip, n, o, d, ins, con, skip, flags):
    dist = hip.empty(n, np.float32)
    tid = hip.empty(n, np.int32)
    bary = hip.empty((n, 2), np.float32)
    cnt = hip.empty((n, 2), np.uint32)
    rc = getattr(lib, name)(ctx.h, n, o, d, ins, con, skip, flags, dist, tid, bary, cnt)
    assert rc == 0, (name, rc, lib.rx_last_error(ctx.h))
    assert lib.rx_synchronize(ctx.h) == 0, lib.rx_last_error(ctx.h)
    return [hip.get(dist, n, np.float32), hip.get(tid, n, np.int32), hip.get(bary, (n, 2), np.float32),
            hip.get(cnt, (n, 2), np.uint32)]


def test_no_alloc_entry_matches_rx_intersect_dev(rx):
    from renderx_amd._lib import lib

    print("[no-alloc] rx_intersect_dev_noalloc vs rx_intersect_dev", flush=True)
    ctx = rx.Context(0)
    sc = scenes.media_scene(8, 8, freq=40, glass_freq=8)
    ig = rx.IntersectionGPU(ctx=ctx)
    ig.loadKDTreeToShader(rx.KDTree(sc.verts, sc.tris))
    ig.load_legality(sc.tri_obj())
    hip = _Hip()
    rng = np.random.default_rng(3)
    try:
        for n, flags in [(300000, rx.ORDERED), (7001, rx.LEGALITY), (40000, rx.EXHAUSTIVE | rx.LEGALITY)]:
            print(f"  n={n} flags={flags}", flush=True)
            o = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
            d = rng.normal(size=(n, 3))
            d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
            ins = rng.integers(-1, len(sc.objects), n).astype(np.int32)
            con = rng.integers(-1, len(sc.objects), n).astype(np.int32)
            skip = np.where(rng.random(n) < 0.5, rng.integers(0, sc.n_tris, n), -1).astype(np.int32)
            args = [hip.put(x) for x in (o, d, ins, con, skip)]
            ref = _call(lib(), ctx, "rx_intersect_dev", hip, n, *args, flags)
            got = _call(lib(), ctx, "rx_intersect_dev_noalloc", hip, n, *args, flags)
            assert np.array_equal(ref[1], got[1]) and np.array_equal(ref[0].view(np.uint32), got[0].view(np.uint32))
            hit = ref[1] >= 0
            assert np.array_equal(ref[2][hit], got[2][hit]) and np.array_equal(ref[3], got[3])
            assert hit.mean() > 0.3
            hip.free()
    finally:
        hip.free()
