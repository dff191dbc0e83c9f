"""Error paths of the C ABI on the device: a scene or tree load interrupted
by a failed device allocation (forced from the host with
rx_debug_fail_device_alloc) leaves the context without a scene and without a
tree, so rx_render and rx_intersect answer RX_ERR_STATE and launch nothing;
a malformed serialized tree is rejected and clears the previous one; a later
good load works and renders the same image as a fresh context."""
import numpy as np
import pytest

from renderx_amd import scenes

pytestmark = pytest.mark.gpu

RX_ERR_STATE, RX_ERR_NO_MEMORY, RX_ERR_INVALID = -4, -3, -1


def _render_rc(rx, ctx, sc):
    from renderx_amd._lib import lib, ptr

    cam = np.ascontiguousarray(sc.camera.reshape(1))
    p = np.zeros(1, rx.RENDER_PARAMS)
    p["spp"], p["max_depth"] = 2, 4
    mean = np.empty((int(cam["height"][0]), int(cam["width"][0]), 3), np.float32)
    return lib().rx_render(ctx.h, ptr(cam), ptr(p), ptr(mean), None, None, None)


def _intersect_rc(rx, ctx, want=None):
    """rx_intersect's status on a few rays through the scene; with `want`
    (tid array) a successful call must also return those hits."""
    from renderx_amd._lib import lib, ptr

    rng = np.random.default_rng(2)
    o = rng.uniform(0.1, 0.9, (256, 3)).astype(np.float32)
    d = rng.normal(size=(256, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    dist = np.empty(256, np.float32)
    tid = np.empty(256, np.int32)
    rc = lib().rx_intersect(ctx.h, 256, ptr(o), ptr(d), None, None, None, 0, ptr(dist), ptr(tid), None, None)
    if rc == 0 and want is not None:
        assert np.array_equal(tid, want)
    return rc, tid


def test_interrupted_loads_leave_no_half_built_state(rx):
    from renderx_amd._lib import lib, ptr

    print("[error paths] failed device allocations during rx_load_scene", flush=True)
    ctx = rx.Context(0)
    sc = scenes.media_scene(16, 16, freq=8, glass_freq=4)  # global tree + a medium tree + objects + emitters
    good = rx.Context(0)
    ref = rx.PathTracerHIP(sc, ctx=good).renderPixels(spp=16, mode=rx.MODE_NEE, seed=3)["sum"]
    _, want = _intersect_rc(rx, good)
    small = scenes.cornell(8, 8)
    arr = lambda s: (np.ascontiguousarray(s.verts, np.float32), np.ascontiguousarray(s.tris, np.uint32),
                     np.ascontiguousarray(s.objects))
    v, t, o = arr(sc)
    sv, st, so = arr(small)
    failures, tree_kept = 0, 0
    try:
        for k in range(64):  # make the k-th device allocation of the load fail, until a load gets through
            # a smaller scene first, so that every buffer of the big load must grow
            assert lib().rx_load_scene(ctx.h, len(sv), ptr(sv), len(st), ptr(st), len(so), ptr(so), 0) == 0
            lib().rx_debug_fail_device_alloc(k)
            rc = lib().rx_load_scene(ctx.h, len(v), ptr(v), len(t), ptr(t), len(o), ptr(o), 0)
            lib().rx_debug_fail_device_alloc(-1)
            if rc == 0:
                break
            failures += 1
            assert rc == RX_ERR_NO_MEMORY, (k, rc)
            assert _render_rc(rx, ctx, sc) == RX_ERR_STATE, k
            # the global tree is either absent or whole (a later stage failed): never half-built
            irc, _ = _intersect_rc(rx, ctx, want)
            assert irc in (0, RX_ERR_STATE), (k, irc)
            tree_kept += irc == 0
    finally:
        lib().rx_debug_fail_device_alloc(-1)
    print(f"  {failures} interrupted loads ({tree_kept} after the global tree was complete), each without a scene")
    assert failures >= 5 and 0 < tree_kept < failures
    pt = rx.PathTracerHIP(sc, ctx=ctx)
    assert np.array_equal(pt.renderPixels(spp=16, mode=rx.MODE_NEE, seed=3)["sum"], ref)


def test_malformed_tree_is_rejected_and_clears_the_old_one(rx, oracle_mod):
    from renderx_amd._lib import lib, ptr

    print("[error paths] malformed serialized trees", flush=True)
    ctx = rx.Context(0)
    sc = scenes.cornell(8, 8)
    arrs = [np.ascontiguousarray(a, np.float32) for a in oracle_mod.Tree(sc.verts, sc.tris, 0).serialize()]
    ig = rx.IntersectionGPU(ctx=ctx)
    ig.loadKDTreeToShader(arrs)
    assert _intersect_rc(rx, ctx)[0] == 0
    nodes = arrs[0].reshape(-1, 4).copy()
    interior = np.nonzero(nodes[:, 1] >= 0)[0]
    bad = nodes.copy()
    bad[interior[1] + 1, 0] = 0  # a left child whose parent pointer names another node
    rc = lib().rx_load_tree(ctx.h, len(bad), ptr(bad), ptr(arrs[1]), ptr(arrs[2]), len(arrs[3]),
                            ptr(arrs[3]), ptr(arrs[4]), ptr(arrs[5]))
    assert rc == RX_ERR_INVALID
    assert _intersect_rc(rx, ctx)[0] == RX_ERR_STATE
    ig.loadKDTreeToShader(arrs)
    assert _intersect_rc(rx, ctx)[0] == 0


def test_guided_record_index_limit_is_checked(rx):
    """Guided training ranks one record per pixel and bounce with 32-bit
    indices: a render whose pixels * (max_depth + 1) reaches 2^31 is rejected
    before anything is allocated."""
    from renderx_amd._lib import lib, ptr

    print("[error paths] guided record limit", flush=True)
    ctx = rx.Context(0)
    sc = scenes.cornell(8, 8)
    pt = rx.PathTracerHIP(sc, ctx=ctx)
    pt.guide_init(train_passes=1, batch_size=256, steps_per_pass=1)
    cam = np.ascontiguousarray(sc.camera.reshape(1)).copy()
    cam["width"], cam["height"] = 16384, 16384  # 2^28 pixels
    p = np.zeros(1, rx.RENDER_PARAMS)
    p["spp"], p["max_depth"], p["mode"] = 1, 7, rx.MODE_GUIDED
    assert lib().rx_render(ctx.h, ptr(cam), ptr(p), None, None, None, None) == RX_ERR_INVALID
    assert b"2^31" in lib().rx_last_error(ctx.h)
