"""Multi-scene batch renderer (rx_batch_*): every scene's film is bit-identical
to rendering that scene alone with rx_render, whether the passes run eagerly
or replay one recorded hipGraph, and whatever the path-slot budget (which
splits scenes across passes)."""
import math

import numpy as np
import pytest

from renderx_amd import scenes

pytestmark = pytest.mark.gpu

MODE_PT, MODE_NEE = 0, 1


def _small(sc, w, h, spp, max_depth=None):
    cam = sc.camera.copy()
    cam["width"], cam["height"] = w, h
    cam["sight_dist"] = int(round((h / 2) / math.tan(math.radians(20.0))))
    sc.camera = cam
    sc.spp = spp
    if max_depth is not None:
        sc.max_depth = max_depth
    return sc


def _scene_set():
    out = [_small(scenes.cornell(32, 32), 32, 32, 5, max_depth=8)]
    for k, (w, h, spp) in enumerate([(40, 24, 3), (16, 16, 7), (56, 40, 2)]):
        out.append(_small(scenes.soup_scene(k, n_tris=4000 + 3000 * k), w, h, spp))
    out.append(_small(scenes.icosphere_scene(24, 24, freq=12), 24, 24, 4, max_depth=10))
    return out


# (path slots, extend stack depth, extend block size, camera rays: 0 the
# batch default (one per lane), 2 in packets); 4096 slots split the larger
# scenes over several passes
SHAPES = [(0, 0, 0, 0), (4096, 0, 0, 0), (0, 8, 128, 0), (4096, 16, 128, 0), (0, 10, 256, 0), (0, 10, 64, 0),
          (0, 0, 0, 2)]


@pytest.mark.parametrize("mode", [MODE_PT, MODE_NEE])
@pytest.mark.parametrize("graph", [0, 1, 2])
def test_batch_images_equal_single_renders(rx, gpu_ctx, mode, graph):
    print(f"[batch] mode={mode} graph={graph}", flush=True)
    scs = _scene_set()
    single = []
    for k, sc in enumerate(scs):
        pt = rx.PathTracerHIP(sc, ctx=gpu_ctx)
        single.append(pt.renderPixels(mode=mode, seed=100 + k))
    br = rx.BatchRenderer(ctx=gpu_ctx)
    for k, sc in enumerate(scs):
        assert br.add(sc, mode=mode, seed=100 + k) == k
    try:
        for slots, depth, block, lanes in SHAPES:
            if graph == 2 and (depth or block or lanes):
                continue  # the persistent launch has the default extend shape
            pt.set_tuning(path_slots=slots, use_graph=graph, stack_depth=depth, block_size=block, camera_rays=lanes)
            st = br.render()
            assert st["samples"] == sum(sc.camera["width"] * sc.camera["height"] * sc.spp for sc in scs)
            for k in range(len(scs)):
                img = br.image(k)
                a = single[k]
                assert np.array_equal(img["sum"], a["sum"]), (k, slots)
                assert np.array_equal(img["sumsq"], a["sumsq"]), (k, slots)
                assert np.array_equal(img["mean"], a["mean"]), (k, slots)
            print(f"  slots={slots} depth={depth} block={block} lanes={lanes}: {len(scs)} scenes identical, {st['rays']} rays "
                  f"in {st['ms_total']:.2f} ms")
    finally:
        pt.set_tuning()


def test_batch_autotune(rx, gpu_ctx):
    """rx_batch_autotune times the candidate shapes on the first passes, keeps
    one that the context accepts, and the films stay bit-identical to the
    single renders under it."""
    print("[batch] autotune", flush=True)
    scs = _scene_set()
    single = [rx.PathTracerHIP(sc, ctx=gpu_ctx).renderPixels(mode=MODE_NEE, seed=100 + k) for k, sc in enumerate(scs)]
    br = rx.BatchRenderer(ctx=gpu_ctx)
    for k, sc in enumerate(scs):
        br.add(sc, mode=MODE_NEE, seed=100 + k)
    try:
        t, ms = br.autotune(passes=2)
        print(f"  chose {t}, {ms * 1e6:.2f} ns per camera path")
        assert ms > 0 and t["use_graph"] in (0, 1, 2)
        br.render()
        for k in range(len(scs)):
            assert np.array_equal(br.image(k)["sum"], single[k]["sum"]), k
    finally:
        br.set_tuning()


def test_batch_rejects_mixed_modes_and_media(rx, gpu_ctx):
    br = rx.BatchRenderer(ctx=gpu_ctx)
    br.add(_small(scenes.cornell(16, 16), 16, 16, 1), mode=MODE_PT)
    with pytest.raises(rx.RxError):
        br.add(_small(scenes.cornell(16, 16), 16, 16, 1), mode=MODE_NEE)
    with pytest.raises(rx.RxError):
        br.add(_small(scenes.media_scene(16, 16, freq=8, glass_freq=4), 16, 16, 1), mode=MODE_PT)


# Eight scenes of the 64-scene workload at their true triangle counts (11 k to
# 1.83 M, 4.0 M in all), at reduced resolution and spp
REAL_SOUPS = [51, 26, 60, 18, 21, 2, 61, 13]


def test_batch_of_real_soup_scenes(rx, oracle_mod, gpu_ctx):
    """The batch renderer on scenes of the 64-scene workload at their true
    sizes: every film bit-identical to the scene rendered alone, and the
    smallest one within 3 sigma of the oracle's PathTracer."""
    from tests.stats_util import clt_mask, three_sigma

    print("[batch] 8 real soup scenes", flush=True)
    scs = [_small(scenes.soup_scene(i), 24, 24, 4) for i in REAL_SOUPS]
    print(f"  triangles: {[sc.n_tris for sc in scs]}", flush=True)
    br = rx.BatchRenderer(ctx=gpu_ctx)
    for k, sc in enumerate(scs):
        br.add(sc, mode=MODE_PT, seed=500 + k)
    try:
        br.set_tuning(use_graph=1)
        st = br.render()
    finally:
        br.set_tuning()
    assert st["samples"] == sum(24 * 24 * 4 for _ in scs)
    for k, sc in enumerate(scs):
        a = rx.PathTracerHIP(sc, ctx=gpu_ctx).renderPixels(mode=MODE_PT, seed=500 + k)
        img = br.image(k)
        assert np.array_equal(img["sum"], a["sum"]) and np.array_equal(img["sumsq"], a["sumsq"]), k
    print(f"  batch {st['ms_total']:.1f} ms, {st['rays']} rays; all films identical", flush=True)
    # the smallest scene against the oracle
    sc = _small(scenes.soup_scene(REAL_SOUPS[0]), 16, 16, 1)
    n = 32768
    g = rx.PathTracerHIP(sc, ctx=gpu_ctx).renderPixels(spp=n, mode=MODE_PT, seed=7)
    osum, osq = oracle_mod.SceneOracle(sc).render(n, seed=8)
    mask = clt_mask(g["sum"], g["sumsq"], osum, osq)
    frac, zmean, _ = three_sigma(g["sum"], g["sumsq"], n, osum, osq, n, True, mask=mask)
    print(f"  soup {REAL_SOUPS[0]} ({sc.n_tris} tris) vs oracle: within 3 sigma {frac:.4f}, mean z {zmean:+.3f}, "
          f"excluded {1 - mask.mean():.4f}")
    # a sparse soup lit by a small light: many pixels see the light on few
    # paths, so their channels rest on too few effective samples for a z-test
    # (clt_mask); the image mean, over every pixel, must agree as well
    gm, om = g["sum"] / n, osum / n
    se = np.sqrt((np.maximum(g["sumsq"] / n - gm * gm, 0).sum() + np.maximum(osq / n - om * om, 0).sum()) / n) / gm.size
    print(f"  image mean gpu {gm.mean():.5f} oracle {om.mean():.5f} ({(gm.mean() - om.mean()) / se:+.2f} se)")
    assert mask.mean() >= 0.85 and frac >= 0.995 and abs(zmean) < 0.1
    assert abs(gm.mean() - om.mean()) < 4 * se
