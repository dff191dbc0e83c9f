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


@pytest.mark.parametrize("mode", [MODE_PT, MODE_NEE])
@pytest.mark.parametrize("graph", [0, 1])
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
        for slots in (0, 4096):  # 4096 slots split the larger scenes over several passes
            pt.set_tuning(path_slots=slots, use_graph=graph)
            st = br.render()
            assert st["samples"] == sum(sc.camera["width"] * sc.camera["height"] * sc.spp for sc in scs)
            for k in range(len(scs)):
                img = br.image(k)
                a = single[k]
                assert np.array_equal(img["sum"], a["sum"]), (k, slots)
                assert np.array_equal(img["sumsq"], a["sumsq"]), (k, slots)
                assert np.array_equal(img["mean"], a["mean"]), (k, slots)
            print(f"  slots={slots}: {len(scs)} scenes identical, {st['rays']} rays in {st['ms_total']:.2f} ms")
    finally:
        pt.set_tuning()


def test_batch_rejects_mixed_modes_and_media(rx, gpu_ctx):
    br = rx.BatchRenderer(ctx=gpu_ctx)
    br.add(_small(scenes.cornell(16, 16), 16, 16, 1), mode=MODE_PT)
    with pytest.raises(rx.RxError):
        br.add(_small(scenes.cornell(16, 16), 16, 16, 1), mode=MODE_NEE)
    with pytest.raises(rx.RxError):
        br.add(_small(scenes.media_scene(16, 16, freq=8, glass_freq=4), 16, 16, 1), mode=MODE_PT)
