"""GPU bidirectional tracer (RX_MODE_BPT, renderx_amd/csrc/rx_bpt.hip)
against the oracle's restated BidirectionalPathTracer
(src/BidirectionalPathTracer.cpp): every pixel mean within 3 sigma over the
pixel-channels whose estimates rest on at least three effective samples,
mean z near 0, image means within 4 standard errors; on the Cornell box and
on a Cornell box with a glass sphere (caustics through specular chains, light
subpaths refracting without the eye path's (n_i/n_t)^2).  With the
registered landing pixel the image is checked against a long GPU PathTracer
render instead, an estimator that shares none of the connection code.  On
the media scene and on a Cornell box inside a medium that holds the camera
and the light, the subpaths scatter in a homogeneous medium and the
connections take its transmittance, phase function and distance densities.
Plus
determinism, independence from the pass size, and the rejections."""
import math

import numpy as np
import pytest

from renderx_amd import scenes
from tests.stats_util import clt_mask, three_sigma

pytestmark = pytest.mark.gpu

MODE_PT, MODE_BPT = 0, 3
RX_RENDER_BPT_REGISTERED = 4
GPU_SPP = 1 << 18

# oracle samples per pixel: light-tracing contributions make the per-pixel
# distributions heavy-tailed, so the oracle's means need many samples before
# their errors are close to normal
CASES = {
    "cornell": (lambda: scenes.cornell(16, 16, spp=16, max_depth=8, rr_start_depth=0), 16384),
    "glass": (lambda: scenes.glass_scene(16, 16, max_depth=8), 8192),
    # the media scene (homogeneous medium filling the box, glass sphere, Phong
    # sphere): subpaths scatter in the medium, connections inside it carry its
    # transmittance, phase function and p_medium / P_surface densities
    "media": (lambda: scenes.media_scene(16, 16, max_depth=6, freq=12, glass_freq=6), 32768),
    # the Cornell box inside a medium that also holds the camera and the light
    "fog": (lambda: scenes.fog_room(16, 16), 16384),
}
# In the two media scenes light subpaths scatter anywhere in the volume, and
# scattering points close to the camera splat large light-traced values (the
# 1/d^2 of the camera connection); with boundary crossings and glass in the
# media scene as well, their per-pixel distributions are heavier-tailed than
# the surface scenes' (up to 3 % of the media scene's pixel-channels fail the
# CLT mask at these sample counts): bars of 98.5 % within 3 sigma and
# |mean z| < 0.15 there.
BARS = {"media": (0.985, 0.15), "fog": (0.985, 0.15)}
_oracle_cache = {}


def _oracle(oracle_mod, name):
    if name not in _oracle_cache:
        make, on = CASES[name]
        sc = make()
        s, q = oracle_mod.SceneOracle(sc).render(on, mode=3, seed=5)
        _oracle_cache[name] = (sc, s, q, on)
    return _oracle_cache[name]


def _check(name, gsum, gsq, gn, osum, osq, on, frac_min=0.995, zmax=0.1, pooled=True):
    """pooled: both renders sample the same estimator, so the per-pixel
    variance is estimated from both (dominated by the longer render)."""
    gm, om = gsum / gn, osum / on
    gv, ov = np.maximum(gsq / gn - gm * gm, 0), np.maximum(osq / on - om * om, 0)
    mask = clt_mask(gsum, gsq, osum, osq)
    frac, zmean, z = three_sigma(gsum, gsq, gn, osum, osq, on, pooled=pooled, mask=mask)
    se = math.sqrt(gv.sum() / gn + ov.sum() / on) / gm.size
    d = gm.mean() - om.mean()
    excl = 1.0 - mask.mean()
    print(f"  {name}: within 3 sigma {frac:.4f}, mean z {zmean:+.4f}, max |z| {np.abs(z[mask]).max():.2f}, "
          f"excluded {excl:.4f}, image mean gpu {gm.mean():.5f} reference {om.mean():.5f} ({d / se:+.2f} se)",
          flush=True)
    assert np.isfinite(gsum).all()
    assert excl <= 0.05
    assert frac >= frac_min
    assert abs(zmean) < zmax
    assert abs(d) < 4 * se


@pytest.mark.parametrize("name", list(CASES))
def test_image_matches_oracle_bpt(rx, oracle_mod, gpu_ctx, name):
    # a pixel's per-sample value includes rare large light-traced
    # contributions, so its mean's error is skewed at the oracle's sample
    # counts: 99 % of the pixel-channels within 3 sigma, with the mean z and
    # the image mean catching any systematic shift
    print(f"[bpt parity] {name}", flush=True)
    sc, osum, osq, on = _oracle(oracle_mod, name)
    r = rx.BidirectionalPathTracerHIP(sc, ctx=gpu_ctx).renderPixels(spp=GPU_SPP, seed=11)
    assert r["stats"]["samples"] == 16 * 16 * GPU_SPP
    frac_min, zmax = BARS.get(name, (0.99, 0.1))
    _check(name, r["sum"], r["sumsq"], GPU_SPP, osum, osq, on, frac_min=frac_min, zmax=zmax)


def test_registered_matches_gpu_pathtracer(rx, gpu_ctx):
    # Two differences of the reference's estimators are kept out of this
    # comparison.  No Phong object: the PathTracer's uniform lobe pick halves
    # a Phong surface's reflectance while the connections evaluate the lobe
    # sum (src/SceneObject.cpp:373-380 vs :349-359).  A high max_depth: each
    # subpath stops at max_depth bounces, so joined paths can be nearly twice
    # as long as the PathTracer's and the bidirectional image is brighter
    # wherever the depth limit cuts off light (through the glass at depth 8).
    print("[bpt registered] glass + diffuse sphere vs GPU PathTracer", flush=True)
    sc = scenes.glass_scene(16, 16, max_depth=20, phong=False)
    pn = 1 << 21
    ref = rx.PathTracerHIP(sc, ctx=gpu_ctx).renderPixels(spp=pn, mode=MODE_PT, seed=3)
    r = rx.BidirectionalPathTracerHIP(sc, ctx=gpu_ctx).renderPixels(spp=GPU_SPP, seed=4,
                                                                     flags=RX_RENDER_BPT_REGISTERED)
    _check("glass registered", r["sum"], r["sumsq"], GPU_SPP, ref["sum"], ref["sumsq"], pn, frac_min=0.99,
           pooled=False)


def test_deterministic_and_pass_size_invariant(rx, gpu_ctx):
    print("[bpt determinism]", flush=True)
    sc = scenes.glass_scene(16, 16, max_depth=8)
    pt = rx.BidirectionalPathTracerHIP(sc, ctx=gpu_ctx)
    a = pt.renderPixels(spp=64, seed=9)
    b = pt.renderPixels(spp=64, seed=9)
    pt.set_tuning(path_slots=4096)  # 8 samples per pass instead of 64
    c = pt.renderPixels(spp=64, seed=9)
    pt.set_tuning()
    d = pt.renderPixels(spp=64, seed=10)
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sumsq"], b["sumsq"])
    assert np.array_equal(a["sum"], c["sum"]) and np.array_equal(a["sumsq"], c["sumsq"])
    assert not np.array_equal(a["sum"], d["sum"])


def test_rejections(rx, gpu_ctx):
    sc = scenes.cornell(8, 8)
    with pytest.raises(rx.RxError, match="max_depth"):
        rx.BidirectionalPathTracerHIP(sc, ctx=gpu_ctx).renderPixels(spp=1, max_depth=31)
