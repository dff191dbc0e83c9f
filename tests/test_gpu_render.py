"""GPU wavefront renderer against the oracle's restated PathTracer
(src/PathTracer.cpp:17-89): every pixel mean within 3 sigma (pass fraction
>= 99.5 % of the pixel-channels whose estimates rest on at least three
effective samples, mean z near 0, image means within 4 standard errors),
plus determinism, resumable sample counters and the closed-form furnace
sphere."""
import math

import numpy as np
import pytest

from renderx_amd import scenes
from tests.stats_util import closed_form_z, clt_mask, three_sigma

pytestmark = pytest.mark.gpu

MODE_PT, MODE_NEE = 0, 1
RES = 24
GPU_SPP = 65536


def _check(name, gsum, gsq, gn, osum, osq, on, frac_min=0.995, zmax=0.1, pooled=True, ovar=None, extra=()):
    """Per pixel-channel z-test over the channels where the normal
    approximation holds (clt_mask; at most 5 % may be excluded), plus the
    image means within 4 standard errors.  ovar: a better estimate of the
    oracle estimator's per-pixel variance (e.g. pooled with a long GPU render
    of the same estimator, whose sums go in `extra`)."""
    gm, om = gsum / gn, osum / on
    gv = np.maximum(gsq / gn - gm * gm, 0)
    ov = np.maximum(osq / on - om * om, 0) if ovar is None else ovar
    mask = clt_mask(gsum, gsq, osum, osq, *extra)
    if ovar is None:
        frac, zmean, z = three_sigma(gsum, gsq, gn, osum, osq, on, pooled, mask=mask)
    else:
        se_p = np.sqrt(gv / gn + ov / on)
        z = np.where(se_p > 0, (gm - om) / np.where(se_p > 0, se_p, 1), np.where(gm == om, 0.0, np.inf))
        zs = z[mask]
        frac, zmean = float((np.abs(zs) <= 3).mean()), float(np.clip(zs, -5, 5).mean())
    se = math.sqrt(gv.sum() / gn + ov.sum() / on) / gm.size
    d = gm.mean() - om.mean()
    excl = 1.0 - mask.mean()
    print(f"  {name}: within 3 sigma {frac:.4f}, mean z {zmean:+.4f}, max |z| {np.abs(z[mask]).max():.2f}, "
          f"excluded {excl:.4f}, image mean gpu {gm.mean():.5f} oracle {om.mean():.5f} ({d / se:+.2f} se)",
          flush=True)
    assert np.isfinite(gsum).all()
    assert excl <= 0.05
    assert frac >= frac_min
    assert abs(zmean) < zmax
    assert abs(d) < 4 * se


def _render(rx, ctx, scene, spp, mode, seed=11, **kw):
    pt = rx.PathTracerHIP(scene, ctx=ctx)
    return pt.renderPixels(spp=spp, mode=mode, seed=seed, **kw)


CASES = {
    "cornell": (lambda: scenes.cornell(RES, RES, spp=16, max_depth=8, rr_start_depth=3), 16384),
    "media": (lambda: scenes.media_scene(RES, RES, spp=16, max_depth=16, freq=24, glass_freq=8), 32768),
    # a grey scene: the three channels of a pixel fail together, so more pixels
    "icosphere": (lambda: scenes.icosphere_scene(32, 32, spp=16, max_depth=10, freq=24), 32768),
    # the participating-media workload at its real geometry (499,280-triangle
    # displaced sphere, 5,120-triangle glass sphere), at reduced resolution
    "media_full": (lambda: scenes.media_scene(16, 16, spp=16, max_depth=16, freq=158, glass_freq=16), 16384),
}
_oracle_cache = {}
_gpu_pt_cache = {}


def _oracle(oracle_mod, name):
    if name not in _oracle_cache:
        make, on = CASES[name]
        sc = make()
        s, q = oracle_mod.SceneOracle(sc).render(on, seed=5)
        _oracle_cache[name] = (sc, s, q, on)
    return _oracle_cache[name]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("mode", [MODE_PT, MODE_NEE])
def test_image_matches_oracle_pathtracer(rx, oracle_mod, gpu_ctx, name, mode):
    print(f"[render parity] {name} mode={mode}", flush=True)
    sc, osum, osq, on = _oracle(oracle_mod, name)
    r = _render(rx, gpu_ctx, sc, GPU_SPP, mode)
    if mode == MODE_PT:
        _gpu_pt_cache[name] = r
        _check(f"{name} mode {mode}", r["sum"], r["sumsq"], GPU_SPP, osum, osq, on, pooled=True)
    else:
        # the oracle runs the PathTracer estimator: estimate its variance from
        # the oracle's samples pooled with the long GPU PathTracer render of the
        # same scene (validated just above)
        pt = _gpu_pt_cache.get(name) or _render(rx, gpu_ctx, sc, GPU_SPP, MODE_PT)
        m = pt["sum"] / GPU_SPP
        om = osum / on
        ovar = (np.maximum(pt["sumsq"] / GPU_SPP - m * m, 0) * GPU_SPP +
                np.maximum(osq / on - om * om, 0) * on) / (GPU_SPP + on)
        _check(f"{name} mode {mode}", r["sum"], r["sumsq"], GPU_SPP, osum, osq, on, ovar=ovar,
               extra=(pt["sum"], pt["sumsq"]))
    st = r["stats"]
    assert st["samples"] == osum.shape[0] * osum.shape[1] * GPU_SPP and st["rays"] >= st["extend_rays"] > 0


def test_cornell_nee_matches_pathtracertest(rx, oracle_mod, gpu_ctx):
    """On the all-diffuse Cornell box PathTracerTest (NEE+MIS) is consistent with
    PathTracer up to the camera cosine/vignetting factor (~1e-4 here) and the
    depth limit: PathTracerTest also connects the vertex at max_depth to a
    light, one segment beyond the PathTracer's paths (about +0.15 % of this
    image at max depth 8), so the depth is set where Russian roulette has
    ended nearly every path.  The oracle's NEE samples are heavy-tailed, so
    it takes 16384 of them per pixel."""
    print("[render parity] cornell NEE vs PathTracerTest", flush=True)
    sc = scenes.cornell(RES, RES, spp=16, max_depth=30, rr_start_depth=0)
    so = oracle_mod.SceneOracle(sc)
    osum, osq = so.render(16384, seed=6, mode=1)
    r = _render(rx, gpu_ctx, sc, GPU_SPP, MODE_NEE, rr_start_depth=0)
    _check("cornell NEE vs PTtest", r["sum"], r["sumsq"], GPU_SPP, osum, osq, 16384, pooled=False)


def test_render_is_deterministic_and_resumable(rx, gpu_ctx):
    print("[determinism]", flush=True)
    sc = scenes.media_scene(40, 40, freq=16, glass_freq=6)
    pt = rx.PathTracerHIP(sc, ctx=gpu_ctx)
    a = pt.renderPixels(spp=64, mode=MODE_NEE, seed=3)
    b = pt.renderPixels(spp=64, mode=MODE_NEE, seed=3)
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sumsq"], b["sumsq"])
    c = pt.renderPixels(spp=64, mode=MODE_NEE, seed=4)
    assert not np.array_equal(a["sum"], c["sum"])
    # the same sample indices split over two calls give the same samples
    h1 = pt.renderPixels(spp=40, mode=MODE_NEE, seed=3)
    h2 = pt.renderPixels(spp=24, mode=MODE_NEE, seed=3, sample_offset=40)
    assert np.allclose(h1["sum"] + h2["sum"], a["sum"], rtol=1e-12, atol=1e-12)


def _furnace_scene(rho, n_view=32, freq=24, cap_cos=0.9):
    """Camera inside a closed diffuse sphere whose polar cap (z > cap_cos) emits Le=1.
    On a sphere the irradiance from any patch is uniform over the sphere, so every
    wall point has radiance L_w = rho*(f + (1-f) L_w) with f = A_cap/A_sphere."""
    v, t = scenes.icosphere(freq)
    c = v[t].mean(axis=1)
    cap = c[:, 2] > cap_cos
    t = t[:, [0, 2, 1]]  # inward normals
    b = scenes._Builder()
    b.add(v, t[~cap], scenes.DIFFUSE, (rho, rho, rho))
    b.add(v, t[cap], scenes.EMISSIVE, (1.0, 1.0, 1.0))
    vv, tt, oo = b.finish()
    cam = scenes.make_camera([0.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0, 1, 0], n_view, n_view)
    P = vv.astype(np.float64)[tt]
    areas = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    f = areas[oo[1]["tri_begin"]:].sum() / areas.sum()
    return scenes.Scene("furnace", vv, tt, oo, cam, 64, 40, 0), f


@pytest.mark.parametrize("mode", [MODE_PT, MODE_NEE])
def test_furnace_sphere_closed_form(rx, gpu_ctx, mode):
    print(f"[furnace] mode={mode}", flush=True)
    rho = 0.5
    sc, f = _furnace_scene(rho)
    n = 4096
    r = _render(rx, gpu_ctx, sc, n, mode, max_depth=40)
    m = r["sum"] / n
    expect = rho * f / (1 - rho * (1 - f))
    got = m.mean()
    se = math.sqrt(max((r["sumsq"] / n - m * m).mean(), 0) / n / m.size)
    print(f"  f={f:.4f} expect {expect:.6f} got {got:.6f} (se {se:.2e})")
    # tessellation makes the sphere identity approximate: 1 % on top of the noise
    assert abs(got - expect) < 0.01 * expect + 5 * se


@pytest.mark.parametrize("mode", [MODE_PT, MODE_NEE])
def test_absorbing_slab_closed_form(rx, gpu_ctx, mode):
    """A purely absorbing, index-matched medium slab in front of an emitter:
    every pixel converges to Le exp(-sigma_a l) (Beer-Lambert), through the
    medium's distance sampling, its P_surface origin density and the
    transmittance of the segment (src/SceneVPMObject.cpp:4-251,
    include/SceneVPMObject.h:27-47)."""
    print(f"[absorbing slab] mode={mode}", flush=True)
    sc = scenes.absorbing_slab(8, 8)
    n = 65536
    r = _render(rx, gpu_ctx, sc, n, mode)
    zmax, zmean = closed_form_z(r["sum"] / n, r["sumsq"] / n, n, sc.meta["expected"])
    print(f"  max |z| {zmax:.2f}, image-mean z {zmean:.2f}")
    assert zmax < 5.0 and zmean < 4.0


def test_camera_packets_full_size(rx, gpu_ctx):
    """Camera rays traced as packets of 64 (the default) and one per lane hit
    the same triangles at the same distances: on the 1,003,526-triangle
    icosphere scene at 1024x1024 the two films are bit-identical, and the ray
    counts agree."""
    print("[packets] icosphere 1M tris, 1024^2", flush=True)
    sc = scenes.icosphere_scene(1024, 1024, max_depth=10)
    pt = rx.PathTracerHIP(sc, ctx=gpu_ctx)
    try:
        pt.set_tuning(camera_rays=2)
        a = pt.renderPixels(spp=4, mode=MODE_PT, seed=21)
        pt.set_tuning(camera_rays=1)
        b = pt.renderPixels(spp=4, mode=MODE_PT, seed=21)
    finally:
        pt.set_tuning()
    print(f"  packets {a['stats']['ms_total']:.2f} ms, lanes {b['stats']['ms_total']:.2f} ms")
    assert a["stats"]["rays"] == b["stats"]["rays"]
    assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sumsq"], b["sumsq"])


@pytest.mark.parametrize("name", ["icosphere", "glass"])
def test_tail_launch_gives_identical_images(rx, gpu_ctx, name):
    """Plain path tracing with the last bounces of every path carried by one
    launch (k_tail, from rx_render_tuning.tail_bounce on) against one extend
    and one shade launch per bounce: the films are bit-identical and every
    ray is traced once, on the 1,003,526-triangle icosphere and on the glass
    scene (paths inside the refractive sphere)."""
    print(f"[tail] {name}", flush=True)
    sc = scenes.icosphere_scene(256, 256, max_depth=10) if name == "icosphere" else scenes.glass_scene(64, 64)
    pt = rx.PathTracerHIP(sc, ctx=gpu_ctx)
    try:
        pt.set_tuning(tail_bounce=64)  # above max_depth: no tail launch
        a = pt.renderPixels(spp=8, mode=MODE_PT, seed=5)
        for tb in [0, 1, 2, 5]:
            pt.set_tuning(tail_bounce=tb)
            b = pt.renderPixels(spp=8, mode=MODE_PT, seed=5)
            print(f"  tail from bounce {tb or 'default'}: {b['stats']['ms_total']:.2f} ms "
                  f"(per-bounce launches {a['stats']['ms_total']:.2f} ms)")
            assert b["stats"]["rays"] == a["stats"]["rays"], tb
            assert b["stats"]["segments"] == a["stats"]["segments"], tb
            assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sumsq"], b["sumsq"]), tb
    finally:
        pt.set_tuning()


def test_per_bounce_stats_add_up(rx, gpu_ctx):
    """rx_render_stats' per-bounce breakdown: the counted work of each bounce
    sums to the render's totals, bounce 0 traces one camera ray per sample,
    and the breakdown does not depend on how the samples are split into
    passes (one pass against four); the timed breakdown sums to ms_extend and
    puts the tail launch at its first bounce."""
    print("[per-bounce stats]", flush=True)
    sc = scenes.icosphere_scene(256, 256, max_depth=10)
    pt = rx.PathTracerHIP(sc, ctx=gpu_ctx)
    try:
        npaths = 256 * 256 * 8
        counted = []
        for slots in (1 << 22, 1 << 17):
            pt.set_tuning(path_slots=slots)
            st = pt.renderPixels(spp=8, mode=MODE_PT, seed=9, want_sums=False, flags=rx.RENDER_COUNT_WORK)["stats"]
            print(f"  {slots} slots: {st['extend_launches']} launches, bounce 0 box tests {st['box_tests_bounce'][0]}")
            assert st["extend_launches"] == 11 * ((npaths + slots - 1) // slots)
            assert sum(st["extend_rays_bounce"]) == st["extend_rays"]
            assert sum(st["box_tests_bounce"]) == st["box_tests"]
            assert sum(st["tri_tests_bounce"]) == st["tri_tests"]
            assert st["extend_rays_bounce"][0] == npaths
            assert all(v == 0 for v in st["extend_rays_bounce"][11:])
            counted.append(st)
        for k in ("extend_rays_bounce", "box_tests_bounce", "tri_tests_bounce"):
            assert counted[0][k] == counted[1][k], k
        pt.set_tuning(tail_bounce=3)
        st = pt.renderPixels(spp=8, mode=MODE_PT, seed=9, want_sums=False, flags=rx.RENDER_TIMING)["stats"]
        msb = st["ms_extend_bounce"]
        print("  ms per bounce " + " ".join(f"{m:.3f}" for m in msb[:5]))
        assert all(m > 0 for m in msb[:4]) and all(m == 0 for m in msb[4:])
        assert abs(sum(msb) - st["ms_extend"]) <= 1e-6 * max(1.0, st["ms_extend"])
    finally:
        pt.set_tuning()


def test_launch_tuning_gives_identical_images(rx, gpu_ctx):
    """rx_render_tuning (samples per pass, resident blocks, the extend
    kernel's stack depth and block size) and the autotuner change only the
    time: the film is bit-identical."""
    print("[tuning]", flush=True)
    sc = scenes.media_scene(40, 40, freq=16, glass_freq=6)
    pt = rx.PathTracerHIP(sc, ctx=gpu_ctx)
    pt.set_tuning()
    a = pt.renderPixels(spp=48, mode=MODE_NEE, seed=3)
    try:
        for slots, blk, depth, bsz in [(4096, 0, 0, 0), (40 * 40 * 5, 2, 0, 0), (1 << 20, 3, 0, 0), (0, 0, 6, 128),
                                       (0, 0, 8, 0), (4096, 0, 12, 128), (0, 0, 16, 0), (0, 0, 10, 64),
                                       (40 * 40 * 5, 0, 0, 256)]:
            pt.set_tuning(path_slots=slots, blocks_per_cu=blk, stack_depth=depth, block_size=bsz)
            b = pt.renderPixels(spp=48, mode=MODE_NEE, seed=3)
            assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sumsq"], b["sumsq"]), (slots, blk, depth, bsz)
        pt.set_tuning(camera_rays=1)  # camera rays one per lane instead of in packets
        b = pt.renderPixels(spp=48, mode=MODE_NEE, seed=3)
        assert np.array_equal(a["sum"], b["sum"]) and np.array_equal(a["sumsq"], b["sumsq"])
        for depth, bsz in [(7, 0), (16, 256), (0, 32)]:  # shapes without an instantiation
            with pytest.raises(rx.RxError):
                pt.set_tuning(stack_depth=depth, block_size=bsz)
        best, ms = pt.autotune(spp=16, mode=MODE_NEE)
        print(f"  autotuned {best} in {ms:.2f} ms")
        assert best["path_slots"] > 0 and pt.get_tuning() == best
        c = pt.renderPixels(spp=48, mode=MODE_NEE, seed=3)
        assert np.array_equal(a["sum"], c["sum"])
    finally:
        pt.set_tuning()
