"""Mathematical pins of the oracle functions that GPU-vs-oracle agreement
alone cannot pin (both sides are restatements by the same hand):

* the refraction step of RefractiveMaterial::scatter and SceneVPMObject
  (src/RefractiveMaterial.cpp:189-232, src/SceneVPMObject.cpp:44-75,176-209;
  rotMat of include/smallFuncs.h:11-24) against Snell's law, coplanarity,
  the side of the surface, the direction of the tangential component, the
  total-internal-reflection boundary and the textbook Fresnel reflectance;
* the medium densities p_medium / P_surface (include/SceneVPMObject.h:27-47)
  against the exponential transmittance and the normalisation
  int_0^d p_medium(s) ds + P_surface(d) = 1;
* the glossy lobe (include/GlossyBSDF.h:24-33): ks at the mirror direction,
  ks cos^n of the angle to it elsewhere, zero below the surface;
* a purely absorbing, index-matched slab in front of an emitter, rendered by
  the oracle's PathTracer: L = Le exp(-sigma_a l) per pixel.

Each pin is shown to fail when the matching mistake is planted in the oracle
(oracle.mutation): a flipped rotation sign, an inverted index ratio, the
complement for P_surface, p_medium without its sigma_t factor, the glossy lobe
about the wrong axis.
"""
import numpy as np
import pytest

from tests.stats_util import closed_form_z

IORS = [(1.0, 1.5), (1.5, 1.0), (1.0, 1.33), (1.33, 2.4), (2.4, 1.0), (1.2, 1.2)]


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _refraction_violations(orc, kind, n_rays=3000, seed=0):
    """Counts the rays whose refracted direction breaks one of the physical laws."""
    rng = np.random.default_rng(seed)
    bad = {"tir": 0, "snell": 0, "plane": 0, "side": 0, "tangent": 0, "fresnel": 0, "unit": 0}
    for k in range(n_rays):
        n_i, n_t = IORS[k % len(IORS)]
        n = _unit(rng.normal(size=3)).astype(np.float32)
        d = _unit(rng.normal(size=3)).astype(np.float32)
        n64, d64 = n.astype(np.float64), d.astype(np.float64)
        cos_i = abs(float(d64 @ n64))
        sin_i = np.sqrt(max(0.0, 1 - cos_i * cos_i))
        if cos_i > 0.9999 or cos_i < 1e-3:  # normal and grazing incidence: hand cases below
            continue
        ok, t, er = orc.refract_dir(d, n, n_i, n_t, kind)
        s_t = n_i / n_t * sin_i
        if abs(s_t - 1) < 1e-4:
            continue
        if ok != (s_t < 1):
            bad["tir"] += 1
            continue
        if not ok:
            continue
        t = t.astype(np.float64)
        if abs(np.linalg.norm(t) - 1) > 1e-5:
            bad["unit"] += 1
        cos_t = abs(t @ n64)
        sin_t = np.linalg.norm(np.cross(t, n64))
        if abs(n_i * sin_i - n_t * sin_t) > 3e-5 * max(n_i, n_t):
            bad["snell"] += 1
        plane = _unit(np.cross(n64, d64))
        if abs(t @ plane) > 3e-5:
            bad["plane"] += 1
        if np.sign(t @ n64) != np.sign(d64 @ n64):  # the refracted ray crosses the surface
            bad["side"] += 1
        if sin_i > 1e-2 and (t - (t @ n64) * n64) @ (d64 - (d64 @ n64) * n64) <= 0:
            bad["tangent"] += 1
        rs = ((n_i * cos_i - n_t * cos_t) / (n_i * cos_i + n_t * cos_t)) ** 2
        rp = ((n_t * cos_i - n_i * cos_t) / (n_t * cos_i + n_i * cos_t)) ** 2
        if abs(er - 0.5 * (rs + rp)) > 2e-5 + 2e-4 * 0.5 * (rs + rp):
            bad["fresnel"] += 1
    return bad


@pytest.mark.parametrize("kind", [0, 1])
def test_refraction_obeys_snell(oracle_mod, kind):
    bad = _refraction_violations(oracle_mod, kind)
    assert sum(bad.values()) == 0, bad


@pytest.mark.parametrize("kind", [0, 1])
def test_refraction_hand_cases(oracle_mod, kind):
    n = np.array([0, 0, 1], np.float32)
    # 45 degrees into glass: sin(theta_t) = sin 45 / 1.5
    d = _unit(np.array([1, 0, -1], np.float64)).astype(np.float32)
    ok, t, er = oracle_mod.refract_dir(d, n, 1.0, 1.5, kind)
    st = np.sin(np.pi / 4) / 1.5
    assert ok and np.allclose(t, [st, 0, -np.sqrt(1 - st * st)], atol=2e-6)
    # out of glass beyond the critical angle (41.8 deg): total internal reflection
    d = _unit(np.array([np.sin(np.radians(45)), 0, np.cos(np.radians(45))])).astype(np.float32)
    ok, _, _ = oracle_mod.refract_dir(d, n, 1.5, 1.0, kind)
    assert not ok
    # index-matched: straight through, no reflection
    d = _unit(np.array([0.3, -0.4, -0.8])).astype(np.float32)
    ok, t, er = oracle_mod.refract_dir(d, n, 1.3, 1.3, kind)
    assert ok and np.allclose(t, d, atol=2e-6) and abs(er) < 1e-6
    if kind == 0:  # normal incidence (only RefractiveMaterial clamps the cosine)
        ok, t, er = oracle_mod.refract_dir(np.array([0, 0, -1], np.float32), n, 1.0, 1.5, kind)
        assert ok and np.allclose(t, [0, 0, -1], atol=1e-6) and abs(er - 0.04) < 1e-6


@pytest.mark.parametrize("mut,law", [("MUT_ROTATION_SIGN", "tangent"), ("MUT_INVERTED_RATIO", "snell")])
def test_refraction_pins_catch_planted_mistakes(oracle_mod, mut, law):
    with oracle_mod.mutation(getattr(oracle_mod, mut)):
        bad = _refraction_violations(oracle_mod, 0, n_rays=600)
    assert bad[law] > 50, bad


SIGMAS = [(0.55, 0.55, 0.55), (1.0, 0.3, 2.5), (0.05, 0.05, 0.05)]


def _medium_violations(orc, sigma_t):
    st = np.array(sigma_t, np.float64)
    w = np.array([0.212671, 0.715160, 0.072169])
    bad = 0
    for d in [0.01, 0.3, 1.0, 2.7, 6.0]:
        _, P = orc.medium_densities(st, d)
        if abs(P - w @ np.exp(-st * d)) > 2e-6:
            bad += 1
        if np.ptp(st) == 0 and abs(P - np.exp(-st[0] * d)) > 2e-6:  # grey: plain transmittance
            bad += 1
        # int_0^d p_medium(s) ds + P_surface(d) = 1 (composite Simpson, 2001 nodes)
        s = np.linspace(0.0, d, 2001)
        pm = np.array([orc.medium_densities(st, float(x))[0] for x in s], np.float64)
        h = d / 2000
        integral = h / 3 * (pm[0] + pm[-1] + 4 * pm[1:-1:2].sum() + 2 * pm[2:-1:2].sum())
        if abs(integral + P - 1) > 1e-5:
            bad += 1
    pm0, P0 = orc.medium_densities(st, 0.0)
    if abs(pm0 - w @ st) > 1e-6 or abs(P0 - 1) > 1e-6:
        bad += 1
    return bad


@pytest.mark.parametrize("sigma_t", SIGMAS)
def test_medium_densities(oracle_mod, sigma_t):
    assert _medium_violations(oracle_mod, sigma_t) == 0


@pytest.mark.parametrize("mut", ["MUT_P_SURFACE_COMPLEMENT", "MUT_P_MEDIUM_NO_SIGMA"])
def test_medium_pins_catch_planted_mistakes(oracle_mod, mut):
    with oracle_mod.mutation(getattr(oracle_mod, mut)):
        assert _medium_violations(oracle_mod, SIGMAS[1]) >= 3


def _glossy_violations(orc, n_cases=400, seed=3):
    rng = np.random.default_rng(seed)
    bad = 0
    for k in range(n_cases):
        n = _unit(rng.normal(size=3))
        d = _unit(rng.normal(size=3))
        if d @ n > -0.05:
            d = d - 2 * (d @ n) * n if d @ n > 0 else d
            if d @ n > -0.05:
                continue
        ks = rng.uniform(0.1, 1.0, 3)
        e = float(rng.choice([1.0, 10.0, 50.0]))
        r = d - 2 * (d @ n) * n
        f = orc.glossy_eval(n, d, r, ks, e)
        if not np.allclose(f, ks, rtol=2e-5):
            bad += 1
        # rotate r by alpha towards the normal or away within the plane (n, r): stays above the surface
        axis = _unit(np.cross(r, n)) if np.linalg.norm(np.cross(r, n)) > 1e-3 else _unit(np.cross(r, [1, 0, 0]))
        alpha = rng.uniform(0.02, 0.6)
        w = r * np.cos(alpha) + np.cross(axis, r) * np.sin(alpha) + axis * (axis @ r) * (1 - np.cos(alpha))
        w = _unit(w)
        f = orc.glossy_eval(n, d, w, ks, e).astype(np.float64)
        c = w @ r
        if w @ n >= 1e-4:
            expect = ks * c ** e
            if not np.allclose(f, expect, rtol=2e-3, atol=1e-7):
                bad += 1
        elif w @ n < -1e-4 and np.any(f != 0):
            bad += 1
        # below the surface: zero
        wb = _unit(w - 2 * max(w @ n, 0.05) * n) if w @ n > 0 else w
        if wb @ n < -1e-4 and np.any(orc.glossy_eval(n, d, wb, ks, e) != 0):
            bad += 1
    return bad


def test_glossy_lobe_value(oracle_mod):
    assert _glossy_violations(oracle_mod) == 0


def test_glossy_pin_catches_planted_mistake(oracle_mod):
    with oracle_mod.mutation(oracle_mod.MUT_GLOSSY_AXIS):
        assert _glossy_violations(oracle_mod, n_cases=100) > 30


def test_absorbing_slab_oracle(oracle_mod):
    """The oracle's PathTracer through a purely absorbing, index-matched slab:
    each pixel converges to Le exp(-sigma_a l) (Beer-Lambert)."""
    from renderx_amd import scenes

    sc = scenes.absorbing_slab(8, 8)
    spp = 4096
    s, q = oracle_mod.SceneOracle(sc).render(spp, seed=11)
    zmax, zmean = closed_form_z(s / spp, q / spp, spp, sc.meta["expected"])
    assert zmax < 5.0 and zmean < 4.0, (zmax, zmean)
    # and the same render with the transmittance written wrongly fails
    with oracle_mod.mutation(oracle_mod.MUT_P_SURFACE_COMPLEMENT):
        s, q = oracle_mod.SceneOracle(sc).render(256, seed=11)
    assert abs((s / 256).mean() - sc.meta["expected"].mean()) > 0.1
