"""Distribution pins of the oracle's restated samplers against closed forms:
cosine-power lobe (include/CosineSphericalSampler.h, src/RandGenerator.cpp:152-166),
Henyey-Greenstein (include/HGPhaseSampler.h:13-34), exponential distance
(include/HomoMediaDistSampler.h:9-14) and the dielectric Fresnel term
(src/RefractiveMaterial.cpp:224-231)."""
import numpy as np
import pytest


@pytest.mark.parametrize("n_exp", [1.0, 50.0])
def test_cosine_lobe_moments(oracle_mod, n_exp):
    d = oracle_mod.sampler_draws(0, 200000, n_exp).astype(np.float64)
    c = d[:, 1]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    # p(cos) = (n+1) cos^n on [0,1]  ->  E[cos] = (n+1)/(n+2), E[cos^2] = (n+1)/(n+3)
    assert abs(c.mean() - (n_exp + 1) / (n_exp + 2)) < 4 * c.std() / np.sqrt(len(c)) + 1e-6
    assert abs((c * c).mean() - (n_exp + 1) / (n_exp + 3)) < 5e-3
    # azimuth uniform
    phi = np.arctan2(d[:, 2], d[:, 0])
    h, _ = np.histogram(phi, bins=16, range=(-np.pi, np.pi))
    exp = len(phi) / 16
    assert ((h - exp) ** 2 / exp).sum() < 45  # chi2, 15 dof, p ~ 1e-4


@pytest.mark.parametrize("g", [0.0, 0.7, -0.3])
def test_hg_mean_cosine(oracle_mod, g):
    d = oracle_mod.sampler_draws(1, 200000, g).astype(np.float64)
    c = d[:, 1]
    assert abs(c.mean() - g) < 5 * c.std() / np.sqrt(len(c))
    # chi2 of the cosine histogram against the HG marginal
    edges = np.linspace(-1, 1, 21)

    def cdf(x):
        if abs(g) < 1e-5:
            return (x + 1) / 2
        return (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * x) - 1 / (1 + g))

    p = np.diff(cdf(edges))
    h, _ = np.histogram(c, bins=edges)
    e = p * len(c)
    m = e > 5
    assert ((h[m] - e[m]) ** 2 / e[m]).sum() < 60


def test_medium_distance_mean(oracle_mod):
    s = oracle_mod.sampler_draws(2, 200000, 0.55)[:, 0].astype(np.float64)
    assert abs(s.mean() - 1 / 0.55) < 5 * s.std() / np.sqrt(len(s))
    assert abs(np.median(s) - np.log(2) / 0.55) < 0.02


def test_fresnel_normal_incidence(oracle_mod):
    er = oracle_mod.sampler_draws(3, 1, 1.5, 1.0)[0, 0]
    assert abs(er - 0.04) < 1e-6  # ((n-1)/(n+1))^2
    er60 = oracle_mod.sampler_draws(3, 1, 1.5, 0.5)[0, 0]
    ci, ct = 0.5, np.sqrt(1 - (np.sin(np.arccos(0.5)) / 1.5) ** 2)
    rs = ((ci - 1.5 * ct) / (ci + 1.5 * ct)) ** 2
    rp = ((1.5 * ci - ct) / (1.5 * ci + ct)) ** 2
    assert abs(er60 - (rs + rp) / 2) < 1e-5
