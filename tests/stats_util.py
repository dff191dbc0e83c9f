"""Per-pixel statistical comparison of two renders given FP64 sums and sums of squares."""
import numpy as np


def mean_var(s, q, n):
    m = s / n
    v = np.maximum(q / n - m * m, 0.0) * n / max(n - 1, 1)
    return m, v


def z_scores(s1, q1, n1, s2, q2, n2, pooled=True):
    """Per pixel-channel z of the mean difference.  Under the null both renders
    sample the same per-pixel distribution, so the pooled variance (dominated by
    the larger sample) is the robust estimate for heavy-tailed path samples."""
    m1, v1 = mean_var(s1, q1, n1)
    m2, v2 = mean_var(s2, q2, n2)
    if pooled:
        v = (v1 * n1 + v2 * n2) / (n1 + n2)
        se = np.sqrt(v / n1 + v / n2)
    else:
        se = np.sqrt(v1 / n1 + v2 / n2)
    diff = m1 - m2
    z = np.zeros_like(diff)
    nz = se > 0
    z[nz] = diff[nz] / se[nz]
    # both estimates exactly equal with zero variance (e.g. black pixels): z = 0;
    # zero variance with a difference is an infinite z
    z[(~nz) & (diff != 0)] = np.inf
    return z


def n_eff(s, q):
    """Effective number of contributing samples (sum x)^2 / sum x^2: n for a
    constant estimator, k for k equal non-zero samples among zeros.  A
    pixel-channel whose estimate rests on a handful of rare samples has an
    unreliable variance estimate, so the normal approximation behind a
    z-test does not hold there."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(q > 0, s * s / np.where(q > 0, q, 1), 0.0)


def clt_mask(*sums, min_eff=3):
    """Pixel-channels where every (sum, sumsq) pair has at least min_eff
    effective samples, or where all of them are exactly zero.  An estimate
    resting on fewer than three samples has no usable variance estimate."""
    ok = np.ones(np.shape(sums[0]), bool)
    zero = np.ones(np.shape(sums[0]), bool)
    for s, q in zip(sums[0::2], sums[1::2]):
        ok &= n_eff(s, q) >= min_eff
        zero &= (s == 0) & (q == 0)
    return ok | zero


def three_sigma(s1, q1, n1, s2, q2, n2, pooled=True, mask=None):
    """Fraction of pixel-channels within 3 sigma, mean z (z clipped to +-5 so
    that it measures a systematic shift, not outliers) and the z array, over
    `mask` (default: all).  Use pooled=False when the two renders use
    different estimators."""
    z = z_scores(s1, q1, n1, s2, q2, n2, pooled)
    sel = np.ones(z.shape, bool) if mask is None else mask
    zs = z[sel]
    return float((np.abs(zs) <= 3).mean()), float(np.clip(zs, -5, 5).mean()), z


def rel_mse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))


def closed_form_z(mean, sumsq_mean, spp, expected):
    """Per pixel-channel z of a render's means against a closed-form image
    expected (H, W): returns (max |z|, |z| of the image mean)."""
    var = np.maximum(sumsq_mean - mean * mean, 0)
    se = np.sqrt(var / spp) + 1e-9
    z = (mean - expected[..., None]) / se
    tot_se = np.sqrt((var / spp).sum()) / mean.size
    return float(np.abs(z).max()), float(abs(mean.mean() - expected.mean()) / tot_se)
