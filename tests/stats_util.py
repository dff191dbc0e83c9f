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


def three_sigma(s1, q1, n1, s2, q2, n2, pooled=True):
    """Fraction of pixel-channels within 3 sigma, mean z, and the z array.
    Use pooled=False when the two renders use different estimators."""
    z = z_scores(s1, q1, n1, s2, q2, n2, pooled)
    fin = np.isfinite(z)
    return float((np.abs(z) <= 3).mean()), float(z[fin].mean()), z


def rel_mse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))
