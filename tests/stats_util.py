"""Degree-distribution statistics shared by the parity tests."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_DD = [(1_000_000, 4, 0.5), (1_000_000, 8, 0.2), (300_000, 16, 0.5), (300_000, 2, 0.9)]


def read_dd(path):
    """'Degree,Count' CSV (cuPPA.cu:148-160, PAHelper.hpp:57-62) -> (x, c)."""
    with open(path) as f:
        assert f.readline().strip() == "Degree,Count"
        rows = [tuple(map(int, line.strip().split(","))) for line in f if line.strip()]
    x = np.array([r[0] for r in rows], dtype=np.int64)
    c = np.array([r[1] for r in rows], dtype=np.float64)
    return x, c


def ref_dd(n, d, p):
    return read_dd(os.path.join(GOLDEN, f"ref_dd_n{n}_d{d}_p{p}.csv"))


def from_hist(hist):
    hist = np.asarray(hist)
    x = np.nonzero(hist)[0]
    return x.astype(np.int64), hist[x].astype(np.float64)


def ks_distance(x1, c1, x2, c2):
    """Two-sample Kolmogorov-Smirnov distance between two degree histograms."""
    xs = np.union1d(x1, x2)
    f1 = np.cumsum(np.bincount(np.searchsorted(xs, x1), weights=c1, minlength=xs.size)) / c1.sum()
    f2 = np.cumsum(np.bincount(np.searchsorted(xs, x2), weights=c2, minlength=xs.size)) / c2.sum()
    return float(np.abs(f1 - f2).max())


def ks_critical(n1, n2, alpha_coeff=1.95):
    """c(alpha) * sqrt((n1 + n2) / (n1 n2)); 1.95 is alpha = 0.001."""
    return alpha_coeff * np.sqrt((n1 + n2) / (n1 * n2))


def powerlaw_gamma(x, c, xmin):
    """Discrete power-law MLE of the tail exponent above xmin (Clauset,
    Shalizi and Newman 2009, eq. 3.7 approximation)."""
    m = x >= xmin
    return float(1.0 + c[m].sum() / (c[m] * np.log(x[m] / (xmin - 0.5))).sum())
