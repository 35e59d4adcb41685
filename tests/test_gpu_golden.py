"""GPU generator against the committed golden fixtures and the reference
model's degree distributions (no /root/reference needed at run time)."""
import hashlib
import json
import os

import numpy as np
import pytest

from stats_util import GOLDEN, REF_DD, from_hist, ks_critical, ks_distance, powerlaw_gamma, ref_dd

pytestmark = pytest.mark.gpu


def test_tiny_graphs_fixture(ctx):
    for g in json.load(open(os.path.join(GOLDEN, "tiny_graphs.json"))):
        out = ctx.generate(g["n"], g["d"], p_thresh=g["p_threshold"], seed=g["seed"])
        assert out.targets.cpu().tolist() == g["targets"], g


def test_n1e6_d4_digest(ctx):
    dg = json.load(open(os.path.join(GOLDEN, "n1e6_d4_p0.5_digest.json")))
    g = ctx.generate(dg["n"], dg["d"], dg["p"], dg["seed"])
    G = g.targets.cpu().numpy()
    assert hashlib.sha256(G.tobytes()).hexdigest() == dg["sha256"]
    assert ctx.stats()["retries"] == dg["retries"]


@pytest.mark.parametrize("n,d,p", REF_DD)
def test_gpu_degree_distribution_matches_reference_model(ctx, n, d, p):
    g = ctx.generate(n, d, p, 1)
    hist, _, _ = ctx.degree_histogram(g)
    x2, c2 = from_hist(hist.cpu().numpy())
    x1, c1 = ref_dd(n, d, p)
    assert c2.sum() == n
    assert ks_distance(x1, c1, x2, c2) < ks_critical(c1.sum(), c2.sum())
    assert abs(powerlaw_gamma(x1, c1, 4 * d) - powerlaw_gamma(x2, c2, 4 * d)) < 0.15
