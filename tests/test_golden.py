"""Oracle against the committed golden fixtures, and statistical parity of the
model with the reference's own serial copy model (CPU only)."""
import hashlib
import json
import os

import numpy as np
import pytest

from stats_util import GOLDEN, REF_DD, from_hist, ks_critical, ks_distance, powerlaw_gamma, ref_dd


def test_tiny_graphs_fixture(oracle_mod):
    for g in json.load(open(os.path.join(GOLDEN, "tiny_graphs.json"))):
        assert oracle_mod.p_threshold(g["p"]) == g["p_threshold"]
        G, _, _ = oracle_mod.generate(g["n"], g["d"], g["p_threshold"], g["seed"])
        assert G.tolist() == g["targets"], g


def test_n1e6_d4_digest(oracle_mod):
    dg = json.load(open(os.path.join(GOLDEN, "n1e6_d4_p0.5_digest.json")))
    n, d = dg["n"], dg["d"]
    G, _, retries = oracle_mod.generate(n, d, oracle_mod.p_threshold(dg["p"]), dg["seed"])
    assert hashlib.sha256(G.tobytes()).hexdigest() == dg["sha256"]
    assert retries == dg["retries"]
    for s, rows in dg["slices"].items():
        s = int(s)
        assert G[s * d:(s + 16) * d].tolist() == rows


@pytest.mark.parametrize("n,d,p", REF_DD)
def test_degree_distribution_matches_reference_model(oracle_mod, n, d, p):
    """Same (n, d, p) as the reference's serial copy model run: the degree
    distributions must be indistinguishable (two-sample KS at alpha = 0.001)
    and the fitted tail exponents close."""
    G, _, _ = oracle_mod.generate(n, d, oracle_mod.p_threshold(p), 1)
    deg, _, _ = oracle_mod.degrees(n, d, G)
    x2, c2 = from_hist(np.bincount(deg.astype(np.int64)))
    x1, c1 = ref_dd(n, d, p)
    assert ks_distance(x1, c1, x2, c2) < ks_critical(c1.sum(), c2.sum())
    xmin = 4 * d
    assert abs(powerlaw_gamma(x1, c1, xmin) - powerlaw_gamma(x2, c2, xmin)) < 0.15


def test_tail_exponent_near_copying_model_theory(oracle_mod):
    """Kumar et al.: in-degree tail exponent (2 - p) / (1 - p)."""
    for d, p, want in [(4, 0.5, 3.0), (8, 0.2, 2.25)]:
        n = 1_000_000
        G, _, _ = oracle_mod.generate(n, d, oracle_mod.p_threshold(p), 3)
        deg, _, _ = oracle_mod.degrees(n, d, G)
        x, c = from_hist(np.bincount(deg.astype(np.int64)))
        assert abs(powerlaw_gamma(x, c, 8 * d) - want) < 0.12


def test_reference_fixtures_are_reproducible(oracle_mod, tmp_path):
    if not os.path.exists(oracle_mod.REF_BIN):
        pytest.skip("reference build absent")
    n, d, p = REF_DD[3]
    dd, _ = oracle_mod.run_reference_scm(n, d, p, str(tmp_path))
    x, c = ref_dd(n, d, p)
    assert dd == {int(a): int(b) for a, b in zip(x, c)}
