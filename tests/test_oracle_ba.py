"""The oracle's Barabasi-Albert generator (the reference's SerialBA,
src/PA/SPA/SerialBA.cpp:68-187, with keyed draws) pinned on the CPU:

* against an independent pure-Python copy-model restatement over self-loop
  seed rows at p = 1/2, which is the form the GPU runs (bit for bit);
* against the exact probabilities of node d+1's first target, fixed by the
  endpoint counts of the seed (a clique seed or a biased endpoint draw fails);
* against invariants (distinct earlier targets, the degree sum 2d(n-d));
* against the reference's own generator (built from its sources by
  `make -C oracle ref`; committed "Degree,Count" fixtures): two-sample KS
  and tail exponent;
* against the closed-form BA degree distribution 2m(m+1)/(k(k+1)(k+2)).
"""
import os

import numpy as np
import pytest
from scipy import stats

from stats_util import GOLDEN, from_hist, ks_critical, ks_distance, powerlaw_gamma, read_dd

REF_DD_BA = [(1_000_000, 1), (1_000_000, 4), (300_000, 16)]


@pytest.mark.parametrize("n,d,seed", [(2, 1, 0), (9, 1, 3), (20, 2, 1), (40, 3, 2), (64, 5, 9),
                                      (48, 8, 4), (40, 16, 2**33 + 5)])
def test_ba_matches_copy_model_form(oracle_mod, n, d, seed):
    G, _ = oracle_mod.generate_ba(n, d, seed)
    want = oracle_mod.py_generate(n, d, 1 << 31, seed, self_loop_seed=True)
    assert G.tolist() == want.tolist()


@pytest.mark.parametrize("d", [1, 3, 8])
def test_ba_first_step_probabilities(oracle_mod, d):
    """Node d+1 draws over the 2d(d+1) endpoints of nodes 0..d: seed node
    v < d owns 2d + 1 of them, node d owns d."""
    samples = 6000
    first = np.array([oracle_mod.generate_ba(d + 2, d, s)[0][(d + 1) * d] for s in range(samples)])
    counts = np.bincount(first, minlength=d + 1)
    assert counts.size == d + 1
    prob = np.array([(2 * d + 1) / (2 * d * (d + 1))] * d + [1 / (2 * (d + 1))])
    chi2 = float((((counts - samples * prob) ** 2) / (samples * prob)).sum())
    assert chi2 < stats.chi2.ppf(0.999, d)


@pytest.mark.parametrize("n,d,seed", [(5000, 1, 1), (20000, 4, 2), (3000, 16, 3)])
def test_ba_invariants(oracle_mod, n, d, seed):
    G, _ = oracle_mod.generate_ba(n, d, seed)
    rows = G.reshape(n, d)
    assert (rows[:d] == np.arange(d)[:, None]).all()
    assert rows[d].tolist() == list(range(d))
    t = np.arange(n)[:, None]
    assert ((rows[d + 1:] >= 0) & (rows[d + 1:] < t[d + 1:])).all()
    srt = np.sort(rows[d + 1:], axis=1)
    assert (np.diff(srt, axis=1) > 0).all()
    deg, hist, mx = oracle_mod.degrees_ba(n, d, G, hist_len=n)
    assert int(deg.sum()) == 2 * d * (n - d)
    assert (deg[d:] >= d).all()
    assert mx == int(deg.max()) and int(hist.sum()) == n


@pytest.mark.parametrize("n,d", REF_DD_BA)
def test_ba_degree_distribution_matches_reference(oracle_mod, n, d):
    G, _ = oracle_mod.generate_ba(n, d, 1)
    deg, _, _ = oracle_mod.degrees_ba(n, d, G)
    x2, c2 = from_hist(np.bincount(deg.astype(np.int64)))
    x1, c1 = read_dd(os.path.join(GOLDEN, f"ref_dd_ba_n{n}_d{d}.csv"))
    assert c1.sum() == c2.sum() == n
    assert ks_distance(x1, c1, x2, c2) < ks_critical(c1.sum(), c2.sum())
    assert abs(powerlaw_gamma(x1, c1, 4 * d) - powerlaw_gamma(x2, c2, 4 * d)) < 0.1


def test_ba_degree_closed_form(oracle_mod):
    n, d = 1_000_000, 4
    G, _ = oracle_mod.generate_ba(n, d, 5)
    deg, _, _ = oracle_mod.degrees_ba(n, d, G)
    h = np.bincount(deg.astype(np.int64)) / n
    k = np.arange(d, d + 10)
    assert np.allclose(h[k], 2 * d * (d + 1) / (k * (k + 1) * (k + 2)), rtol=0.04)


def test_reference_ba_fixtures_are_reproducible(oracle_mod, tmp_path):
    if not os.path.exists(oracle_mod.REF_BA_BIN):
        pytest.skip("reference build absent")
    n, d = REF_DD_BA[1]
    dd, _ = oracle_mod.run_reference_sba(n, d, str(tmp_path))
    x, c = read_dd(os.path.join(GOLDEN, f"ref_dd_ba_n{n}_d{d}.csv"))
    assert dd == {int(a): int(b) for a, b in zip(x, c)}


def test_ba_row_verifier_accepts_oracle_and_catches_planted_mistakes(oracle_mod):
    n, d, seed = 30000, 4, 11
    G, _ = oracle_mod.generate_ba(n, d, seed)
    assert oracle_mod.verify_rows(n, d, 0, seed, G, threads=4, model=1) is None
    # the copy-model check rejects the self-loop seed rows
    assert oracle_mod.verify_rows(n, d, 1 << 31, seed, G, 0, d + 1) == 0
    bad = G.copy()
    bad[(20000 * d) + 2] = (bad[(20000 * d) + 2] + 1) % 20000
    assert oracle_mod.verify_rows(n, d, 0, seed, bad, threads=4, model=1) == 20000
    bad = G.copy()
    bad[1 * d] = 0
    assert oracle_mod.verify_rows(n, d, 0, seed, bad, 0, d + 1, model=1) == 1
