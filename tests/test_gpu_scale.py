"""GPU parity at full size.

Config 2 (n = 2^28, d = 8, p = 0.5, seed = 7):
  * the first 2^24 rows against the sequential oracle, bit for bit;
  * a deterministic subsample of later node ranges ([2^k, 2^k + 4096), the
    last 4096 nodes and 16 Philox-placed ranges) against the closure oracle;
  * every row against the oracle's row-consistency verifier (each row must be
    what the model yields from the GPU's own earlier rows; all rows passing is
    equivalent to bit-exactness by induction over t);
  * size-independent properties: targets of t in [0, t), distinct within a
    row, row_ptr[t] = t*d, the degree histogram equal to an independent torch
    recount, sum(hist) = n and sum(deg * hist) = d(d+1) + 2(n-d-1)d.
Config 3 (n = 2^30, d = 8, p = 0.2, seed = 11): the first 2^22 rows bit for
bit against the sequential oracle, and every row through the
row-consistency verifier; the stats' node count.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, D, P, SEED = 1 << 28, 8, 0.5, 7
PREFIX = 1 << 24
THREADS = min(16, os.cpu_count() or 1)


def sample_ranges(n, width=4096, extra=16, key=SEED):
    """[2^k, 2^k+width) for 2^k+width <= n, the last `width` nodes, and
    `extra` ranges placed by keyed Philox draws."""
    import oracle
    out = []
    k = 4
    while (1 << k) + width <= n:
        out.append((1 << k, width))
        k += 1
    out.append((n - width, width))
    for i in range(extra):
        x = oracle.philox((i, 0xC0FFEE, 0, 0), (key, 0))
        start = ((x[0] << 32) | x[1]) % (n - width)
        out.append((int(start), width))
    return out


@pytest.fixture(scope="module")
def big_graph(ctx):
    import torch
    print(f"\n[scale] generating n={N} d={D} p={P} seed={SEED}", flush=True)
    g = ctx.generate(N, D, P, SEED)
    yield g
    del g
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def big_host(big_graph):
    return big_graph.targets.cpu().numpy()


def test_prefix_and_subsample_match_oracle(big_host, oracle_mod):
    T = oracle_mod.p_threshold(P)
    prefix, _, _ = oracle_mod.generate(PREFIX, D, T, SEED)
    assert np.array_equal(big_host[:PREFIX * D], prefix)
    clo = oracle_mod.Closure(N, D, T, SEED, prefix=prefix)
    for start, width in sample_ranges(N):
        if start + width <= PREFIX:
            continue
        want = clo.rows(start, width)
        got = big_host[start * D:(start + width) * D]
        assert np.array_equal(got, want), f"rows [{start}, {start + width})"


def test_every_row_consistent(big_host, oracle_mod):
    T = oracle_mod.p_threshold(P)
    bad = oracle_mod.verify_rows(N, D, T, SEED, big_host, threads=THREADS)
    assert bad is None, f"first inconsistent row {bad}"


def test_structure_invariants(big_graph):
    import torch
    n, d = N, D
    rows = big_graph.targets.view(n, d)
    step = 1 << 24
    for lo in range(d + 1, n, step):
        hi = min(n, lo + step)
        r = rows[lo:hi]
        t = torch.arange(lo, hi, device=r.device, dtype=torch.int64).unsqueeze(1)
        assert bool((r >= 0).all()) and bool((r.to(torch.int64) < t).all())
        s, _ = torch.sort(r, dim=1)
        assert bool((s[:, 1:] != s[:, :-1]).all())
    rp = big_graph.row_ptr
    assert int(rp[0]) == 0 and int(rp[-1]) == n * d
    assert bool((rp[1:] - rp[:-1] == d).all())


def test_full_degree_histogram(ctx, big_graph):
    import torch
    n, d = N, D
    hist, mx, deg = ctx.degree_histogram(big_graph, want_degree=True)
    tg = big_graph.targets[(d + 1) * d:]
    indeg = torch.zeros(n, dtype=torch.int64, device=tg.device)
    step = 1 << 28
    for lo in range(0, tg.numel(), step):
        indeg += torch.bincount(tg[lo:lo + step].to(torch.int64), minlength=n)
    want_deg = indeg + d
    assert bool((deg.to(torch.int64) == want_deg).all())
    want_hist = torch.bincount(want_deg)
    assert mx == int(want_deg.max())
    assert torch.equal(hist, want_hist)
    h = hist.to(torch.float64)
    x = torch.arange(h.numel(), device=h.device, dtype=torch.float64)
    assert int(h.sum()) == n
    assert int((h * x).sum()) == d * (d + 1) + 2 * (n - d - 1) * d


@pytest.mark.timeout(900)
def test_n2p30_every_row_consistent(ctx, big_graph, oracle_mod):
    import torch
    n, d, p, seed = 1 << 30, 8, 0.2, 11
    T = oracle_mod.p_threshold(p)
    print(f"\n[scale] generating n={n} d={d} p={p} seed={seed}", flush=True)
    g = ctx.generate(n, d, p, seed, row_ptr=False)
    assert ctx.stats()["nodes"] == n - d - 1
    host = g.targets.cpu().numpy()
    del g
    torch.cuda.empty_cache()
    want, _, _ = oracle_mod.generate(1 << 22, d, T, seed)
    assert np.array_equal(host[:(1 << 22) * d], want)
    del want
    print("[scale] verifying every row of n=2^30", flush=True)
    bad = oracle_mod.verify_rows(n, d, T, seed, host, threads=THREADS)
    assert bad is None, f"first inconsistent row {bad}"


def test_d16_n2p24_every_row_consistent(ctx, oracle_mod):
    """d = 16, p = 0.1 at n = 2^24: the 160 KB LDS prefix (2,500 nodes) and the
    persistent kernel after it. The first 2^20 rows bit for bit against the
    sequential oracle, every row through the row-consistency verifier."""
    import torch
    n, d, p, seed = 1 << 24, 16, 0.1, 5
    T = oracle_mod.p_threshold(p)
    print(f"\n[scale] generating n={n} d={d} p={p} seed={seed}", flush=True)
    g = ctx.generate(n, d, p, seed)
    host = g.targets.cpu().numpy()
    assert ctx.stats()["nodes"] == n - d - 1
    del g
    torch.cuda.empty_cache()
    want, _, _ = oracle_mod.generate(1 << 20, d, T, seed)
    assert np.array_equal(host[:(1 << 20) * d], want)
    assert oracle_mod.verify_rows(n, d, T, seed, host, threads=THREADS) is None
