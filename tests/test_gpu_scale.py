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
row-consistency verifier; the stats' node count. Its degree pass and in-edge
CSR (2^33 slots: positions past 2^32) against torch recounts: every degree,
every in-list's length, every list ascending, the sources' multiset through
two checksums, and the exact lists of hub and tail nodes.
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


@pytest.mark.timeout(900)
def test_n2p30_degree_and_in_csr(ctx):
    import torch
    n, d, p, seed = 1 << 30, 8, 0.2, 11
    ctx.trim()
    torch.cuda.empty_cache()
    print(f"\n[scale] generating n={n} d={d} p={p} seed={seed}", flush=True)
    g = ctx.generate(n, d, p, seed, row_ptr=False)
    tg = g.targets
    step = 1 << 28
    # degrees: deg(v) = d + slots of nodes t >= d+1 naming v
    print("[scale] degree pass against torch.bincount", flush=True)
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    cnt = torch.zeros(n, dtype=torch.int64, device=tg.device)
    first = (d + 1) * d
    for lo in range(first, tg.numel(), step):
        cnt += torch.bincount(tg[lo:min(lo + step, tg.numel())].to(torch.int64), minlength=n)
    assert bool((deg.to(torch.int64) == cnt + d).all())
    assert mx == int(cnt.max()) + d
    assert torch.equal(hist, torch.bincount(cnt + d))
    del deg, hist
    ctx.trim()
    cnt += torch.bincount(tg[:first].to(torch.int64), minlength=n)  # in-edges of all slots
    # exact lists of a few nodes, from the targets themselves (a row names v at most once)
    probe = [0, 1, d, d + 1, 1000, 1 << 20, n // 2, n - 1]
    # (torch.nonzero of more than 2^32 elements overflows: chunks of 2^28)
    want_lists = {}
    for v in probe:
        parts = [torch.nonzero(tg[lo:lo + step] == v).flatten() + lo for lo in range(0, tg.numel(), step)]
        want_lists[v] = (torch.cat(parts) // d).to(torch.int32).cpu()
    print("[scale] in-edge CSR", flush=True)
    in_ptr, in_src = ctx.in_csr(g)
    ctx.trim()
    del g, tg
    torch.cuda.empty_cache()
    total = n * d
    assert int(in_ptr[0]) == 0 and int(in_ptr[n]) == total
    assert torch.equal(in_ptr[1:] - in_ptr[:-1], cnt)
    del cnt
    for v in probe:
        assert torch.equal(in_src[int(in_ptr[v]):int(in_ptr[v + 1])].cpu(), want_lists[v]), f"in-list of {v}"
    # every list strictly ascending: a non-increasing step only where a list starts
    for lo in range(0, total - 1, step):
        hi = min(lo + step, total - 1)
        x = in_src[lo:hi + 1].to(torch.int64)
        bad = (x[1:] - x[:-1]) <= 0  # position lo + 1 + i
        a = int(torch.searchsorted(in_ptr, torch.tensor([lo + 1], device=in_ptr.device)))
        b = int(torch.searchsorted(in_ptr, torch.tensor([hi + 1], device=in_ptr.device)))
        starts = in_ptr[a:b] - (lo + 1)
        bad[starts] = False
        assert not bool(bad.any()), f"a list is not ascending in [{lo}, {hi}]"
    # each node is the source of exactly d slots: two checksums of the sources
    s1 = s2 = 0
    for lo in range(0, total, step):
        x = in_src[lo:lo + step].to(torch.int64)
        s1 += int(x.sum())
        s2 += int(((x & 4095) ** 2).sum())
    t = torch.arange(n, dtype=torch.int64, device=in_src.device)
    assert s1 == d * (n * (n - 1) // 2)
    assert s2 == d * int(((t & 4095) ** 2).sum())
    del in_ptr, in_src, t
    torch.cuda.empty_cache()


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
