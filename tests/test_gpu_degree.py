"""Degree pass on the GPU against numpy counting, across the pass's regimes:
no partition (n <= 2^15), one partition level (n <= 2^23), two levels (one
matrix count up to 2^30 nodes, two above), and leaves of more than 2^20 slots
that are counted in pieces. Target arrays are synthetic (heavy-tailed towards
node 0, like a preferential-attachment graph, with a few out-of-range values
that must be skipped), so the cases cost no generation time; generated graphs
are covered in test_gpu_generate and test_gpu_scale."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _targets(n, d, seed, power):
    rng = np.random.default_rng(seed)
    G = np.empty(n * d, np.int32)
    for i in range(d + 1):
        G[i * d:(i + 1) * d] = [j + (j >= i) for j in range(d)]
    t = np.arange(d + 1, n, dtype=np.int64)
    G[(d + 1) * d:] = (t[:, None] * rng.random((len(t), d)) ** power).astype(np.int32).ravel()
    if n * d > (d + 1) * d + 8:
        bad = rng.integers((d + 1) * d, n * d, 6)
        G[bad[:3]] = -1
        G[bad[3:]] = n + 5
    return G


@pytest.mark.parametrize("n,d,power", [
    (300, 4, 3),                 # one leaf, no partition
    (20000, 3, 3),
    (32768, 2, 3),               # exactly one leaf
    (32769, 2, 3),               # one partition level, two leaves
    (1 << 20, 2, 3),
    (3 << 20, 5, 3),
    ((1 << 22) + 77777, 2, 3),
    (1 << 23, 1, 3),             # one level, 256 buckets (the most)
    ((1 << 23) + 1, 1, 3),       # two levels, two buckets
    (9_000_000, 1, 3),
    ((1 << 24) + 5, 8, 4),       # two levels, hub leaves split into pieces
])
def test_degrees_match_counting(ctx, n, d, power):
    import torch
    import cuppa
    G = _targets(n, d, 7, power)
    g = cuppa.Graph(n, d, torch.from_numpy(G).cuda(), None)
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    v = G[(d + 1) * d:]
    v = v[(v >= 0) & (v < n)]
    want = d + np.bincount(v, minlength=n)
    got = deg.cpu().numpy().astype(np.int64)
    assert np.array_equal(got, want)
    assert mx == int(want.max())
    assert np.array_equal(hist.cpu().numpy(), np.bincount(want))


def test_degrees_above_2p30_nodes(ctx):
    """More than 2^30 nodes: 129 level-1 buckets, counted by two matrix
    launches of at most 128 buckets. The targets are drawn on the GPU and the
    expected degrees come from torch.bincount, so the case costs seconds."""
    import torch
    import cuppa
    n, d = (1 << 30) + (1 << 20) + 3, 1
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    t = torch.arange(n, dtype=torch.float64, device="cuda")
    u = torch.rand(n, dtype=torch.float64, device="cuda", generator=gen)
    G = (t * u ** 3).to(torch.int32)
    del t, u
    G[:2] = torch.tensor([1, 0], dtype=torch.int32, device="cuda")  # the seed rows of d = 1
    G[n - 1] = -1                                                  # skipped
    G[n - 7] = n + 11                                              # skipped
    g = cuppa.Graph(n, d, G, None)
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    v = G[2:].to(torch.int64)
    v = v[(v >= 0) & (v < n)]
    want = torch.bincount(v, minlength=n).to(torch.int64) + d
    del v
    assert torch.equal(deg.to(torch.int64), want)
    assert mx == int(want.max().item())
    assert torch.equal(hist.cpu(), torch.bincount(want).cpu())
    del G, g, deg, want
    ctx.trim()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("node", [0, 10_000, 5_000_000])
@pytest.mark.parametrize("extra", [7, 8])
def test_degree_past_32_bits_is_reported(ctx, node, extra):
    """Every slot names one node, so deg = d + (n - d - 1) * d = 8 * (n - 8)
    at d = 8: 2^32 - 8 at n = 2^29 + 7, which the 32-bit degree holds, and
    2^32 at n = 2^29 + 8, which it cannot: that call returns PPA_ERR_RANGE
    instead of the count modulo 2^32. Node 0 is one of the hub nodes the
    count pass counts itself; nodes 10,000 and 5,000,000 are counted in leaves
    split into pieces (added to the degree with atomics)."""
    import torch
    import cuppa
    d = 8
    n = (1 << 29) + extra
    G = torch.full((n * d,), node, dtype=torch.int32, device="cuda")
    g = cuppa.Graph(n, d, G, None)
    if extra == 8:
        with pytest.raises(cuppa.PPAError) as e:
            ctx.degree_histogram(g, hist_len=16)
        assert e.value.code == -11
    else:
        hist, mx, deg = ctx.degree_histogram(g, hist_len=16, want_degree=True)
        assert mx == (1 << 32) - 8
        assert int(deg[node].item()) & 0xFFFFFFFF == (1 << 32) - 8
        assert int(hist[d].item()) == n - 1  # every other node keeps its d out-slots
        del deg, hist
    del G, g
    ctx.trim()
    torch.cuda.empty_cache()


def test_histogram_of_degrees(ctx):
    """ppa_histogram_of_degrees over degrees left by ppa_degree_histogram:
    same histogram and maximum, truncated below hist_len, maximum alone with
    no histogram."""
    import ctypes
    import torch
    import cuppa
    n, d = 3 << 20, 5
    G = _targets(n, d, 11, 3)
    g = cuppa.Graph(n, d, torch.from_numpy(G).cuda(), None)
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    want = np.bincount(deg.cpu().numpy().astype(np.int64))
    assert np.array_equal(hist.cpu().numpy(), want) and mx == want.size - 1
    L = cuppa.lib()
    out_max = torch.zeros(1, dtype=torch.int64, device="cuda")
    for hist_len in (0, 7, want.size, want.size + 100):
        h = torch.full((max(hist_len, 1),), -1, dtype=torch.int64, device="cuda")
        rc = L.ppa_histogram_of_degrees(ctx.handle, n, ctypes.c_void_p(deg.data_ptr()),
                                        ctypes.c_void_p(h.data_ptr()) if hist_len else None, hist_len,
                                        ctypes.c_void_p(out_max.data_ptr()), None)
        assert rc == 0
        assert int(out_max.item()) == mx
        if hist_len:
            w = np.zeros(hist_len, np.int64)
            k = min(hist_len, want.size)
            w[:k] = want[:k]
            assert np.array_equal(h.cpu().numpy(), w)
    assert L.ppa_histogram_of_degrees(ctx.handle, n, None, None, 0, None, None) == -1


def test_scratch_budget_fallback_and_trim(ctx, oracle_mod):
    """A context whose scratch budget is below the partition buffers (one
    16-bit copy of the slots at this n; one int32 and one 16-bit copy from
    n > 2^23 on) counts with one atomic per slot instead, holds no work
    memory, and gets the same degrees; its in-edge CSR is refused with
    PPA_ERR_BUDGET. A context without a cap keeps the buffers until
    ppa_ctx_trim."""
    import torch
    import cuppa
    n, d, p, seed = 300_000, 8, 0.3, 17
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    deg_want, _, mx_want = oracle_mod.degrees(n, d, G)
    hist_want = np.bincount(deg_want.astype(np.int64))
    small = cuppa.Context(0, scratch_budget=n * d)
    try:
        assert small.memory() == {"held": 0, "budget": n * d}
        g = small.generate(n, d, p_thresh=T, seed=seed)
        hist, mx, deg = small.degree_histogram(g, want_degree=True)
        assert small.memory()["held"] == 0  # the atomic kernels ran: no partition buffers
        assert np.array_equal(deg.cpu().numpy().astype(np.uint64), deg_want)
        assert mx == mx_want
        assert np.array_equal(hist.cpu().numpy(), hist_want)
        with pytest.raises(cuppa.PPAError) as e:
            small.in_csr(g)
        assert e.value.code == -6  # PPA_ERR_BUDGET
    finally:
        small.close()
    g = ctx.generate(n, d, p_thresh=T, seed=seed)
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    assert ctx.memory()["held"] >= 2 * (n - d - 1) * d
    ctx.trim()
    assert ctx.memory()["held"] == 0
    hist2, mx2, deg2 = ctx.degree_histogram(g, want_degree=True)  # works again after the trim
    assert torch.equal(deg, deg2) and torch.equal(hist, hist2) and mx == mx2 == mx_want
