"""In-edge CSR on the GPU (ppa_in_csr) against the oracle's counting sort:
generated graphs of one to four radix passes (n up to 2^24 + 5), and a
synthetic target array with out-of-range values that must be skipped."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d,p,seed", [
    (200, 3, 0.5, 1),            # one 8-bit pass
    (60_000, 5, 0.2, 2),         # two passes
    (3_000_000, 8, 0.5, 3),      # three passes
    ((1 << 24) + 5, 2, 0.1, 4),  # four passes
])
def test_in_csr_matches_oracle(ctx, oracle_mod, n, d, p, seed):
    T = oracle_mod.p_threshold(p)
    g = ctx.generate(n, d, p_thresh=T, seed=seed)
    G = g.targets.cpu().numpy()
    want_ptr, want_src = oracle_mod.in_csr(n, d, G)
    ptr, src = ctx.in_csr(g)
    assert np.array_equal(ptr.cpu().numpy(), want_ptr)
    assert np.array_equal(src.cpu().numpy(), want_src)


def test_in_csr_skips_out_of_range(ctx, oracle_mod):
    import torch
    import cuppa
    n, d = 700_000, 6
    rng = np.random.default_rng(9)
    G = (np.arange(n * d) // d * rng.random(n * d) ** 3).astype(np.int32)
    bad = rng.integers(0, n * d, 40)
    G[bad[:20]] = -1
    G[bad[20:]] = n + 3
    g = cuppa.Graph(n, d, torch.from_numpy(G).cuda(), None)
    want_ptr, want_src = oracle_mod.in_csr(n, d, G)
    ptr, src = ctx.in_csr(g)
    assert np.array_equal(ptr.cpu().numpy(), want_ptr)
    assert np.array_equal(src.cpu().numpy(), want_src)


@pytest.mark.parametrize("n,d,p,seed", [(300, 4, 0.5, 1), (70_000, 16, 0.1, 2), (2_500_000, 3, 0.3, 3)])
def test_undirected_matches_oracle(ctx, oracle_mod, n, d, p, seed):
    g = ctx.generate(n, d, p_thresh=oracle_mod.p_threshold(p), seed=seed)
    G = g.targets.cpu().numpy()
    want_ptr, want_adj = oracle_mod.undirected_csr(n, d, G)
    ptr, adj = ctx.undirected_csr(g)
    assert np.array_equal(ptr.cpu().numpy(), want_ptr)
    assert np.array_equal(adj.cpu().numpy(), want_adj)


def test_undirected_rejects_incomplete(ctx):
    import torch
    import cuppa
    n, d = 1000, 2
    G = np.zeros(n * d, np.int32)
    G[5] = -1
    with pytest.raises(cuppa.PPAError):
        ctx.undirected_csr(cuppa.Graph(n, d, torch.from_numpy(G).cuda(), None))


@pytest.mark.parametrize("case", ["few_values", "all_zero", "none_valid"])
def test_in_csr_wide_gaps(ctx, oracle_mod, case):
    """Targets with few distinct values leave long runs of nodes no slot
    names; the whole wavefront fills those runs of in_ptr."""
    import torch
    import cuppa
    n, d = 3_000_000, 2
    rng = np.random.default_rng(21)
    if case == "few_values":
        G = rng.choice(np.array([0, 17, 50, 1000, 999_999, n - 1], np.int32), n * d)
    elif case == "all_zero":
        G = np.zeros(n * d, np.int32)
    else:
        G = np.full(n * d, -1, np.int32)
    g = cuppa.Graph(n, d, torch.from_numpy(G).cuda(), None)
    want_ptr, want_src = oracle_mod.in_csr(n, d, G)
    ptr, src = ctx.in_csr(g)
    assert np.array_equal(ptr.cpu().numpy(), want_ptr)
    total = int(want_ptr[-1])
    assert np.array_equal(src.cpu().numpy()[:total], want_src[:total])
