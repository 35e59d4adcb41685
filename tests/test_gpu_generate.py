"""GPU parity: the HIP generator against the CPU oracle, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu(ctx, n, d, T, seed):
    import cuppa
    g = ctx.generate(n, d, p_thresh=T, seed=seed)
    return g.targets.cpu().numpy(), g.row_ptr.cpu().numpy()


CASES = [
    # n, d, p, seed
    (2, 1, 0.5, 1),          # clique only
    (3, 1, 0.5, 1),
    (6, 4, 0.5, 3),          # one generated node
    (1000, 1, 0.0, 5),
    (1000, 1, 1.0, 5),
    (5000, 3, 0.0, 9),       # copies only: every target in the seed clique
    (20000, 8, 0.5, 7),
    (20000, 16, 0.2, 11),
    (50000, 2, 0.9, 123),
    (100000, 5, 0.35, 2**40 + 17),
]


@pytest.mark.parametrize("n,d,p,seed", CASES)
def test_small_graphs_bit_exact(ctx, oracle_mod, n, d, p, seed):
    T = oracle_mod.p_threshold(p)
    want, _, _ = oracle_mod.generate(n, d, T, seed)
    got, rp = _gpu(ctx, n, d, T, seed)
    assert np.array_equal(got, want)
    assert np.array_equal(rp, np.arange(n + 1, dtype=np.int64) * d)
    st = ctx.stats()
    assert st["nodes"] == n - d - 1


def test_n1e6_d4_bit_exact(ctx, oracle_mod):
    n, d, p, seed = 1_000_000, 4, 0.5, 42
    T = oracle_mod.p_threshold(p)
    want, _, _ = oracle_mod.generate(n, d, T, seed)
    got, _ = _gpu(ctx, n, d, T, seed)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tuning", [
    dict(block_threads=64, chunk_nodes=64),
    dict(block_threads=256, chunk_nodes=1024),
    dict(block_threads=128, chunk_nodes=128, max_blocks=1),
    dict(block_threads=64, chunk_nodes=64, max_blocks=3),
    dict(block_threads=256, chunk_nodes=512, blocks_per_cu=1),
    dict(early_nodes=1),             # every row through the LDS-staged publication
    dict(early_nodes=1 << 30),       # every slot published as accepted
    dict(early_nodes=1024, chunk_nodes=64, max_blocks=2),
    dict(prefix_nodes=1),            # no LDS-resident prefix: the persistent kernel from node d+1
    dict(prefix_nodes=300),          # a prefix shorter than the default
])
def test_launch_shape_independent(ctx, oracle_mod, tuning):
    n, d, p, seed = 200_000, 8, 0.2, 99
    T = oracle_mod.p_threshold(p)
    want, _, _ = oracle_mod.generate(n, d, T, seed)
    try:
        ctx.set_tuning(watchdog_ms=20000, **tuning)
        got, _ = _gpu(ctx, n, d, T, seed)
    finally:
        ctx.set_tuning(watchdog_ms=20000)
    assert np.array_equal(got, want)


def test_degree_histogram(ctx, oracle_mod):
    n, d, p, seed = 300_000, 6, 0.4, 5
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    deg_want, _, mx_want = oracle_mod.degrees(n, d, G)
    hist_want = np.bincount(deg_want.astype(np.int64))
    import cuppa
    g = ctx.generate(n, d, p_thresh=T, seed=seed)
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    assert mx == mx_want
    assert np.array_equal(deg.cpu().numpy().astype(np.uint64), deg_want)
    assert np.array_equal(hist.cpu().numpy(), hist_want)


def test_async_launch_owns_the_context(ctx, oracle_mod):
    """While an async launch is pending, calls that share the context's
    scratch return PPA_ERR_AGAIN instead of racing it; ppa_ctx_sync ends it."""
    import ctypes
    import torch
    import cuppa
    n, d, p, seed = 300_000, 6, 0.4, 5
    T = oracle_mod.p_threshold(p)
    desc = cuppa.make_desc(n, d, p_thresh=T, seed=seed)
    t1 = torch.empty(n * d, dtype=torch.int32, device="cuda")
    t2 = torch.empty(n * d, dtype=torch.int32, device="cuda")
    L = cuppa.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.ppa_generate_async(ctx.handle, ctypes.byref(desc), ctypes.c_void_p(t1.data_ptr()), None, s) == 0
    assert L.ppa_generate_async(ctx.handle, ctypes.byref(desc), ctypes.c_void_p(t2.data_ptr()), None, s) == -9
    assert L.ppa_degree_histogram(ctx.handle, ctypes.byref(desc), ctypes.c_void_p(t1.data_ptr()), None, None, 0,
                                  None, s) == -9
    assert L.ppa_ctx_sync(ctx.handle, s) == 0
    want, _, _ = oracle_mod.generate(n, d, T, seed)
    assert np.array_equal(t1.cpu().numpy(), want)
    assert L.ppa_generate(ctx.handle, ctypes.byref(desc), ctypes.c_void_p(t2.data_ptr()), None, s) == 0
    assert np.array_equal(t2.cpu().numpy(), want)


@pytest.mark.parametrize("n,d,p,seed", [
    (2500, 16, 0.1, 1),    # exactly the 160 KB prefix (40000 / 16 nodes)
    (2501, 16, 0.1, 2),    # one node for the persistent kernel
    (1020, 16, 0.5, 9),
    (4444, 9, 0.2, 10),    # d = 9: the smallest d of the 160 KB prefix
    (2040, 8, 0.5, 3),     # exactly the 64 KB prefix (16320 / 8 nodes)
    (16320, 1, 0.3, 4),    # d = 1: 16320 nodes in LDS, 16 per lane
    (16321, 1, 0.3, 5),
    (3000, 5, 0.0, 6),     # copies only
    (5000, 3, 1.0, 7),     # direct only
    (100_000, 16, 0.1, 8),
    (1250, 32, 0.1, 11),   # exactly the 160 KB prefix at d = 32 (32-lane groups)
    (1251, 32, 0.1, 12),
    (2352, 17, 0.2, 13),   # d = 17: 40000 / 17 nodes
])
def test_lds_prefix_boundaries(ctx, oracle_mod, n, d, p, seed):
    """The prefix kernel's nodes and the persistent kernel's continue the same
    graph: with the prefix as large as LDS allows (40000 / d nodes for
    d >= 9, 16320 / d below), the default, and none."""
    T = oracle_mod.p_threshold(p)
    want, _, want_retries = oracle_mod.generate(n, d, T, seed)
    for off in (40000, 0, 1):
        ctx.set_tuning(watchdog_ms=20000, prefix_nodes=off)
        try:
            got, rp = _gpu(ctx, n, d, T, seed)
            st = ctx.stats()
        finally:
            ctx.set_tuning(watchdog_ms=20000)
        assert np.array_equal(got, want), off
        assert st["nodes"] == n - d - 1 and st["retries"] == want_retries


def test_watchdog_turns_a_long_run_into_a_timeout(ctx, oracle_mod):
    """A 1 ms budget on a graph that takes ~10 ms: the kernels see the
    expired budget, every wavefront exits, the call returns PPA_ERR_TIMEOUT,
    and the context generates correctly afterwards."""
    import cuppa
    n, d, p, seed = 20_000_000, 16, 0.1, 3
    ctx.set_tuning(watchdog_ms=1)
    try:
        with pytest.raises(cuppa.PPAError) as e:
            ctx.generate(n, d, p, seed=seed)
        assert e.value.code == -5  # PPA_ERR_TIMEOUT
    finally:
        ctx.set_tuning(watchdog_ms=20000)
    T = oracle_mod.p_threshold(p)
    want, _, _ = oracle_mod.generate(50_000, d, T, seed)
    got, _ = _gpu(ctx, 50_000, d, T, seed)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,d,p,seed,cuts", [
    (200_000, 8, 0.2, 3, [1_000, 50_000, 120_000]),
    (60_000, 16, 0.1, 4, [500, 30_000]),     # the first range ends inside the LDS prefix
    (10_000, 4, 0.5, 5, [5]),                # the first range is the seed clique alone
    (10_000, 8, 0.5, 7, [3, 6, 20]),         # ranges inside the seed clique generate nothing
    (300_000, 1, 0.3, 6, [100_000]),
])
def test_generate_in_ranges(ctx, oracle_mod, n, d, p, seed, cuts):
    """ppa_generate_range: consecutive node ranges give ppa_generate's graph,
    and after each range's sync the rows below its end are final."""
    import torch
    import cuppa
    T = oracle_mod.p_threshold(p)
    want, _, _ = oracle_mod.generate(n, d, T, seed)
    desc = cuppa.make_desc(n, d, None, seed, T)
    targets = torch.full((n * d,), -7, dtype=torch.int32, device="cuda")
    rp = torch.full((n + 1,), -7, dtype=torch.int64, device="cuda")
    bounds = [0] + cuts + [n]
    for a, b in zip(bounds, bounds[1:]):
        ctx.generate_range(desc, targets, rp, a, b)
        ctx.sync()
        assert ctx.stats()["nodes"] == max(0, b - max(a, d + 1)), (a, b)
        assert np.array_equal(targets[:b * d].cpu().numpy(), want[:b * d]), (a, b)
    assert np.array_equal(rp.cpu().numpy(), np.arange(n + 1, dtype=np.int64) * d)


def test_generate_range_must_continue(ctx):
    """A range that does not start where the previous one ended, or follows
    a completed graph, is rejected before anything is launched."""
    import torch
    import cuppa
    n, d = 50_000, 6
    desc = cuppa.make_desc(n, d, 0.3, 9)
    targets = torch.empty(n * d, dtype=torch.int32, device="cuda")
    bad = []
    for first, last in ((0, 0), (10, 5), (0, n + 1)):
        with pytest.raises(cuppa.PPAError) as e:
            ctx.generate_range(desc, targets, None, first, last)
        bad.append(e.value.code)
    ctx.generate_range(desc, targets, None, 0, 20_000)
    ctx.sync()
    with pytest.raises(cuppa.PPAError) as e:       # a gap
        ctx.generate_range(desc, targets, None, 20_001, n)
    bad.append(e.value.code)
    other = cuppa.make_desc(n, d, 0.3, 10)
    with pytest.raises(cuppa.PPAError) as e:       # another graph
        ctx.generate_range(other, targets, None, 20_000, n)
    bad.append(e.value.code)
    ctx.generate_range(desc, targets, None, 20_000, n)
    ctx.sync()
    with pytest.raises(cuppa.PPAError) as e:       # the graph is complete
        ctx.generate_range(desc, targets, None, n - 1, n)
    bad.append(e.value.code)
    assert bad == [-1] * 6


def test_failed_range_is_not_continued(ctx, oracle_mod):
    """A range whose kernel ends with an error leaves incomplete rows, so
    the next range of that graph is refused; the context then generates a
    new graph, range by range, as the oracle does. The failure is the
    kernels' own deadline (1 ms against ~25 ms of work), which every
    wavefront reads within 64 passes, so each kernel exits normally."""
    import torch
    import cuppa
    n, d = 1 << 26, 16
    desc = cuppa.make_desc(n, d, 0.1, 12)
    targets = torch.empty(n * d, dtype=torch.int32, device="cuda")
    ctx.set_tuning(watchdog_ms=1)
    try:
        ctx.generate_range(desc, targets, None, 0, n // 2)
        with pytest.raises(cuppa.PPAError) as e:
            ctx.sync()
        assert e.value.code == -5  # PPA_ERR_TIMEOUT
    finally:
        ctx.set_tuning(watchdog_ms=20000)
    with pytest.raises(cuppa.PPAError) as e:
        ctx.generate_range(desc, targets, None, n // 2, n)
    assert e.value.code == -1  # PPA_ERR_INVALID
    del targets
    m, T, seed = 30_000, oracle_mod.p_threshold(0.3), 13
    want, _, _ = oracle_mod.generate(m, 8, T, seed)
    small = cuppa.make_desc(m, 8, None, seed, T)
    out = torch.full((m * 8,), -7, dtype=torch.int32, device="cuda")
    for a, b in ((0, 12_000), (12_000, m)):
        ctx.generate_range(small, out, None, a, b)
        ctx.sync()
    assert np.array_equal(out.cpu().numpy(), want)
