"""Out-degrees 17..32 (the lane-per-node generator's largest): single graphs through the LDS prefix
(32-lane groups, duplicate checks through ds_bpermute) and the persistent
generator, bit for bit against the oracle; their degree pass, in-edge and
undirected CSR and ranged generation. Batches refuse them, and the server
runs them as lone launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d,p,seed", [
    (18, 17, 0.5, 1),           # the seed clique and one generated node
    (3_000, 32, 0.0, 2),        # copies only: many redraws among the first nodes
    (30_000, 17, 0.3, 3),
    (20_000, 24, 0.1, 4),
    (50_000, 32, 0.5, 5),
    (500_000, 20, 0.2, 6),
])
def test_wide_d_bit_exact(ctx, oracle_mod, n, d, p, seed):
    T = oracle_mod.p_threshold(p)
    want, _, want_retries = oracle_mod.generate(n, d, T, seed)
    g = ctx.generate(n, d, p_thresh=T, seed=seed)
    assert np.array_equal(g.targets.cpu().numpy(), want)
    assert np.array_equal(g.row_ptr.cpu().numpy(), np.arange(n + 1, dtype=np.int64) * d)
    st = ctx.stats()
    assert st["nodes"] == n - d - 1 and st["retries"] == want_retries


@pytest.mark.parametrize("n,d,p,seed,tuning", [
    # every row through the LDS-staged publication (no slot-by-slot early
    # phase): 4-byte pieces for odd d, 16-byte pieces with 5..8 per lane for
    # d = 20..32
    (60_000, 3, 0.3, 31, dict(early_nodes=1)),
    (60_000, 5, 0.1, 32, dict(early_nodes=1)),
    (40_000, 17, 0.3, 33, dict(early_nodes=1)),
    (40_000, 20, 0.2, 34, dict(early_nodes=1)),
    (30_000, 31, 0.5, 35, dict(early_nodes=1)),
    (30_000, 32, 0.1, 36, dict(early_nodes=1)),
    # one workgroup: rows finish far apart, so the three LDS flush windows
    # overflow and the rest are stored by their own lanes
    (20_000, 32, 0.2, 37, dict(early_nodes=1, max_blocks=1, block_threads=64)),
    (20_000, 7, 0.2, 38, dict(early_nodes=1, max_blocks=1, block_threads=64)),
])
def test_staged_publication_bit_exact(ctx, oracle_mod, n, d, p, seed, tuning):
    T = oracle_mod.p_threshold(p)
    want, _, want_retries = oracle_mod.generate(n, d, T, seed)
    try:
        ctx.set_tuning(watchdog_ms=20000, **tuning)
        g = ctx.generate(n, d, p_thresh=T, seed=seed)
        st = ctx.stats()
    finally:
        ctx.set_tuning(watchdog_ms=20000)
    assert np.array_equal(g.targets.cpu().numpy(), want)
    assert st["nodes"] == n - d - 1 and st["retries"] == want_retries
    # the staged path published rows
    assert st["staged_rows"] > 0


def test_wide_d_barabasi_albert(ctx, oracle_mod):
    import cuppa
    n, d, seed = 40_000, 24, 7
    g = ctx.generate(n, d, seed=seed, model=cuppa.MODEL_BA)
    want, _ = oracle_mod.generate_ba(n, d, seed)
    assert np.array_equal(g.targets.cpu().numpy(), want)


def test_wide_d_degrees_and_csr_forms(ctx, oracle_mod):
    n, d, p, seed = 60_000, 28, 0.4, 8
    T = oracle_mod.p_threshold(p)
    g = ctx.generate(n, d, p_thresh=T, seed=seed)
    G = g.targets.cpu().numpy()
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    wdeg, _, wmx = oracle_mod.degrees(n, d, G)
    assert np.array_equal(deg.cpu().numpy().astype(np.uint64), wdeg)
    assert mx == wmx
    assert np.array_equal(hist.cpu().numpy(), np.bincount(wdeg.astype(np.int64), minlength=len(hist)))
    in_ptr, in_src = ctx.in_csr(g)
    want_ptr, want_src = oracle_mod.in_csr(n, d, G)
    assert np.array_equal(in_ptr.cpu().numpy(), want_ptr)
    assert np.array_equal(in_src.cpu().numpy(), want_src)
    ptr, adj = ctx.undirected_csr(g)
    want_ptr, want_adj = oracle_mod.undirected_csr(n, d, G)
    assert np.array_equal(ptr.cpu().numpy(), want_ptr)
    assert np.array_equal(adj.cpu().numpy(), want_adj)


def test_wide_d_in_ranges(ctx, oracle_mod):
    import torch
    import cuppa
    n, d, p, seed = 40_000, 32, 0.2, 9
    T = oracle_mod.p_threshold(p)
    want, _, _ = oracle_mod.generate(n, d, T, seed)
    desc = cuppa.make_desc(n, d, None, seed, T)
    targets = torch.full((n * d,), -7, dtype=torch.int32, device="cuda")
    for a, b in ((0, 10_000), (10_000, n)):
        ctx.generate_range(desc, targets, None, a, b)
        ctx.sync()
        assert np.array_equal(targets[:b * d].cpu().numpy(), want[:b * d]), (a, b)


def test_batches_refuse_wide_d(ctx, oracle_mod):
    import torch
    import cuppa
    T = oracle_mod.p_threshold(0.5)
    descs, tot, rot = cuppa.batch_layout([(1_000, 8, T, 1), (1_000, 17, T, 2)])
    targets = torch.empty(tot, dtype=torch.int32, device="cuda")
    row_ptr = torch.empty(rot, dtype=torch.int64, device="cuda")
    with pytest.raises(cuppa.PPAError) as e:
        cuppa.Batch(ctx, descs, targets, row_ptr)
    assert e.value.code == -1  # PPA_ERR_INVALID


def test_server_runs_wide_d_alone(ctx, oracle_mod):
    """With solo_nodes above every request, requests of d <= 16 are batched
    and those of d > 16 launch alone; all come out as the oracle's graphs."""
    import cuppa
    rng = np.random.default_rng(11)
    specs = []
    for i in range(24):
        d = int(rng.choice([4, 8, 16, 20, 32]))
        n = int(rng.integers(d + 2, 20_000))
        specs.append((i, n, d, oracle_mod.p_threshold(0.3), 700 + i))
    srv = cuppa.Server(ctx, budget_bytes=1 << 30, streams=4, batch_nodes=1 << 16, solo_nodes=1 << 20)
    for s in specs:
        assert srv.submit(*s) == 0
    pending = {s[0]: s for s in specs}
    for _ in range(2_000_000):
        srv.step()
        for rid in list(pending):
            r = srv.poll(rid)
            if r.state in (cuppa.REQ_DONE, cuppa.REQ_FAILED):
                assert r.state == cuppa.REQ_DONE, (rid, r.status)
                _, n, d, T, seed = pending.pop(rid)
                got = cuppa.read_device(r.targets, n * d, np.int32)
                want, _, _ = oracle_mod.generate(n, d, T, seed)
                assert np.array_equal(got, want), (rid, n, d)
                assert srv.release(rid) == 0
        if not pending:
            break
    srv.close()
    assert not pending
