"""GPU parity on the benchmark's own mixes, at their real sizes.

Batched ensemble (the bench's 4096 graphs: n log-uniform in [1e3, 1e7], d
1..16, p in [0.05, 0.95], seed 1000 + i), one captured launch:
  * every graph with n <= 1e5 bit for bit against the sequential oracle;
  * every row of every graph with 1e5 < n <= 1e6 through the oracle's
    row-consistency verifier (all rows passing is bit-exactness, by induction
    over t);
  * every row of every graph with n > 1e6 through the verifier as well, so
    the whole ensemble is checked;
  * every graph: targets of t in [0, t), distinct within a row, the seed
    clique, row_ptr = targets_offset + t*d, and its completion flag.
The served mix is checked in test_gpu_configs_served.py (its 200 GB pool
needs the ensemble's buffers freed first).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
SMALL_N = 100_000
MID_N = 1_000_000


@pytest.fixture(scope="module")
def ensemble(ctx):
    import torch
    import cuppa
    from cuppa import workloads
    torch.cuda.empty_cache()
    specs = workloads.batch_specs(4096)
    descs, tot, rot = cuppa.batch_layout(specs)
    print(f"\n[ensemble] {len(specs)} graphs, {tot} slots", flush=True)
    targets = torch.empty(tot, dtype=torch.int32, device="cuda")
    row_ptr = torch.empty(rot, dtype=torch.int64, device="cuda")
    b = cuppa.Batch(ctx, descs, targets, row_ptr)
    b.run()
    done = b.completion_ms()
    b.close()
    st = ctx.stats()
    yield specs, descs, targets, row_ptr, st, done
    del targets, row_ptr
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
def test_ensemble_counts_and_flags(ensemble):
    specs, _, _, _, st, done = ensemble
    assert st["nodes"] == sum(n - d - 1 for n, d, _, _ in specs)
    assert all(x >= 0 for x in done)


@pytest.mark.timeout(900)
def test_ensemble_structure_every_graph(ensemble):
    import torch
    specs, descs, targets, row_ptr, _, _ = ensemble
    for (n, d, _, _), ds in zip(specs, descs):
        rows = targets[ds.targets_offset:ds.targets_offset + n * d].view(n, d)
        clique = torch.tensor([[j + (j >= i) for j in range(d)] for i in range(d + 1)], dtype=torch.int32,
                              device=rows.device)
        assert torch.equal(rows[:d + 1], clique), (n, d)
        gen = rows[d + 1:]
        t = torch.arange(d + 1, n, device=rows.device, dtype=torch.int64).unsqueeze(1)
        assert bool((gen >= 0).all()) and bool((gen.to(torch.int64) < t).all()), (n, d)
        if d > 1:
            s, _ = torch.sort(gen, dim=1)
            assert bool((s[:, 1:] != s[:, :-1]).all()), (n, d)
        rp = row_ptr[ds.row_ptr_offset:ds.row_ptr_offset + n + 1]
        want = ds.targets_offset + torch.arange(n + 1, device=rp.device, dtype=torch.int64) * d
        assert torch.equal(rp, want), (n, d)


@pytest.mark.timeout(900)
def test_ensemble_small_graphs_bit_exact(ensemble, oracle_mod):
    specs, descs, targets, _, _, _ = ensemble
    small = [(s, ds) for s, ds in zip(specs, descs) if s[0] <= SMALL_N]
    print(f"\n[ensemble] {len(small)} graphs with n <= {SMALL_N} against the oracle", flush=True)
    got = {i: targets[ds.targets_offset:ds.targets_offset + s[0] * s[1]].cpu().numpy()
           for i, (s, ds) in enumerate(small)}

    def check(i):
        n, d, T, seed = small[i][0]
        want, _, _ = oracle_mod.generate(n, d, T, seed)
        return bool(np.array_equal(got[i], want))

    with ThreadPoolExecutor(THREADS) as pool:
        ok = list(pool.map(check, range(len(small))))
    bad = [small[i][0] for i, g in enumerate(ok) if not g]
    assert not bad, bad[:5]


@pytest.mark.timeout(900)
def test_ensemble_mid_graphs_every_row(ensemble, oracle_mod):
    import torch
    specs, descs, targets, _, _, _ = ensemble
    mid = [(s, ds) for s, ds in zip(specs, descs) if SMALL_N < s[0] <= MID_N]
    print(f"\n[ensemble] {len(mid)} graphs with {SMALL_N} < n <= {MID_N}: every row", flush=True)
    host = torch.empty(MID_N * 16, dtype=torch.int32, pin_memory=True)
    for (n, d, T, seed), ds in mid:
        m = n * d
        host[:m].copy_(targets[ds.targets_offset:ds.targets_offset + m])
        bad = oracle_mod.verify_rows(n, d, T, seed, host[:m].numpy(), threads=THREADS)
        assert bad is None, (n, d, T, seed, bad)


@pytest.mark.timeout(900)
def test_ensemble_large_graphs_every_row(ensemble, oracle_mod):
    import torch
    specs, descs, targets, _, _, _ = ensemble
    large = [(s, ds) for s, ds in zip(specs, descs) if s[0] > MID_N]
    print(f"\n[ensemble] {len(large)} graphs with n > {MID_N}: every row", flush=True)
    host = torch.empty(max(s[0] * s[1] for s, _ in large), dtype=torch.int32, pin_memory=True)
    for (n, d, T, seed), ds in large:
        m = n * d
        host[:m].copy_(targets[ds.targets_offset:ds.targets_offset + m])
        bad = oracle_mod.verify_rows(n, d, T, seed, host[:m].numpy(), threads=THREADS)
        assert bad is None, (n, d, T, seed, bad)
