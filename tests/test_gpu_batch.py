"""Batched ensembles in one captured persistent launch, and the autotuner.

Every graph of a batch must equal the oracle's graph for its own (n, d, p,
seed), whatever the batch around it, the launch shape and the number of
times the captured hipGraph is replayed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(oracle_mod, count, seed=0, nmax=60_000):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        d = int(rng.integers(1, 17))
        n = int(rng.integers(d + 1, nmax)) if i % 7 else d + 1 + int(rng.integers(0, 3))  # tiny graphs too
        p = float(rng.uniform(0.0, 1.0))
        specs.append((n, d, oracle_mod.p_threshold(p), 1000 + i))
    return specs


def _alloc(specs):
    import torch
    import cuppa
    descs, tot, rot = cuppa.batch_layout(specs)
    targets = torch.full((tot,), -7, dtype=torch.int32, device="cuda")
    row_ptr = torch.full((rot,), -7, dtype=torch.int64, device="cuda")
    return descs, targets, row_ptr


def _check(oracle_mod, specs, descs, targets, row_ptr):
    tg = targets.cpu().numpy()
    rp = row_ptr.cpu().numpy() if row_ptr is not None else None
    for (n, d, T, seed), ds in zip(specs, descs):
        want, _, _ = oracle_mod.generate(n, d, T, seed)
        got = tg[ds.targets_offset:ds.targets_offset + n * d]
        assert np.array_equal(got, want), (n, d, T, seed)
        if rp is not None:
            r = rp[ds.row_ptr_offset:ds.row_ptr_offset + n + 1]
            assert np.array_equal(r, ds.targets_offset + np.arange(n + 1, dtype=np.int64) * d)


def test_batch_matches_oracle_per_graph(ctx, oracle_mod):
    import cuppa
    specs = _specs(oracle_mod, 60)
    descs, targets, row_ptr = _alloc(specs)
    b = cuppa.Batch(ctx, descs, targets, row_ptr)
    b.run()
    b.close()
    st = ctx.stats()
    assert st["nodes"] == sum(n - d - 1 for n, d, _, _ in specs)
    _check(oracle_mod, specs, descs, targets, row_ptr)


def test_batch_graph_replays(ctx, oracle_mod):
    import torch
    import cuppa
    specs = _specs(oracle_mod, 24, seed=3)
    descs, targets, row_ptr = _alloc(specs)
    b = cuppa.Batch(ctx, descs, targets, row_ptr)
    for rep in range(4):
        targets.fill_(-3)
        row_ptr.fill_(-3)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        _check(oracle_mod, specs, descs, targets, row_ptr)
    b.close()


@pytest.mark.parametrize("tuning", [dict(chunk_nodes=64, block_threads=64), dict(chunk_nodes=1024),
                                    dict(max_blocks=1, block_threads=128), dict(blocks_per_cu=1, early_nodes=1),
                                    dict(prefix_nodes=1),        # no LDS-resident prefixes
                                    dict(prefix_nodes=16320)])   # every graph's prefix as large as LDS holds
def test_batch_launch_shape_independent(ctx, oracle_mod, tuning):
    import cuppa
    specs = _specs(oracle_mod, 20, seed=5, nmax=30_000)
    descs, targets, row_ptr = _alloc(specs)
    try:
        ctx.set_tuning(watchdog_ms=20000, **tuning)
        b = cuppa.Batch(ctx, descs, targets, None)
        b.run()
        b.close()
    finally:
        ctx.set_tuning(watchdog_ms=20000)
    _check(oracle_mod, specs, descs, targets, None)


def test_one_shot_generate_batch(ctx, oracle_mod):
    import ctypes
    import cuppa
    specs = _specs(oracle_mod, 8, seed=9)
    descs, targets, row_ptr = _alloc(specs)
    rc = cuppa.lib().ppa_generate_batch(ctx.handle, descs, len(descs), ctypes.c_void_p(targets.data_ptr()),
                                        ctypes.c_void_p(row_ptr.data_ptr()), cuppa._stream_handle(None))
    assert rc == 0
    _check(oracle_mod, specs, descs, targets, row_ptr)


def test_batch_rejects_overlap(ctx, oracle_mod):
    import ctypes
    import cuppa
    specs = _specs(oracle_mod, 3, seed=1)
    descs, targets, row_ptr = _alloc(specs)
    descs[1].targets_offset = descs[0].targets_offset + 1
    h = ctypes.c_void_p()
    rc = cuppa.lib().ppa_batch_create(ctx.handle, descs, len(descs), ctypes.c_void_p(targets.data_ptr()), None,
                                      ctypes.byref(h))
    assert rc == -1


def test_autotune_single_and_batch(ctx, oracle_mod):
    import cuppa
    import torch
    n, d, p, seed = 200_000, 8, 0.2, 4
    T = oracle_mod.p_threshold(p)
    descs, targets, row_ptr = _alloc([(n, d, T, seed)])
    before = ctx.get_tuning()
    try:
        best = cuppa.autotune(ctx, descs, targets, row_ptr)
        assert best["chunk_nodes"] in (0, 64, 128, 256, 512, 1024)
        again = cuppa.autotune(ctx, descs, targets, row_ptr)  # cached
        assert again == {**best, "watchdog_ms": before["watchdog_ms"]}
        ctx.generate_into(descs[0], targets, row_ptr)
        _check(oracle_mod, [(n, d, T, seed)], descs, targets, row_ptr)
        specs = _specs(oracle_mod, 16, seed=11, nmax=20_000)
        bd, bt, br = _alloc(specs)
        cuppa.autotune(ctx, bd, bt, br)
        b = cuppa.Batch(ctx, bd, bt, br)
        b.run()
        b.close()
        _check(oracle_mod, specs, bd, bt, br)
    finally:
        ctx.set_tuning(watchdog_ms=20000)


def test_batch_launch_refused_while_a_launch_is_pending(ctx, oracle_mod):
    """A batch launch resets the context's scratch, so it must not start under
    a pending ppa_generate_async: PPA_ERR_AGAIN, and the pending graph still
    comes out bit for bit."""
    import ctypes
    import torch
    import cuppa
    n, d, p, seed = 300_000, 8, 0.2, 21
    T = oracle_mod.p_threshold(p)
    desc = cuppa.make_desc(n, d, None, seed, T)
    targets = torch.empty(n * d, dtype=torch.int32, device="cuda")
    specs = _specs(oracle_mod, 6, seed=13)
    descs, bt, br = _alloc(specs)
    b = cuppa.Batch(ctx, descs, bt, br)
    try:
        ctx.generate_into(desc, targets, None, sync=False)
        assert cuppa.lib().ppa_batch_launch(b._h, cuppa._stream_handle(None)) == -9
        rc = cuppa.lib().ppa_autotune(ctx.handle, descs, len(descs), ctypes.c_void_p(bt.data_ptr()),
                                      ctypes.c_void_p(br.data_ptr()), cuppa._stream_handle(None),
                                      ctypes.byref(cuppa.Tuning()))
        assert rc == -9
        ctx.sync()
        want, _, _ = oracle_mod.generate(n, d, T, seed)
        assert np.array_equal(targets.cpu().numpy(), want)
        b.run()  # free again
        _check(oracle_mod, specs, descs, bt, br)
    finally:
        b.close()


def test_batch_completion_times(ctx, oracle_mod):
    """Every graph's completion is observed, no later than the launch's end."""
    import time
    import cuppa
    specs = _specs(oracle_mod, 40, seed=17)
    descs, targets, row_ptr = _alloc(specs)
    b = cuppa.Batch(ctx, descs, targets, row_ptr)
    t0 = time.perf_counter()
    b.run()
    wall = (time.perf_counter() - t0) * 1e3
    done = b.completion_ms()
    b.close()
    assert len(done) == len(specs)
    assert all(0.0 <= x <= wall + 1.0 for x in done), (min(done), max(done), wall)
    _check(oracle_mod, specs, descs, targets, row_ptr)
