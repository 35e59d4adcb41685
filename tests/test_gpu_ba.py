"""GPU parity of the Barabasi-Albert model (PPA_MODEL_BA, `-A sba`): the HIP
generator against the oracle's restatement of the reference's SerialBA
(src/PA/SPA/SerialBA.cpp:68-187) in its endpoint-array form, bit for bit,
alone, in batches and in the server next to copy-model graphs; degrees as
PAHelper::outputSBADegreeDistribution counts them (PAHelper.hpp:69-147)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cuppa-public_amd", "bin", "ppa_runner")


def _gpu_ba(ctx, n, d, seed):
    import cuppa
    g = ctx.generate(n, d, seed=seed, model=cuppa.MODEL_BA)
    return g


@pytest.mark.parametrize("n,d,seed", [
    (2, 1, 0),                 # seed rows only
    (3, 1, 1),
    (5, 4, 2),
    (6, 4, 3),                 # one generated node
    (1000, 1, 5),              # a tree
    (20000, 8, 7),
    (20000, 16, 11),
    (100000, 5, 2**40 + 17),
    (1_000_000, 4, 42),
])
def test_ba_bit_exact(ctx, oracle_mod, n, d, seed):
    g = _gpu_ba(ctx, n, d, seed)
    want, _ = oracle_mod.generate_ba(n, d, seed)
    assert np.array_equal(g.targets.cpu().numpy(), want)
    assert np.array_equal(g.row_ptr.cpu().numpy(), np.arange(n + 1, dtype=np.int64) * d)
    assert ctx.stats()["nodes"] == n - d - 1


@pytest.mark.parametrize("tuning", [
    dict(block_threads=64, chunk_nodes=64, max_blocks=2),
    dict(early_nodes=1),
    dict(early_nodes=1 << 30),
])
def test_ba_launch_shape_independent(ctx, oracle_mod, tuning):
    n, d, seed = 200_000, 8, 99
    want, _ = oracle_mod.generate_ba(n, d, seed)
    ctx.set_tuning(watchdog_ms=20000, **tuning)
    try:
        g = _gpu_ba(ctx, n, d, seed)
    finally:
        ctx.set_tuning(watchdog_ms=20000)
    assert np.array_equal(g.targets.cpu().numpy(), want)


@pytest.mark.parametrize("n,d,seed", [(7, 4, 1), (50_000, 1, 2), (300_000, 16, 3), (3_000_000, 8, 4)])
def test_ba_degrees(ctx, oracle_mod, n, d, seed):
    g = _gpu_ba(ctx, n, d, seed)
    hist, mx, deg = ctx.degree_histogram(g, want_degree=True)
    G = g.targets.cpu().numpy()
    wdeg, whist, wmx = oracle_mod.degrees_ba(n, d, G, hist_len=mx + 1)
    assert np.array_equal(deg.cpu().numpy().astype(np.uint64), wdeg)
    assert np.array_equal(hist.cpu().numpy().astype(np.uint64), whist)
    assert mx == wmx
    assert int(wdeg.sum()) == 2 * d * (n - d)


@pytest.mark.parametrize("n,d,seed", [(1 << 24, 8, 5),    # no LDS prefix: 4 copy draws per node
                                      (1 << 22, 16, 6)])  # the 160 KB prefix: 8 per node
def test_ba_rows_consistent_at_scale(ctx, oracle_mod, n, d, seed):
    """Every row of a 2^22..2^24-node graph is what the model draws from the
    rows before it (the oracle's row check in copy-model form, pinned to the
    endpoint form on the CPU), and the degree sum is 2d(n-d)."""
    g = _gpu_ba(ctx, n, d, seed)
    G = g.targets.cpu().numpy()
    assert oracle_mod.verify_rows(n, d, 0, seed, G, threads=16, model=1) is None
    _, mx, deg = ctx.degree_histogram(g, want_degree=True)
    assert int(deg.cpu().numpy().astype(np.int64).sum()) == 2 * d * (n - d)


def test_ba_batch_next_to_copy_graphs(ctx, oracle_mod):
    import torch
    import cuppa
    rng = np.random.default_rng(3)
    specs = []
    for i in range(40):
        d = int(rng.integers(1, 17))
        n = int(rng.integers(d + 1, 50_000)) if i % 9 else d + 1
        model = cuppa.MODEL_BA if i % 2 else cuppa.MODEL_COPY
        specs.append((n, d, oracle_mod.p_threshold(float(rng.uniform())), 500 + i, model))
    descs, tot, rot = cuppa.batch_layout(specs)
    targets = torch.full((tot,), -7, dtype=torch.int32, device="cuda")
    row_ptr = torch.full((rot,), -7, dtype=torch.int64, device="cuda")
    b = cuppa.Batch(ctx, descs, targets, row_ptr)
    b.run()
    b.run()  # a replay gives the same graphs
    b.close()
    tg = targets.cpu().numpy()
    for (n, d, T, seed, model), ds in zip(specs, descs):
        if model == cuppa.MODEL_BA:
            want, _ = oracle_mod.generate_ba(n, d, seed)
        else:
            want, _, _ = oracle_mod.generate(n, d, T, seed)
        assert np.array_equal(tg[ds.targets_offset:ds.targets_offset + n * d], want), (n, d, seed, model)


def test_ba_requests_in_the_server(ctx, oracle_mod):
    import cuppa
    rng = np.random.default_rng(4)
    specs = []
    for i in range(30):
        d = int(rng.choice([2, 4, 8, 16]))
        n = max(d + 1, int(np.clip(rng.lognormal(np.log(2e4), 1.0), 20, 120_000)))
        model = cuppa.MODEL_BA if i % 3 else cuppa.MODEL_COPY
        specs.append((i, n, d, oracle_mod.p_threshold(0.3), 900 + i, model))
    srv = cuppa.Server(ctx, budget_bytes=1 << 30, streams=4, batch_nodes=1 << 18, solo_nodes=1 << 17)
    for s in specs:
        assert srv.submit(*s) == 0
    pending = {s[0]: s for s in specs}
    for _ in range(2_000_000):
        srv.step()
        for rid in list(pending):
            r = srv.poll(rid)
            if r.state in (cuppa.REQ_DONE, cuppa.REQ_FAILED):
                assert r.state == cuppa.REQ_DONE, (rid, r.status)
                _, n, d, T, seed, model = pending.pop(rid)
                got = cuppa.read_device(r.targets, n * d, np.int32)
                if model == cuppa.MODEL_BA:
                    want, _ = oracle_mod.generate_ba(n, d, seed)
                else:
                    want, _, _ = oracle_mod.generate(n, d, T, seed)
                assert np.array_equal(got, want), (rid, model)
                assert srv.release(rid) == 0
        if not pending:
            break
    srv.close()
    assert not pending


def test_ba_in_csr_and_no_undirected_form(ctx, oracle_mod):
    import cuppa
    n, d, seed = 40_000, 6, 8
    g = _gpu_ba(ctx, n, d, seed)
    in_ptr, in_src = ctx.in_csr(g)
    G = g.targets.cpu().numpy()
    want_ptr, want_src = oracle_mod.in_csr(n, d, G)
    assert np.array_equal(in_ptr.cpu().numpy(), want_ptr)
    assert np.array_equal(in_src.cpu().numpy(), want_src)
    with pytest.raises(cuppa.PPAError) as e:
        ctx.undirected_csr(g)
    assert e.value.code == -1  # PPA_ERR_INVALID


def test_runner_sba(tmp_path, oracle_mod):
    n, d, seed = 60_000, 5, 31
    r = subprocess.run([RUNNER, "-A", "sba", "-n", str(n), "-x", str(d), "--seed", str(seed), "--show-dd", "1",
                        "--dd-filename", "dd.csv", "-o", str(tmp_path) + "/", "--edges-file",
                        str(tmp_path / "g.edges")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout.splitlines()
    assert "SBA" in out
    assert any(line.startswith("Computation Time : ") for line in out)
    assert f"MAXIMUM SIZE: {2 * n * d}" in out and "Degree of each vertex is computed!" in out
    log = (tmp_path / "SBA.log").read_text().splitlines()
    assert log[:3] == ["SBA", str(n), str(d)] and "Memory Allocated" in log
    G, _ = oracle_mod.generate_ba(n, d, seed)
    deg, _, _ = oracle_mod.degrees_ba(n, d, G)
    want = np.bincount(deg.astype(np.int64))
    with open(tmp_path / "dd.csv") as f:
        assert f.readline().strip() == "Degree,Count"
        got = {int(a): int(b) for a, b in (line.strip().split(",") for line in f)}
    assert got == {i: int(c) for i, c in enumerate(want) if c}
    edges = np.loadtxt(tmp_path / "g.edges", dtype=np.int64)
    assert edges.shape == ((n - d) * d, 2)
    assert np.array_equal(edges[:, 0], np.repeat(np.arange(d, n), d))
    assert np.array_equal(edges[:, 1], G[d * d:])


def test_ba_in_ranges(ctx, oracle_mod):
    """Barabasi-Albert graphs generated range by range (ppa_generate_range),
    the first range ending inside the 160 KB prefix (d = 16)."""
    import torch
    import cuppa
    n, d, seed = 30_000, 16, 12
    want, _ = oracle_mod.generate_ba(n, d, seed)
    desc = cuppa.make_desc(n, d, seed=seed, model=cuppa.MODEL_BA)
    targets = torch.full((n * d,), -7, dtype=torch.int32, device="cuda")
    for a, b in ((0, 700), (700, 9_000), (9_000, n)):
        ctx.generate_range(desc, targets, None, a, b)
        ctx.sync()
        assert np.array_equal(targets[:b * d].cpu().numpy(), want[:b * d]), (a, b)
