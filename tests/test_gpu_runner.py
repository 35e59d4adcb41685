"""The C++ host driver on the GPU (`-A cuppa` / `-A cuppa-adaptive`): same
flags, stdout lines and files as the reference driver, graph equal to the
oracle's."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cuppa-public_amd", "bin", "ppa_runner")


def _read_dd(path):
    with open(path) as f:
        assert f.readline().strip() == "Degree,Count"
        return {int(a): int(b) for a, b in (line.strip().split(",") for line in f)}


@pytest.mark.parametrize("algo", ["cuppa", "cuppa-adaptive"])
def test_runner_gpu_outputs(tmp_path, oracle_mod, algo):
    n, d, p, seed = 30_000, 6, 0.3, 77
    r = subprocess.run([RUNNER, "-A", algo, "-n", str(n), "-x", str(d), "-a", str(p), "--seed", str(seed),
                        "--show-dd", "1", "--dd-filename", "dd.csv", "-o", str(tmp_path) + "/",
                        "--show-time-n-stages", "1", "--edges-file", str(tmp_path / "g.edges"), "-t", "128"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Computation Time: " in r.stdout and "Ended" in r.stdout
    G, _, _ = oracle_mod.generate(n, d, oracle_mod.p_threshold(p), seed)
    deg, _, _ = oracle_mod.degrees(n, d, G)
    want = np.bincount(deg.astype(np.int64))
    assert _read_dd(tmp_path / "dd.csv") == {i: int(c) for i, c in enumerate(want) if c}
    edges = np.loadtxt(tmp_path / "g.edges", dtype=np.int64)
    nclique = d * (d + 1) // 2
    assert np.array_equal(edges[nclique:, 1], G[(d + 1) * d:])
    csv = "StepwiseTiming-Original.csv" if algo == "cuppa" else "StepwiseTiming-Adaptive.csv"
    lines = (tmp_path / csv).read_text().splitlines()
    assert lines[0] == "Round,nBlocks,nThreads,Edges,MaxQ,Initialization Time,Computation Time"
    assert len(lines) == 2 and lines[1].split(",")[3] == str(n * d)


def test_runner_gpu_csr_files(tmp_path, oracle_mod):
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_abi import read_csr
    n, d, p, seed = 40_000, 5, 0.25, 13
    r = subprocess.run([RUNNER, "-A", "cuppa", "-n", str(n), "-x", str(d), "-a", str(p), "--seed", str(seed),
                        "--csr-file", str(tmp_path / "out.csr"), "--in-csr-file", str(tmp_path / "in.csr")],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    magic, n2, nnz, d2, kind, seed2, thr, ptr, idx = read_csr(tmp_path / "out.csr")
    assert (magic, n2, nnz, d2, kind, seed2, thr) == (b"PPACSR1\0", n, n * d, d, 0, seed, T)
    assert np.array_equal(ptr, np.arange(n + 1, dtype=np.int64) * d)
    assert np.array_equal(idx, G)
    want_ptr, want_src = oracle_mod.in_csr(n, d, G)
    magic, n2, nnz, d2, kind, _, _, ptr, idx = read_csr(tmp_path / "in.csr")
    assert (n2, nnz, kind) == (n, n * d, 1)
    assert np.array_equal(ptr, want_ptr)
    assert np.array_equal(idx, want_src)


def test_runner_csr_written_while_generating(tmp_path, oracle_mod):
    """Above 2^20 nodes `-A cuppa --csr-file` generates the graph in ranges of
    2^20 nodes and writes each one's rows while the next is generated; the
    file is the graph's CSR all the same."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_abi import read_csr
    n, d, p, seed = 2_500_000, 3, 0.4, 21
    r = subprocess.run([RUNNER, "-A", "cuppa", "-n", str(n), "-x", str(d), "-a", str(p), "--seed", str(seed),
                        "--csr-file", str(tmp_path / "out.csr")],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Computation Time: " in r.stdout and "Ended" in r.stdout
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    magic, n2, nnz, d2, kind, seed2, thr, ptr, idx = read_csr(tmp_path / "out.csr")
    assert (magic, n2, nnz, d2, kind, seed2, thr) == (b"PPACSR1\0", n, n * d, d, 0, seed, T)
    assert np.array_equal(ptr, np.arange(n + 1, dtype=np.int64) * d)
    assert np.array_equal(idx, G)


def test_runner_adaptive_ranges(tmp_path, oracle_mod):
    """`-A cuppa-adaptive` generates one graph as consecutive node ranges
    (cuPPA_2.cu:181-211 semantics): the first of 2^16 nodes, each later one
    twice the previous while the time per node improves, up to n/8; one CSV
    row per range, whose edges add up to the graph's; the graph is the
    oracle's."""
    n, d, p, seed = 2_000_000, 8, 0.2, 21
    r = subprocess.run([RUNNER, "-A", "cuppa-adaptive", "-n", str(n), "-x", str(d), "-a", str(p), "--seed",
                        str(seed), "--show-time-n-stages", "1000", "--edges-file", str(tmp_path / "g.edges")],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Computation Time: " in r.stdout
    lines = (tmp_path / "StepwiseTiming-Adaptive.csv").read_text().splitlines()
    rows = [l.split(",") for l in lines[1:]]
    edges_per_row = [int(row[3]) for row in rows]
    assert len(rows) >= 3
    assert sum(edges_per_row) == n * d
    assert edges_per_row[0] == (1 << 16) * d
    assert edges_per_row[1] == (1 << 17) * d  # the first range always doubles the next
    assert max(edges_per_row) <= (n // 8 + n // 16) * d  # ranges stop doubling at n/8 (the tail joins the last)
    assert all(float(row[6]) > 0 for row in rows)
    G, _, _ = oracle_mod.generate(n, d, oracle_mod.p_threshold(p), seed)
    edges = np.loadtxt(tmp_path / "g.edges", dtype=np.int64)
    assert np.array_equal(edges[d * (d + 1) // 2:, 1], G[(d + 1) * d:])
