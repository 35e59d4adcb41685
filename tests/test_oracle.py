"""Pins the CPU oracle (no GPU needed).

Each check is chosen so that a plausible slip in the oracle -- a swapped
Philox lane, a dropped key bump, a wrong shift in the 64-bit draw, an
off-by-one in the node range, a missing duplicate check -- fails at least one
of them: published known-answer vectors, an independent Python big-integer
restatement, closed forms of special cases and structural invariants.
"""
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

# Random123 known-answer vectors for philox4x32 with 10 rounds:
# (counter, key) -> output
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(oracle_mod, ctr, key, want):
    assert oracle_mod.philox(ctr, key) == want
    assert oracle_mod.py_philox(ctr, key) == want


def test_philox_c_matches_python_restatement(oracle_mod):
    rng = np.random.default_rng(1)
    for _ in range(500):
        ctr = tuple(int(x) for x in rng.integers(0, 2**32, 4, dtype=np.uint64))
        key = tuple(int(x) for x in rng.integers(0, 2**32, 2, dtype=np.uint64))
        assert oracle_mod.philox(ctr, key) == oracle_mod.py_philox(ctr, key)


ROCRAND_PROBE = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
int main() {
    unsigned long long seeds[3] = {0ull, 0x0123456789abcdefull, 0xffffffffffffffffull};
    unsigned long long subs[3] = {0ull, 5ull, 0x00000007000000ffull};
    for (auto s : seeds) for (auto q : subs) {
        rocrand_device::philox4x32_10_engine e(s, q, 0);
        uint4 v = e.next4();
        std::printf("%llu %llu %u %u %u %u\n", s, q, v.x, v.y, v.z, v.w);
    }
}
"""


@pytest.mark.skipif(not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"),
                    reason="rocrand headers absent")
def test_philox_matches_rocrand_header(oracle_mod):
    """rocrand's host-callable engine: seed -> key, subsequence -> counter
    words 2..3, offset 0 -> counter words 0..1 = 0."""
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "p.cpp")
        exe = os.path.join(tmp, "p")
        with open(src, "w") as f:
            f.write(ROCRAND_PROBE)
        hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
        r = subprocess.run([hipcc, "-O1", "-o", exe, src], capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("cannot build rocrand probe: " + r.stderr[-300:])
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [list(map(int, line.split())) for line in out if line.strip()]
    assert len(rows) == 9
    for s, q, *v in rows:
        ctr = (0, 0, q & 0xFFFFFFFF, q >> 32)
        key = (s & 0xFFFFFFFF, s >> 32)
        assert oracle_mod.philox(ctr, key) == tuple(v)


def test_mulhi64_matches_big_integers(oracle_mod):
    rng = np.random.default_rng(2)
    L = oracle_mod.lib()
    cases = [(0, 0), (2**64 - 1, 2**64 - 1), (2**64 - 1, 1), (2**63, 2), (1 << 32, 1 << 32)]
    cases += [(int(a), int(b)) for a, b in rng.integers(0, 2**63, (200, 2), dtype=np.uint64)]
    for a, b in cases:
        assert L.oracle_mulhi64(a, b) == (a * b) >> 64


@pytest.mark.parametrize("p", [0.0, -1.0, 1.0, 1.5, 0.5, 0.2, 0.1, 1e-12, 1 - 2**-40, 0.9, 0.05, 0.95])
def test_p_threshold(oracle_mod, p):
    want = oracle_mod.py_p_threshold(p)
    assert oracle_mod.p_threshold(p) == want
    assert 0 <= want <= 2**32


def test_coin_threshold_edges(oracle_mod):
    assert oracle_mod.p_threshold(0.5) == 2**31
    assert oracle_mod.p_threshold(1.0) == 2**32   # every draw direct
    assert oracle_mod.p_threshold(0.0) == 0       # never direct


def test_attempt_mapping_matches_python(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(300):
        d = int(rng.integers(1, 17))
        t = int(rng.integers(d + 1, 2**31))
        k = int(rng.integers(0, d))
        r = int(rng.integers(0, 5))
        seed = int(rng.integers(0, 2**63))
        T = int(rng.integers(0, 2**32 + 1))
        x = oracle_mod.py_philox((t & 0xFFFFFFFF, t >> 32, k, r), (seed & 0xFFFFFFFF, seed >> 32))
        R = (((x[1] << 32) | x[2]) * (t * d)) >> 64
        assert R < t * d
        direct, u, j = oracle_mod.attempt(seed, T, t, d, k, r)
        assert (direct, u, j) == (x[0] < T, R // d, R % d)


SMALL = [(2, 1, 0.5, 0), (3, 1, 0.5, 1), (7, 6, 0.3, 2), (40, 3, 0.0, 3), (40, 3, 1.0, 4),
         (300, 4, 0.5, 5), (500, 8, 0.2, 6), (200, 16, 0.7, 7), (600, 2, 0.9, 2**40 + 3)]


@pytest.mark.parametrize("n,d,p,seed", SMALL)
def test_c_oracle_matches_python_restatement(oracle_mod, n, d, p, seed):
    T = oracle_mod.p_threshold(p)
    G, rp, _ = oracle_mod.generate(n, d, T, seed, with_row_ptr=True)
    assert np.array_equal(G, oracle_mod.py_generate(n, d, T, seed))
    assert np.array_equal(rp, np.arange(n + 1) * d)


def _check_invariants(G, n, d):
    rows = G.reshape(n, d)
    for i in range(d + 1):
        assert list(rows[i]) == [j + (j >= i) for j in range(d)]
    t = np.arange(n)[:, None]
    body = rows[d + 1:]
    assert (body >= 0).all() and (body < t[d + 1:]).all()
    s = np.sort(body, axis=1)
    assert (s[:, 1:] != s[:, :-1]).all()


@pytest.mark.parametrize("n,d,p,seed", [(100_000, 8, 0.2, 11), (100_000, 1, 0.5, 1), (50_000, 16, 0.05, 9)])
def test_invariants_and_degree_sums(oracle_mod, n, d, p, seed):
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    _check_invariants(G, n, d)
    deg, hist, mx = oracle_mod.degrees(n, d, G, hist_len=1 << 16)
    assert int(hist.sum()) == n
    assert int((hist * np.arange(hist.size, dtype=np.uint64)).sum()) == d * (d + 1) + 2 * (n - d - 1) * d
    assert mx == int(deg.max())
    indeg = np.bincount(G[(d + 1) * d:], minlength=n)
    assert np.array_equal(deg, (indeg + d).astype(np.uint64))


def test_copy_only_stays_in_seed_clique(oracle_mod):
    # p = 0: targets only propagate the seed clique's targets, so all lie in 0..d
    for d in (1, 3, 8):
        G, _, _ = oracle_mod.generate(20_000, d, 0, 5)
        assert G.max() <= d and G.min() >= 0


def test_d1_never_retries(oracle_mod):
    for p in (0.0, 0.3, 1.0):
        _, _, retries = oracle_mod.generate(50_000, 1, oracle_mod.p_threshold(p), 8)
        assert retries == 0


def test_uniform_recursive_tree_in_degrees(oracle_mod):
    """p = 1, d = 1: node t links to a uniform node of [0, t), so
    E[in-degree(v)] = H(n-1) - H(v) for v >= 2 and H(n-1) - 1 for v in {0, 1}
    (the seed edge 0-1 is not counted as in-degree)."""
    n, seeds = 20_000, 40
    acc = np.zeros(n)
    for s in range(seeds):
        G, _, _ = oracle_mod.generate(n, 1, 2**32, 1000 + s)
        acc += np.bincount(G[2:], minlength=n)
    acc /= seeds
    H = np.concatenate([[0.0], np.cumsum(1.0 / np.arange(1, n))])  # H[m] = sum_{i<=m} 1/i
    exp = H[n - 1] - H
    exp[0] = exp[1] = H[n - 1] - 1.0
    # bands keep the statistical error small; a node-range off-by-one moves them
    t = np.arange(2, n, dtype=np.float64)
    for lo, hi in [(0, 2), (2, 20), (20, 200), (200, 2000), (2000, 20000)]:
        got, want = acc[lo:hi].sum(), exp[lo:hi].sum()
        # node t lands in [lo, hi) with probability q_t = |[lo, min(hi, t))| / t,
        # independently over t; the band total is a sum of those Bernoullis
        q = np.clip(np.minimum(hi, t) - lo, 0, None) / t
        sd = math.sqrt(float((q * (1 - q)).sum()) / seeds)
        assert abs(got - want) < 6 * sd + 1e-6, (lo, hi, got, want, sd)


def test_closure_matches_sequential(oracle_mod):
    n, d, p, seed = 200_000, 8, 0.5, 13
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    c = oracle_mod.Closure(n, d, T, seed)
    for start in (9, 1000, 54_321, 150_000, n - 64):
        assert np.array_equal(c.rows(start, 64), G[start * d:(start + 64) * d])
    c2 = oracle_mod.Closure(n, d, T, seed, prefix=G[:50_000 * d])
    assert np.array_equal(c2.rows(n - 500, 500), G[(n - 500) * d:])


def test_verifier_accepts_oracle_and_catches_planted_mistakes(oracle_mod):
    n, d, p, seed = 100_000, 6, 0.3, 21
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    assert oracle_mod.verify_rows(n, d, T, seed, G, threads=4) is None
    rng = np.random.default_rng(5)
    for slot in rng.integers((d + 1) * d, n * d, 20):
        bad = G.copy()
        bad[slot] = (bad[slot] + 1) % max(1, slot // d)
        if bad[slot] == G[slot]:
            continue
        first_bad = oracle_mod.verify_rows(n, d, T, seed, bad, threads=4)
        assert first_bad is not None and first_bad <= slot // d
    seedbad = G.copy()
    seedbad[0] = 0
    assert oracle_mod.verify_rows(n, d, T, seed, seedbad, 0, d + 1) == 0


def test_in_csr_oracle_pinned(oracle_mod):
    """The in-edge CSR oracle (a counting sort) against an independent method
    (numpy's stable argsort of the targets) and against the degree pass:
    in(v) = deg(v) - d, plus d for the seed nodes 0..d."""
    n, d = 5000, 4
    T = oracle_mod.p_threshold(0.3)
    G, _, _ = oracle_mod.generate(n, d, T, 3)
    ptr, src = oracle_mod.in_csr(n, d, G)
    order = np.argsort(G, kind="stable")
    assert np.array_equal(src, (order // d).astype(np.int32))
    assert np.array_equal(np.diff(ptr), np.bincount(G, minlength=n))
    deg, _, _ = oracle_mod.degrees(n, d, G)
    want_in = deg.astype(np.int64) - d
    want_in[: d + 1] += d
    assert np.array_equal(np.diff(ptr), want_in)
    for v in (0, 1, d + 1, n // 2):  # each list ascending and naming v
        lst = src[ptr[v]:ptr[v + 1]]
        assert np.all(np.diff(lst) > 0)
        assert all(v in G[t * d:(t + 1) * d] for t in lst)
    # out-of-range values are skipped
    H = G.copy()
    H[[10, 20]] = -1
    H[30] = n
    ptr2, src2 = oracle_mod.in_csr(n, d, H)
    assert ptr2[n] == n * d - 3
    keep = (H >= 0) & (H < n)
    assert np.array_equal(src2, (np.argsort(np.where(keep, H, n), kind="stable")[: ptr2[n]] // d).astype(np.int32))


def test_undirected_oracle_pinned(oracle_mod):
    """Undirected lists: symmetric as multisets (v in list(u) as often as u in
    list(v)), |list(v)| = deg(v) of the degree pass, each list starting with
    the node's own row."""
    n, d = 4000, 5
    G, _, _ = oracle_mod.generate(n, d, oracle_mod.p_threshold(0.3), 9)
    ptr, adj = oracle_mod.undirected_csr(n, d, G)
    deg, _, _ = oracle_mod.degrees(n, d, G)
    assert np.array_equal(np.diff(ptr), deg.astype(np.int64))
    assert ptr[n] == 2 * n * d - (d + 1) * d
    src = np.repeat(np.arange(n), np.diff(ptr))
    pairs = np.stack([src, adj.astype(np.int64)], 1)
    fwd = np.unique(pairs, axis=0, return_counts=True)
    rev = np.unique(pairs[:, ::-1], axis=0, return_counts=True)
    assert np.array_equal(fwd[0], rev[0]) and np.array_equal(fwd[1], rev[1])
    for v in (0, d, d + 1, n - 1):
        assert np.array_equal(adj[ptr[v]:ptr[v] + d], G[v * d:(v + 1) * d])
