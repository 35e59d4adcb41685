"""C ABI and host driver checks that need no GPU.

* the library loads and exports every function include/ppa.h declares;
* argument validation and status codes;
* the product's serial CPU path (ppa_cpu_generate, `-A scm`) against the oracle;
* the command-line driver: flags, INI file, output formats, error exit.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ppa.h")
RUNNER = os.path.join(ROOT, "cuppa-public_amd", "bin", "ppa_runner")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ppa_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    import cuppa
    L = cuppa.lib()
    L.ppa_cpu_generate.argtypes = [ctypes.POINTER(cuppa.GraphDesc), ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_uint64)]
    L.ppa_cpu_generate.restype = ctypes.c_int
    L.ppa_cpu_degree_histogram.argtypes = [ctypes.POINTER(cuppa.GraphDesc), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.ppa_cpu_degree_histogram.restype = ctypes.c_int
    return L


def test_exports_every_declared_symbol(L):
    names = declared_functions()
    assert len(names) >= 14
    out = subprocess.run(["nm", "-D", "--defined-only", L._name], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (ppa_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(L, n)


def test_binding_covers_the_header():
    import cuppa
    bound = set(cuppa.exported_symbols())
    extra = {"ppa_cpu_generate", "ppa_cpu_degree_histogram"}
    assert set(declared_functions()) <= bound | extra


def test_strerror_and_threshold(L, oracle_mod):
    for code in range(0, -12, -1):
        assert L.ppa_strerror(code) != b"unknown error"
    assert L.ppa_strerror(-99) == b"unknown error"
    for p in (0.0, 1.0, 0.5, 0.2, 0.33333, 1e-9, 2.0, -0.1):
        assert L.ppa_p_threshold(p) == oracle_mod.p_threshold(p)


@pytest.mark.parametrize("n,d,T,ok", [
    (10, 4, 0, True), (5, 4, 0, True), (4, 4, 0, False), (10, 0, 0, False), (100, 16, 0, True),
    (100, 17, 0, True), (100, 32, 0, True), (100, 33, 0, True), (33, 32, 0, True), (32, 32, 0, False),
    (8193, 8192, 0, True), (8192, 8192, 0, False), (9000, 8193, 0, False),
    (2**31 - 1, 8, 0, True), (2**31, 8, 0, False), (10, 2, 2**32, True),
    (10, 2, 2**32 + 1, False)])
def test_check_desc(L, n, d, T, ok):
    import cuppa
    desc = cuppa.GraphDesc(n, d, 0, T, 0, 0, 0)
    assert (L.ppa_check_desc(ctypes.byref(desc)) == 0) == ok


@pytest.mark.parametrize("model,ok", [(0, True), (1, True), (2, False), (2**31, False)])
def test_check_desc_model(L, model, ok):
    import cuppa
    desc = cuppa.GraphDesc(100, 4, model, 0, 0, 0, 0)
    assert (L.ppa_check_desc(ctypes.byref(desc)) == 0) == ok


def test_null_arguments_rejected(L):
    assert L.ppa_check_desc(None) == -1
    assert L.ppa_ctx_destroy(None) == -1
    assert L.ppa_get_stats(None, None) == -1
    import cuppa
    desc = cuppa.GraphDesc(100, 4, 0, 0, 0, 0, 0)
    assert L.ppa_generate_range(None, ctypes.byref(desc), None, None, 0, 50, None) == -1
    assert L.ppa_generate_async(None, ctypes.byref(desc), None, None, None) == -1


def test_ctx_create_reports_missing_gpu(L):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert L.ppa_ctx_create(ctypes.byref(h), 0, 0) == -4  # PPA_ERR_DEVICE, no crash


@pytest.mark.parametrize("n,d,p,seed", [(6, 5, 0.5, 1), (20_000, 1, 0.5, 2), (50_000, 8, 0.2, 11),
                                        (30_000, 16, 0.9, 3), (40_000, 3, 0.0, 4), (1_000_000, 4, 0.5, 42),
                                        (20_000, 32, 0.3, 5), (5_000, 17, 0.0, 6)])
def test_cpu_path_matches_oracle(L, oracle_mod, n, d, p, seed):
    import cuppa
    desc = cuppa.make_desc(n, d, p, seed)
    G = np.empty(n * d, dtype=np.int32)
    rp = np.empty(n + 1, dtype=np.int64)
    retries = ctypes.c_uint64()
    rc = L.ppa_cpu_generate(ctypes.byref(desc), G.ctypes.data, rp.ctypes.data, ctypes.byref(retries))
    assert rc == 0
    want, _, want_retries = oracle_mod.generate(n, d, oracle_mod.p_threshold(p), seed)
    assert np.array_equal(G, want)
    assert retries.value == want_retries
    assert np.array_equal(rp, np.arange(n + 1) * d)
    deg = np.empty(n, dtype=np.uint32)
    hist = np.empty(4096, dtype=np.uint64)
    mx = ctypes.c_uint64()
    assert L.ppa_cpu_degree_histogram(ctypes.byref(desc), G.ctypes.data, deg.ctypes.data, hist.ctypes.data,
                                      hist.size, ctypes.byref(mx)) == 0
    wdeg, whist, wmx = oracle_mod.degrees(n, d, want, hist_len=4096)
    assert np.array_equal(deg.astype(np.uint64), wdeg)
    assert np.array_equal(hist, whist)
    assert mx.value == wmx


@pytest.mark.parametrize("n,d,seed", [(2, 1, 0), (5, 4, 1), (20_000, 1, 2), (50_000, 8, 11),
                                      (30_000, 16, 3), (1_000_000, 4, 42)])
def test_cpu_path_ba_matches_oracle(L, oracle_mod, n, d, seed):
    """PPA_MODEL_BA through the product's serial loop (copy-model form over
    self-loop seed rows) against the oracle's endpoint-array restatement of
    SerialBA.cpp; p_threshold is ignored."""
    import cuppa
    desc = cuppa.make_desc(n, d, p_thresh=12345, seed=seed, model=cuppa.MODEL_BA)
    G = np.empty(n * d, dtype=np.int32)
    retries = ctypes.c_uint64()
    assert L.ppa_cpu_generate(ctypes.byref(desc), G.ctypes.data, None, ctypes.byref(retries)) == 0
    want, want_retries = oracle_mod.generate_ba(n, d, seed)
    assert np.array_equal(G, want)
    assert retries.value == want_retries
    deg = np.empty(n, dtype=np.uint32)
    hist = np.empty(8192, dtype=np.uint64)
    mx = ctypes.c_uint64()
    assert L.ppa_cpu_degree_histogram(ctypes.byref(desc), G.ctypes.data, deg.ctypes.data, hist.ctypes.data,
                                      hist.size, ctypes.byref(mx)) == 0
    wdeg, whist, wmx = oracle_mod.degrees_ba(n, d, want, hist_len=8192)
    assert np.array_equal(deg.astype(np.uint64), wdeg)
    assert np.array_equal(hist, whist)
    assert mx.value == wmx


def run(args, cwd):
    return subprocess.run([RUNNER] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


@pytest.fixture(scope="module")
def runner():
    if not os.path.exists(RUNNER):
        pytest.skip("runner not built")
    return RUNNER


def _read_dd(path):
    with open(path) as f:
        assert f.readline().strip() == "Degree,Count"
        return {int(a): int(b) for a, b in (line.strip().split(",") for line in f)}


def test_runner_scm_outputs(runner, tmp_path, oracle_mod):
    n, d, p, seed = 50_000, 4, 0.5, 42
    r = run(["-A", "scm", "-n", str(n), "-x", str(d), "-a", str(p), "--seed", str(seed), "--show-dd", "1",
             "--dd-filename", "dd.csv", "-o", str(tmp_path) + "/", "--edges-file", str(tmp_path / "g.edges")],
            tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Computation Time : " in r.stdout
    G, _, _ = oracle_mod.generate(n, d, oracle_mod.p_threshold(p), seed)
    deg, _, _ = oracle_mod.degrees(n, d, G)
    want = np.bincount(deg.astype(np.int64))
    assert _read_dd(tmp_path / "dd.csv") == {i: int(c) for i, c in enumerate(want) if c}
    edges = np.loadtxt(tmp_path / "g.edges", dtype=np.int64)
    clique = [(i, j) for i in range(1, d + 1) for j in range(i)]
    assert edges.shape == (len(clique) + (n - d - 1) * d, 2)
    assert [tuple(e) for e in edges[:len(clique)]] == clique
    body = edges[len(clique):]
    assert np.array_equal(body[:, 0], np.repeat(np.arange(d + 1, n), d))
    assert np.array_equal(body[:, 1], G[(d + 1) * d:])
    assert (tmp_path / "SCM.log").exists()


def test_runner_ini_overrides_flags(runner, tmp_path):
    ini = tmp_path / "c.ini"
    ini.write_text("[config]\nalgorithm = scm\nn1 = 3000\nx1 = 3\nalpha = 0.25\nseed = 9\n"
                   f"outputdir = {tmp_path}/\ndistfile1 = x.csv\n")
    r = run(["-n", "10", "-c", str(ini), "--show-dd", "1"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "N:\t3000\tM:\t3" in r.stdout
    dd = _read_dd(tmp_path / "x.csv")
    assert sum(dd.values()) == 3000


def test_runner_rejects_bad_input(runner, tmp_path):
    r = run(["-A", "nope"], tmp_path)
    assert r.returncode == 1 and "Invalid Algorithm" in r.stderr
    r = run(["--no-such-flag", "1"], tmp_path)
    assert r.returncode == 1 and "not expected" in r.stdout
    r = run(["-A", "scm", "-n", "3", "-x", "4"], tmp_path)
    assert r.returncode == 1 and "invalid argument" in r.stderr
    r = run(["-n", "abc"], tmp_path)
    assert r.returncode == 1


def read_csr(path):
    """Reads the driver's binary CSR file: 64-byte header, int64 offsets[n+1],
    int32 indices[nnz]."""
    raw = open(path, "rb").read()
    magic = raw[:8]
    n, nnz, d, kind, seed, thr = np.frombuffer(raw[8:48], dtype=np.dtype(
        [("n", "<u8"), ("nnz", "<u8"), ("d", "<u4"), ("kind", "<u4"), ("seed", "<u8"), ("thr", "<u8")]))[0]
    ptr = np.frombuffer(raw, dtype="<i8", count=int(n) + 1, offset=64)
    idx = np.frombuffer(raw, dtype="<i4", count=int(nnz), offset=64 + 8 * (int(n) + 1))
    assert len(raw) == 64 + 8 * (int(n) + 1) + 4 * int(nnz)
    return magic, int(n), int(nnz), int(d), int(kind), int(seed), int(thr), ptr, idx


def test_runner_scm_csr_file(runner, tmp_path, oracle_mod):
    n, d, p, seed = 20_000, 3, 0.4, 8
    r = run(["-A", "scm", "-n", str(n), "-x", str(d), "-a", str(p), "--seed", str(seed),
             "--csr-file", str(tmp_path / "g.csr")], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    magic, n2, nnz, d2, kind, seed2, thr, ptr, idx = read_csr(tmp_path / "g.csr")
    T = oracle_mod.p_threshold(p)
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    assert magic == b"PPACSR1\0" and (n2, nnz, d2, kind, seed2, thr) == (n, n * d, d, 0, seed, T)
    assert np.array_equal(ptr, np.arange(n + 1, dtype=np.int64) * d)
    assert np.array_equal(idx, G)
