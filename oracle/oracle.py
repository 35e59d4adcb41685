"""CPU oracle bindings -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, and only as the checker. The product path never does.

It wraps ``oracle/ppa_oracle.c`` (a sequential restatement of
src/PA/SPA/SerialCopyModel3.cpp:38-78 fed keyed Philox-4x32-10 draws) with
ctypes, and also holds an independent pure-Python restatement of the same
model (big-integer Philox, list-based graph) used to pin the C code on small
inputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libppa_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "scm_ref")
REF_BA_BIN = os.path.join(HERE, "_ref", "sba_ref")

MAX_RETRY = 1 << 24

_lib = None


def build(quiet: bool = True) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    u64, u32, i32p = ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)
    L.oracle_philox4x32_10.argtypes = [ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]
    L.oracle_philox4x32_10.restype = None
    L.oracle_p_threshold.argtypes = [ctypes.c_double]
    L.oracle_p_threshold.restype = u64
    L.oracle_mulhi64.argtypes = [u64, u64]
    L.oracle_mulhi64.restype = u64
    L.oracle_attempt.argtypes = [u64, u64, u64, u32, u32, u32, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(u64), ctypes.POINTER(u32)]
    L.oracle_attempt.restype = None
    L.oracle_generate.argtypes = [u64, u32, u64, u64, i32p, ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(u64)]
    L.oracle_generate.restype = ctypes.c_int
    L.oracle_generate_ba.argtypes = [u64, u32, u64, i32p, ctypes.POINTER(u64)]
    L.oracle_generate_ba.restype = ctypes.c_int
    L.oracle_degrees_ba.argtypes = [u64, u32, i32p, ctypes.POINTER(u64), ctypes.POINTER(u64), u64,
                                    ctypes.POINTER(u64)]
    L.oracle_degrees_ba.restype = ctypes.c_int
    L.oracle_closure_create.argtypes = [u64, u32, u64, u64]
    L.oracle_closure_create.restype = ctypes.c_void_p
    L.oracle_closure_set_prefix.argtypes = [ctypes.c_void_p, i32p, u64]
    L.oracle_closure_set_prefix.restype = ctypes.c_int
    L.oracle_closure_destroy.argtypes = [ctypes.c_void_p]
    L.oracle_closure_destroy.restype = None
    L.oracle_closure_target.argtypes = [ctypes.c_void_p, u64, u32, i32p]
    L.oracle_closure_target.restype = ctypes.c_int
    L.oracle_closure_rows.argtypes = [ctypes.c_void_p, u64, u64, i32p]
    L.oracle_closure_rows.restype = ctypes.c_int
    L.oracle_closure_size.argtypes = [ctypes.c_void_p]
    L.oracle_closure_size.restype = u64
    L.oracle_degrees.argtypes = [u64, u32, i32p, ctypes.POINTER(u64), ctypes.POINTER(u64), u64,
                                 ctypes.POINTER(u64)]
    L.oracle_degrees.restype = ctypes.c_int
    L.oracle_in_csr.argtypes = [u64, u32, i32p, ctypes.POINTER(ctypes.c_int64), i32p]
    L.oracle_in_csr.restype = ctypes.c_int
    L.oracle_undirected_csr.argtypes = [u64, u32, i32p, ctypes.POINTER(ctypes.c_int64), i32p]
    L.oracle_undirected_csr.restype = ctypes.c_int
    L.oracle_verify_rows.argtypes = [u64, u32, u64, u64, i32p, u64, u64, ctypes.POINTER(u64)]
    L.oracle_verify_rows.restype = ctypes.c_int
    L.oracle_verify_parallel.argtypes = [u64, u32, u64, u64, i32p, u64, u64, ctypes.c_int,
                                         ctypes.POINTER(u64)]
    L.oracle_verify_parallel.restype = ctypes.c_int
    L.oracle_verify_parallel_model.argtypes = [u32, u64, u32, u64, u64, i32p, u64, u64, ctypes.c_int,
                                               ctypes.POINTER(u64)]
    L.oracle_verify_parallel_model.restype = ctypes.c_int
    _lib = L
    return L


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return tuple(o)


def p_threshold(p: float) -> int:
    return int(lib().oracle_p_threshold(p))


def attempt(seed, T, t, d, k, r):
    direct = ctypes.c_int()
    u = ctypes.c_uint64()
    j = ctypes.c_uint32()
    lib().oracle_attempt(seed, T, t, d, k, r, ctypes.byref(direct), ctypes.byref(u), ctypes.byref(j))
    return bool(direct.value), int(u.value), int(j.value)


def generate(n: int, d: int, T: int, seed: int, with_row_ptr: bool = False):
    """Sequential oracle. Returns (targets int32 [n*d], row_ptr or None, retries)."""
    G = np.empty(n * d, dtype=np.int32)
    rp = np.empty(n + 1, dtype=np.int64) if with_row_ptr else None
    retries = ctypes.c_uint64()
    rc = lib().oracle_generate(n, d, T, seed, _ptr(G, ctypes.c_int32),
                               _ptr(rp, ctypes.c_int64) if rp is not None else None,
                               ctypes.byref(retries))
    if rc != 0:
        raise RuntimeError(f"oracle_generate failed: {rc}")
    return G, rp, int(retries.value)


def generate_ba(n: int, d: int, seed: int):
    """Serial Barabasi-Albert (SerialBA.cpp:68-187) in the reference's
    endpoint-array form. Returns (targets int32 [n*d], retries)."""
    G = np.empty(n * d, dtype=np.int32)
    retries = ctypes.c_uint64()
    rc = lib().oracle_generate_ba(n, d, seed, _ptr(G, ctypes.c_int32), ctypes.byref(retries))
    if rc != 0:
        raise RuntimeError(f"oracle_generate_ba failed: {rc}")
    return G, int(retries.value)


def degrees_ba(n: int, d: int, G: np.ndarray, hist_len: int = 0):
    """Barabasi-Albert degrees (PAHelper.hpp:69-147): (deg, hist or None, max)."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    deg = np.empty(n, dtype=np.uint64)
    hist = np.empty(max(hist_len, 1), dtype=np.uint64)
    mx = ctypes.c_uint64()
    rc = lib().oracle_degrees_ba(n, d, _ptr(G, ctypes.c_int32), _ptr(deg, ctypes.c_uint64),
                                 _ptr(hist, ctypes.c_uint64) if hist_len else None, hist_len,
                                 ctypes.byref(mx))
    if rc != 0:
        raise RuntimeError(f"oracle_degrees_ba failed: {rc}")
    return deg, (hist[:hist_len] if hist_len else None), int(mx.value)


class Closure:
    """Memoised target(t, k) for graphs too large to generate sequentially."""

    def __init__(self, n, d, T, seed, prefix: np.ndarray | None = None):
        self.n, self.d = n, d
        self.h = lib().oracle_closure_create(n, d, T, seed)
        if not self.h:
            raise ValueError("bad closure parameters")
        self._prefix = None
        if prefix is not None:
            self.set_prefix(prefix)

    def set_prefix(self, prefix: np.ndarray):
        """prefix: int32 rows 0..m-1 of this graph from the sequential oracle."""
        self._prefix = np.ascontiguousarray(prefix, dtype=np.int32)
        m = self._prefix.size // self.d
        rc = lib().oracle_closure_set_prefix(self.h, _ptr(self._prefix, ctypes.c_int32), m)
        if rc != 0:
            raise RuntimeError(f"oracle_closure_set_prefix failed: {rc}")

    def rows(self, first: int, count: int) -> np.ndarray:
        out = np.empty(count * self.d, dtype=np.int32)
        rc = lib().oracle_closure_rows(self.h, first, count, _ptr(out, ctypes.c_int32))
        if rc != 0:
            raise RuntimeError(f"oracle_closure_rows failed: {rc}")
        return out

    def target(self, t, k) -> int:
        v = ctypes.c_int32()
        rc = lib().oracle_closure_target(self.h, t, k, ctypes.byref(v))
        if rc != 0:
            raise RuntimeError(f"oracle_closure_target failed: {rc}")
        return int(v.value)

    def size(self) -> int:
        return int(lib().oracle_closure_size(self.h))

    def close(self):
        if self.h:
            lib().oracle_closure_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


NO_BAD_ROW = (1 << 64) - 1


def verify_rows(n: int, d: int, T: int, seed: int, G: np.ndarray, first: int = 0, count: int | None = None,
                threads: int = 1, model: int = 0):
    """Row-consistency check of a complete candidate graph G (int32 [n*d]).
    Returns the first row in [first, first+count) that is not what the model
    produces from G's earlier rows, or None if all pass. model 1 checks a
    Barabasi-Albert graph (T ignored)."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    assert G.size == n * d
    if count is None:
        count = n - first
    bad = ctypes.c_uint64()
    rc = lib().oracle_verify_parallel_model(model, n, d, T, seed, _ptr(G, ctypes.c_int32), first, count, threads,
                                            ctypes.byref(bad))
    if rc != 0:
        raise RuntimeError(f"oracle_verify_parallel failed: {rc}")
    return None if bad.value == NO_BAD_ROW else int(bad.value)


def degrees(n: int, d: int, G: np.ndarray, hist_len: int = 0):
    """Returns (deg uint64 [n], hist uint64 [hist_len] or None, max_degree)."""
    deg = np.empty(n, dtype=np.uint64)
    hist = np.empty(max(hist_len, 1), dtype=np.uint64)
    mx = ctypes.c_uint64()
    rc = lib().oracle_degrees(n, d, _ptr(G, ctypes.c_int32), _ptr(deg, ctypes.c_uint64),
                              _ptr(hist, ctypes.c_uint64) if hist_len else None, hist_len,
                              ctypes.byref(mx))
    if rc != 0:
        raise RuntimeError(f"oracle_degrees failed: {rc}")
    return deg, (hist[:hist_len] if hist_len else None), int(mx.value)


def in_csr(n: int, d: int, G: np.ndarray):
    """Returns (in_ptr int64 [n+1], in_src int32 [in_ptr[n]]) of all n*d slots."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    in_ptr = np.empty(n + 1, dtype=np.int64)
    in_src = np.empty(n * d, dtype=np.int32)
    rc = lib().oracle_in_csr(n, d, _ptr(G, ctypes.c_int32), _ptr(in_ptr, ctypes.c_int64),
                             _ptr(in_src, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(f"oracle_in_csr failed: {rc}")
    return in_ptr, in_src[:in_ptr[n]]


def undirected_csr(n: int, d: int, G: np.ndarray):
    """Returns (ptr int64 [n+1], adj int32 [ptr[n]]) of a complete graph."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    ptr = np.empty(n + 1, dtype=np.int64)
    adj = np.empty(max(1, 2 * n * d - (d + 1) * d), dtype=np.int32)
    rc = lib().oracle_undirected_csr(n, d, _ptr(G, ctypes.c_int32), _ptr(ptr, ctypes.c_int64),
                                     _ptr(adj, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(f"oracle_undirected_csr failed: {rc}")
    return ptr, adj[:ptr[n]]


# ---------------------------------------------------------------------------
# Independent pure-Python restatement (small inputs only).
# Written from the Philox paper's round function and the model definition,
# with Python integers throughout, so that a slip in the C code (a dropped
# term, a swapped lane, a wrong shift) shows up as a mismatch.
# ---------------------------------------------------------------------------
_M32 = 0xFFFFFFFF


def py_philox(ctr, key):
    x = list(ctr)
    k0, k1 = key
    for _ in range(10):
        prod_a = 0xD2511F53 * x[0]
        prod_b = 0xCD9E8D57 * x[2]
        x = [((prod_b >> 32) ^ x[1] ^ k0) & _M32, prod_b & _M32,
             ((prod_a >> 32) ^ x[3] ^ k1) & _M32, prod_a & _M32]
        k0 = (k0 + 0x9E3779B9) & _M32
        k1 = (k1 + 0xBB67AE85) & _M32
    return tuple(x)


def py_p_threshold(p: float) -> int:
    if not p > 0.0:
        return 0
    if p >= 1.0:
        return 1 << 32
    # floor(p * 2^32) from the exact rational value of the double p
    from fractions import Fraction
    return int(Fraction(p) * (1 << 32))


def py_generate(n: int, d: int, T: int, seed: int, self_loop_seed: bool = False):
    """Copy model over the clique seed, or, with self_loop_seed, over seed rows
    i < d that hold i (row d still 0..d-1): the copy-model form of the
    Barabasi-Albert generator at T = 2^31."""
    G = [[-1] * d for _ in range(n)]
    for i in range(d + 1):
        for j in range(d):
            G[i][j] = i if self_loop_seed and i < d else j + (1 if j >= i else 0)
    key = (seed & _M32, seed >> 32)
    for t in range(d + 1, n):
        row = G[t]
        for k in range(d):
            r = 0
            while True:
                x = py_philox((t & _M32, t >> 32, k, r), key)
                R = ((((x[1] << 32) | x[2]) * (t * d)) >> 64)
                u, j = divmod(R, d)
                cand = u if x[0] < T else G[u][j]
                if cand not in row[:k]:
                    row[k] = cand
                    break
                r += 1
                assert r < MAX_RETRY
    return np.array([v for row in G for v in row], dtype=np.int32)


def run_reference_scm(n: int, d: int, p: float, workdir: str, timeout: float = 600.0):
    """Runs the reference's own serial copy model (MT19937-64 draws), built by
    ``make -C oracle ref``. Returns (degree->count dict, compute seconds)."""
    if not os.path.exists(REF_BIN):
        raise FileNotFoundError(REF_BIN)
    out = subprocess.run([REF_BIN, str(n), str(d), repr(p), workdir + "/"], check=True,
                         capture_output=True, text=True, timeout=timeout).stdout
    secs = None
    for line in out.splitlines():
        if line.startswith("Computation Time"):
            secs = float(line.split(":")[1])
    dd = {}
    with open(os.path.join(workdir, "scm_dd.csv")) as f:
        next(f)
        for line in f:
            a, b = line.strip().split(",")
            dd[int(a)] = int(b)
    return dd, secs


def run_reference_sba(n: int, d: int, workdir: str, timeout: float = 600.0):
    """Runs the reference's own serial Barabasi-Albert generator (MT19937-64
    draws), built by ``make -C oracle ref``. Returns (degree->count dict,
    compute seconds)."""
    if not os.path.exists(REF_BA_BIN):
        raise FileNotFoundError(REF_BA_BIN)
    out = subprocess.run([REF_BA_BIN, str(n), str(d), workdir + "/"], check=True,
                         capture_output=True, text=True, timeout=timeout).stdout
    secs = None
    for line in out.splitlines():
        if line.startswith("Computation Time"):
            secs = float(line.split(":")[1])
    dd = {}
    with open(os.path.join(workdir, "sba_dd.csv")) as f:
        next(f)
        for line in f:
            a, b = line.strip().split(",")
            dd[int(a)] = int(b)
    return dd, secs
