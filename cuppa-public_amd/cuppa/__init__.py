"""Python binding of the MI355X copy-model generator (ctypes over include/ppa.h).

The C ABI in ``lib/libppa.so`` is the product; this module is a thin host-side
wrapper used by the tests, the benchmark and Python callers. Device memory and
streams come from PyTorch (plumbing only). Nothing here computes graph values:
if the native library is missing, every entry point raises.

It mirrors the reference's generator object (src/PA/cuPPA/cuPPA.cuh:84-105):
``CuPPA(N, M, p)`` with ``generate_graph()`` / ``generate_graph_adaptive()``
(cuPPA_1.cu:9-254, cuPPA_2.cu:9-241) and ``degree_distribution()``
(cuPPA.cu:103-165), except that the graph is returned instead of freed.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libppa.so")

PPA_OK = 0
PPA_MAX_D = 8192
PPA_NARROW_MAX_D = 32  # d above this runs the wavefront-per-node generator
PPA_MAX_N = 0x7FFFFFFF
PPA_MAX_RETRY = 1 << 24

ERRORS = {
    -1: "PPA_ERR_INVALID", -2: "PPA_ERR_RETRY", -3: "PPA_ERR_NOMEM", -4: "PPA_ERR_DEVICE",
    -5: "PPA_ERR_TIMEOUT", -6: "PPA_ERR_BUDGET", -7: "PPA_ERR_NOT_FOUND", -8: "PPA_ERR_CANCELLED",
    -9: "PPA_ERR_AGAIN", -10: "PPA_ERR_INTERNAL", -11: "PPA_ERR_RANGE",
}


class PPAError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = f"{ERRORS.get(code, code)}: {strerror(code)}"
        super().__init__(f"{what}: {msg}" if what else msg)


MODEL_COPY, MODEL_BA = 0, 1  # PPA_MODEL_*


class GraphDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("d", ctypes.c_uint32), ("model", ctypes.c_uint32),
                ("p_threshold", ctypes.c_uint64), ("seed", ctypes.c_uint64),
                ("targets_offset", ctypes.c_uint64), ("row_ptr_offset", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [("block_threads", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("chunk_nodes", ctypes.c_uint32), ("max_blocks", ctypes.c_uint32),
                ("watchdog_ms", ctypes.c_uint32), ("early_nodes", ctypes.c_uint32),
                ("sleep_cap", ctypes.c_uint32), ("prefix_nodes", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint64), ("slots", ctypes.c_uint64), ("draws", ctypes.c_uint64),
                ("retries", ctypes.c_uint64), ("prefetch_misses", ctypes.c_uint64),
                ("polls", ctypes.c_uint64), ("stalls", ctypes.c_uint64), ("claims", ctypes.c_uint64),
                ("sleeps", ctypes.c_uint64), ("staged_rows", ctypes.c_uint64), ("passes", ctypes.c_uint64),
                ("grid_blocks", ctypes.c_uint64),
                ("block_threads", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_uint64), ("streams", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("batch_nodes", ctypes.c_uint64), ("solo_nodes", ctypes.c_uint64), ("plain", ctypes.c_uint32),
                ("solo_grid_div", ctypes.c_uint32), ("keep_results", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("n", ctypes.c_uint64), ("d", ctypes.c_uint32),
                ("model", ctypes.c_uint32), ("p_threshold", ctypes.c_uint64), ("seed", ctypes.c_uint64)]


class Result(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("targets", ctypes.c_void_p),
                ("row_ptr", ctypes.c_void_p), ("n", ctypes.c_uint64), ("d", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)

_lib = None

# (name, restype, argtypes) of every exported entry point
_VP = ctypes.c_void_p
_SIGNATURES = [
    ("ppa_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("ppa_p_threshold", ctypes.c_uint64, [ctypes.c_double]),
    ("ppa_check_desc", ctypes.c_int, [ctypes.POINTER(GraphDesc)]),
    ("ppa_ctx_create", ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.c_uint64]),
    ("ppa_ctx_destroy", ctypes.c_int, [_VP]),
    ("ppa_ctx_trim", ctypes.c_int, [_VP]),
    ("ppa_ctx_memory", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    ("ppa_ctx_set_tuning", ctypes.c_int, [_VP, ctypes.POINTER(Tuning)]),
    ("ppa_ctx_get_tuning", ctypes.c_int, [_VP, ctypes.POINTER(Tuning)]),
    ("ppa_get_stats", ctypes.c_int, [_VP, ctypes.POINTER(Stats)]),
    ("ppa_generate", ctypes.c_int, [_VP, ctypes.POINTER(GraphDesc), _VP, _VP, _VP]),
    ("ppa_generate_async", ctypes.c_int, [_VP, ctypes.POINTER(GraphDesc), _VP, _VP, _VP]),
    ("ppa_generate_noalloc", ctypes.c_int, [ctypes.POINTER(GraphDesc), _VP, _VP, ctypes.POINTER(Stats), _VP]),
    ("ppa_ctx_sync", ctypes.c_int, [_VP, _VP]),
    ("ppa_generate_range", ctypes.c_int,
     [_VP, ctypes.POINTER(GraphDesc), _VP, _VP, ctypes.c_uint64, ctypes.c_uint64, _VP]),
    ("ppa_degree_histogram", ctypes.c_int,
     [_VP, ctypes.POINTER(GraphDesc), _VP, _VP, _VP, ctypes.c_uint64, _VP, _VP]),
    ("ppa_histogram_of_degrees", ctypes.c_int, [_VP, ctypes.c_uint64, _VP, _VP, ctypes.c_uint64, _VP, _VP]),
    ("ppa_in_csr", ctypes.c_int, [_VP, ctypes.POINTER(GraphDesc), _VP, _VP, _VP, _VP]),
    ("ppa_undirected_csr", ctypes.c_int, [_VP, ctypes.POINTER(GraphDesc), _VP, _VP, _VP, _VP]),
    ("ppa_batch_create", ctypes.c_int,
     [_VP, ctypes.POINTER(GraphDesc), ctypes.c_uint64, _VP, _VP, ctypes.POINTER(_VP)]),
    ("ppa_batch_launch", ctypes.c_int, [_VP, _VP]),
    ("ppa_batch_sync", ctypes.c_int, [_VP, _VP]),
    ("ppa_batch_completion_ms", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.c_uint64]),
    ("ppa_batch_destroy", ctypes.c_int, [_VP]),
    ("ppa_generate_batch", ctypes.c_int, [_VP, ctypes.POINTER(GraphDesc), ctypes.c_uint64, _VP, _VP, _VP]),
    ("ppa_autotune", ctypes.c_int,
     [_VP, ctypes.POINTER(GraphDesc), ctypes.c_uint64, _VP, _VP, _VP, ctypes.POINTER(Tuning)]),
    ("ppa_server_create", ctypes.c_int, [_VP, ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)]),
    ("ppa_server_destroy", ctypes.c_int, [_VP]),
    ("ppa_server_submit", ctypes.c_int, [_VP, ctypes.POINTER(Request)]),
    ("ppa_server_step", ctypes.c_int, [_VP]),
    ("ppa_server_poll", ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.POINTER(Result)]),
    ("ppa_server_release", ctypes.c_int, [_VP, ctypes.c_uint64]),
    ("ppa_server_cancel", ctypes.c_int, [_VP, ctypes.c_uint64]),
    ("ppa_server_usage", ctypes.c_int, [_VP] + [ctypes.POINTER(ctypes.c_uint64)] * 4),
    ("ppa_server_replay", ctypes.c_int, [_VP, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double),
                                         ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
]


def lib():
    """Loads lib/libppa.so; raises if it was not built (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    # torch ships its own libamdhip64.so.7. Loaded after ours it would be a
    # second HIP runtime in the process that finds no GPU; loaded first, the
    # library binds to it by soname and both share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return [name for name, _, _ in _SIGNATURES]


def strerror(code: int) -> str:
    try:
        return lib().ppa_strerror(code).decode()
    except ImportError:
        return ERRORS.get(code, str(code))


def p_threshold(p: float) -> int:
    return int(lib().ppa_p_threshold(p))


def make_desc(n: int, d: int, p: float | None = None, seed: int = 0, p_thresh: int | None = None,
              targets_offset: int = 0, row_ptr_offset: int = 0, model: int = MODEL_COPY) -> GraphDesc:
    """Descriptor of one graph; a Barabasi-Albert graph (model=MODEL_BA) needs no p."""
    if p_thresh is None:
        if p is None and model != MODEL_BA:
            raise ValueError("give p or p_thresh")
        p_thresh = p_threshold(p) if p is not None else 0
    return GraphDesc(n, d, model, p_thresh, seed & 0xFFFFFFFFFFFFFFFF, targets_offset, row_ptr_offset)


def _check(rc: int, what: str):
    if rc != PPA_OK:
        raise PPAError(rc, what)


def _stream_handle(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


@dataclass
class Graph:
    n: int
    d: int
    targets: "object"  # torch.int32 [n*d] on the device
    row_ptr: "object"  # torch.int64 [n+1] on the device, or None
    model: int = MODEL_COPY


class Context:
    """One device context (ppa_ctx). scratch_budget caps the work memory of
    the degree pass and the CSR forms (0 = no cap)."""

    def __init__(self, device: int = 0, scratch_budget: int = 0):
        self._h = ctypes.c_void_p()
        self.device = device
        _check(lib().ppa_ctx_create(ctypes.byref(self._h), device, scratch_budget), "ppa_ctx_create")

    def trim(self):
        """Frees the work memory the context keeps between calls (ppa_ctx_trim)."""
        _check(lib().ppa_ctx_trim(self._h), "ppa_ctx_trim")

    def memory(self) -> dict:
        """{'held': bytes of work memory held, 'budget': its cap (0 = none)}."""
        held, budget = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().ppa_ctx_memory(self._h, ctypes.byref(held), ctypes.byref(budget)), "ppa_ctx_memory")
        return {"held": held.value, "budget": budget.value}

    def close(self):
        if self._h:
            lib().ppa_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_tuning(self, block_threads=0, blocks_per_cu=0, chunk_nodes=0, max_blocks=0, watchdog_ms=0,
                   early_nodes=0, sleep_cap=0, prefix_nodes=0):
        t = Tuning(block_threads, blocks_per_cu, chunk_nodes, max_blocks, watchdog_ms, early_nodes, sleep_cap,
                   prefix_nodes)
        _check(lib().ppa_ctx_set_tuning(self._h, ctypes.byref(t)), "ppa_ctx_set_tuning")

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib().ppa_ctx_get_tuning(self._h, ctypes.byref(t)), "ppa_ctx_get_tuning")
        return t.as_dict()

    def stats(self) -> dict:
        s = Stats()
        _check(lib().ppa_get_stats(self._h, ctypes.byref(s)), "ppa_get_stats")
        return s.as_dict()

    def generate_into(self, desc: GraphDesc, targets, row_ptr=None, stream=None, sync=True):
        """targets / row_ptr: device tensors (int32 [n*d], int64 [n+1]) or raw pointers."""
        tp = targets.data_ptr() if hasattr(targets, "data_ptr") else int(targets)
        rp = (row_ptr.data_ptr() if hasattr(row_ptr, "data_ptr") else int(row_ptr)) if row_ptr is not None else None
        fn = lib().ppa_generate if sync else lib().ppa_generate_async
        _check(fn(self._h, ctypes.byref(desc), ctypes.c_void_p(tp), ctypes.c_void_p(rp) if rp else None,
                  _stream_handle(stream)), "ppa_generate")

    def sync(self, stream=None):
        _check(lib().ppa_ctx_sync(self._h, _stream_handle(stream)), "ppa_ctx_sync")

    def generate_range(self, desc: GraphDesc, targets, row_ptr, first: int, last: int, stream=None):
        """Enqueues nodes [first, last) of a graph generated range by range
        (ppa_generate_range); sync() ends it, after which rows < last are final."""
        tp = targets.data_ptr() if hasattr(targets, "data_ptr") else int(targets)
        rp = (row_ptr.data_ptr() if hasattr(row_ptr, "data_ptr") else int(row_ptr)) if row_ptr is not None else None
        _check(lib().ppa_generate_range(self._h, ctypes.byref(desc), ctypes.c_void_p(tp),
                                        ctypes.c_void_p(rp) if rp else None, first, last, _stream_handle(stream)),
               "ppa_generate_range")

    def generate(self, n: int, d: int, p: float | None = None, seed: int = 0, p_thresh: int | None = None,
                 row_ptr: bool = True, stream=None, model: int = MODEL_COPY) -> Graph:
        import torch
        desc = make_desc(n, d, p, seed, p_thresh, model=model)
        _check(lib().ppa_check_desc(ctypes.byref(desc)), "ppa_check_desc")
        dev = torch.device("cuda", self.device)
        targets = torch.empty(n * d, dtype=torch.int32, device=dev)
        rp = torch.empty(n + 1, dtype=torch.int64, device=dev) if row_ptr else None
        self.generate_into(desc, targets, rp, stream)
        return Graph(n, d, targets, rp, model)

    def in_csr(self, g: Graph, stream=None):
        """In-edge CSR (ppa_in_csr): (in_ptr int64 [n+1], in_src int32 [in_ptr[n]])."""
        return _in_csr(self, g, stream)

    def undirected_csr(self, g: Graph, stream=None):
        """Undirected lists (ppa_undirected_csr): (ptr int64 [n+1], adj int32 [ptr[n]])."""
        import torch
        dev = torch.device("cuda", self.device)
        desc = make_desc(g.n, g.d, p_thresh=0, model=g.model)
        ptr = torch.empty(g.n + 1, dtype=torch.int64, device=dev)
        adj = torch.empty(max(1, 2 * g.n * g.d - (g.d + 1) * g.d), dtype=torch.int32, device=dev)
        _check(lib().ppa_undirected_csr(self._h, ctypes.byref(desc), ctypes.c_void_p(g.targets.data_ptr()),
                                        ctypes.c_void_p(ptr.data_ptr()), ctypes.c_void_p(adj.data_ptr()),
                                        _stream_handle(stream)), "ppa_undirected_csr")
        return ptr, adj

    def degree_histogram(self, g: Graph, hist_len: int | None = None, want_degree: bool = False, stream=None):
        """Returns (hist uint64 tensor, max_degree int, degree tensor or None).
        Without hist_len the slots are counted once (ppa_degree_histogram) and
        the histogram is sized from the maximum (ppa_histogram_of_degrees)."""
        import torch
        dev = torch.device("cuda", self.device)
        desc = make_desc(g.n, g.d, p_thresh=0, model=g.model)
        deg = torch.empty(g.n, dtype=torch.int32, device=dev) if (want_degree or hist_len is None) else None
        mx = torch.zeros(1, dtype=torch.int64, device=dev)
        sh = _stream_handle(stream)
        if hist_len is None:
            _check(lib().ppa_degree_histogram(self._h, ctypes.byref(desc), ctypes.c_void_p(g.targets.data_ptr()),
                                              ctypes.c_void_p(deg.data_ptr()), None, 0,
                                              ctypes.c_void_p(mx.data_ptr()), sh), "ppa_degree_histogram")
            hist_len = int(mx.item()) + 1
            hist = torch.zeros(hist_len, dtype=torch.int64, device=dev)
            _check(lib().ppa_histogram_of_degrees(self._h, g.n, ctypes.c_void_p(deg.data_ptr()),
                                                  ctypes.c_void_p(hist.data_ptr()), hist_len,
                                                  ctypes.c_void_p(mx.data_ptr()), sh), "ppa_histogram_of_degrees")
        else:
            hist = torch.zeros(hist_len, dtype=torch.int64, device=dev)
            _check(lib().ppa_degree_histogram(self._h, ctypes.byref(desc), ctypes.c_void_p(g.targets.data_ptr()),
                                              ctypes.c_void_p(deg.data_ptr()) if deg is not None else None,
                                              ctypes.c_void_p(hist.data_ptr()), hist_len,
                                              ctypes.c_void_p(mx.data_ptr()), sh), "ppa_degree_histogram")
        if not want_degree:
            deg = None
        return hist, int(mx.item()), deg


def _in_csr(ctx, g: "Graph", stream=None):
    import torch
    dev = torch.device("cuda", ctx.device)
    desc = make_desc(g.n, g.d, p_thresh=0, model=g.model)
    in_ptr = torch.empty(g.n + 1, dtype=torch.int64, device=dev)
    in_src = torch.empty(g.n * g.d, dtype=torch.int32, device=dev)
    _check(lib().ppa_in_csr(ctx._h, ctypes.byref(desc), ctypes.c_void_p(g.targets.data_ptr()),
                            ctypes.c_void_p(in_ptr.data_ptr()), ctypes.c_void_p(in_src.data_ptr()),
                            _stream_handle(stream)), "ppa_in_csr")
    return in_ptr, in_src[: int(in_ptr[-1].item())]


def batch_layout(specs):
    """specs: list of (n, d, p_threshold, seed[, model]). Returns (GraphDesc
    array, total target slots, total row_ptr entries) with graphs laid out back
    to back."""
    descs = (GraphDesc * len(specs))()
    toff = roff = 0
    for i, spec in enumerate(specs):
        n, d, T, seed = spec[:4]
        model = spec[4] if len(spec) > 4 else MODEL_COPY
        descs[i] = GraphDesc(n, d, model, T, seed & 0xFFFFFFFFFFFFFFFF, toff, roff)
        toff += n * d
        roff += n + 1
    return descs, toff, roff


class Batch:
    """A prepared batch (ppa_batch): descriptor table on the device and a
    captured hipGraph of reset + fill + generate."""

    def __init__(self, ctx: "Context", descs, targets, row_ptr=None):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        self.descs = descs
        self.targets, self.row_ptr = targets, row_ptr
        _check(lib().ppa_batch_create(ctx.handle, descs, len(descs), ctypes.c_void_p(targets.data_ptr()),
                                      ctypes.c_void_p(row_ptr.data_ptr()) if row_ptr is not None else None,
                                      ctypes.byref(self._h)), "ppa_batch_create")

    def launch(self, stream=None):
        _check(lib().ppa_batch_launch(self._h, _stream_handle(stream)), "ppa_batch_launch")

    def sync(self, stream=None):
        _check(lib().ppa_batch_sync(self._h, _stream_handle(stream)), "ppa_batch_sync")

    def run(self, stream=None):
        self.launch(stream)
        self.sync(stream)

    def completion_ms(self):
        """Per-graph ms from the last launch to its graph's completion
        (ppa_batch_completion_ms; -1 for a graph that never completed)."""
        out = (ctypes.c_double * len(self.descs))()
        _check(lib().ppa_batch_completion_ms(self._h, out, len(self.descs)), "ppa_batch_completion_ms")
        return list(out)

    def close(self):
        if self._h:
            lib().ppa_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
        # the output buffers were kept alive only for the native batch
        self.targets = self.row_ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx: "Context", descs, targets, row_ptr=None, stream=None) -> dict:
    t = Tuning()
    _check(lib().ppa_autotune(ctx.handle, descs, len(descs), ctypes.c_void_p(targets.data_ptr()),
                              ctypes.c_void_p(row_ptr.data_ptr()) if row_ptr is not None else None,
                              _stream_handle(stream), ctypes.byref(t)), "ppa_autotune")
    return t.as_dict()


class Server:
    """Continuous-batching request loop (ppa_server)."""

    def __init__(self, ctx: "Context", budget_bytes: int, streams: int = 0, max_batch: int = 0,
                 batch_nodes: int = 0, solo_nodes: int = 0, plain: bool = False, solo_grid_div: int = 0,
                 keep_results: bool = False):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        cfg = ServerConfig(budget_bytes, streams, max_batch, batch_nodes, solo_nodes, 1 if plain else 0,
                           solo_grid_div, 1 if keep_results else 0, 0)
        _check(lib().ppa_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(self._h)), "ppa_server_create")

    def submit(self, rid: int, n: int, d: int, p_thresh: int, seed: int, model: int = MODEL_COPY) -> int:
        return lib().ppa_server_submit(self._h, ctypes.byref(Request(rid, n, d, model, p_thresh, seed)))

    def step(self) -> int:
        rc = lib().ppa_server_step(self._h)
        if rc < 0:
            raise PPAError(rc, "ppa_server_step")
        return rc

    def poll(self, rid: int) -> Result:
        r = Result()
        _check(lib().ppa_server_poll(self._h, rid, ctypes.byref(r)), "ppa_server_poll")
        return r

    def release(self, rid: int) -> int:
        return lib().ppa_server_release(self._h, rid)

    def cancel(self, rid: int) -> int:
        return lib().ppa_server_cancel(self._h, rid)

    def usage(self) -> dict:
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(lib().ppa_server_usage(self._h, *[ctypes.byref(x) for x in v]), "ppa_server_usage")
        return dict(zip(("used", "budget", "queued", "running"), (x.value for x in v)))

    def replay(self, specs, arrival_ms, with_wait=False):
        """specs: list of (id, n, d, p_threshold, seed[, model]). Returns
        (latency_ms list, makespan_ms), or (latency_ms, wait_ms, makespan_ms)
        with with_wait."""
        count = len(specs)
        reqs = (Request * count)(*[Request(sp[0], sp[1], sp[2], sp[5] if len(sp) > 5 else MODEL_COPY, sp[3],
                                           sp[4]) for sp in specs])
        arr = (ctypes.c_double * count)(*arrival_ms)
        lat = (ctypes.c_double * count)()
        wait = (ctypes.c_double * count)()
        mk = ctypes.c_double()
        _check(lib().ppa_server_replay(self._h, reqs, arr, count, lat, wait, ctypes.byref(mk)),
               "ppa_server_replay")
        if with_wait:
            return list(lat), list(wait), mk.value
        return list(lat), mk.value

    def close(self):
        if self._h:
            lib().ppa_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hip = None


def read_device(ptr: int, count: int, np_dtype):
    """Copies `count` elements of device memory at `ptr` into a new numpy array
    (hipMemcpy device-to-host)."""
    import numpy as np
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    out = np.empty(count, dtype=np_dtype)
    if count:
        rc = _hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)  # hipMemcpyDeviceToHost
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
    return out


class CuPPA:
    """Mirror of the reference's ``cuPPA`` object (cuPPA.cuh:84-105).

    ``CuPPA(N, M, p)`` keeps the constructor's argument meaning: N nodes, M
    out-edges per new node, p the probability of a direct edge (the ``-a``
    flag, src/main.cu:198). ``seed`` keys the Philox draws (the reference
    seeds with ``time(0)``, cuPPA_1.cu:39).
    """

    def __init__(self, N: int, M: int, p: float, seed: int = 0, device: int = 0):
        self._N, self._M, self._p, self.seed = N, M, p, seed
        self.ctx = Context(device)
        self.graph: Graph | None = None

    def generate_graph(self) -> Graph:
        self.graph = self.ctx.generate(self._N, self._M, self._p, self.seed)
        return self.graph

    # The adaptive stage sizing of cuPPA_2.cu:181-198 has no counterpart to
    # tune: one persistent launch covers every stage; the output is the same.
    generate_graph_adaptive = generate_graph

    def degree_distribution(self) -> dict:
        if self.graph is None:
            raise RuntimeError("generate_graph() first")
        hist, _, _ = self.ctx.degree_histogram(self.graph)
        h = hist.cpu().tolist()
        return {deg: c for deg, c in enumerate(h) if c > 0}


class SerialBA:
    """Mirror of the reference's ``SerialBA`` object (SerialBA.h:30-51):
    ``SerialBA(n, m)`` is the Barabasi-Albert generator of ``-A sba`` with n
    nodes and m edges per new node, run here on the GPU as a PPA_MODEL_BA
    graph. ``seed`` keys the draws (the reference draws from MT19937-64 with
    its default seed, Random64.hpp:75-76)."""

    def __init__(self, n: int, m: int = 1, seed: int = 0, device: int = 0):
        self._N, self._M, self.seed = n, m, seed
        self.ctx = Context(device)
        self.graph: Graph | None = None

    def generate_graph(self) -> Graph:
        self.graph = self.ctx.generate(self._N, self._M, seed=self.seed, model=MODEL_BA)
        return self.graph

    def degree_distribution(self) -> dict:
        """Degree -> count as PAHelper::outputSBADegreeDistribution writes it
        (PAHelper.hpp:69-147)."""
        if self.graph is None:
            raise RuntimeError("generate_graph() first")
        hist, _, _ = self.ctx.degree_histogram(self.graph)
        h = hist.cpu().tolist()
        return {deg: c for deg, c in enumerate(h) if c > 0}
