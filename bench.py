#!/usr/bin/env python3
"""Benchmark of the MI355X copy-model generator (one JSON line on stdout).

Workload: the single-graph copy model at n = 2^30 nodes, d = 8 slots per node,
p = 0.2, Philox seed 11 (34.4 GB of int32 targets plus 8.6 GB of int64 row
offsets, resident in HBM). One step = one ppa_generate call (sentinel fill +
persistent generation kernel) into the same device buffers. The metric is
edges/s = n*d / step time, with the generation kernel's HBM-roofline fraction
measured live with HIP events recorded on the launch stream.

The CPU baseline runs the reference's own serial copy model
(src/PA/SPA/SerialCopyModel3.cpp, built by `make -C oracle ref`) on one core
over a bounded prefix of the same parameters.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuppa-public_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E vendor figure

WORKLOADS = {
    # name: (n, d, p, seed)
    "n2p30_d8_p0.2": (1 << 30, 8, 0.2, 11),
    "n2p28_d8_p0.5": (1 << 28, 8, 0.5, 7),
    "n1e6_d4_p0.5": (1_000_000, 4, 0.5, 42),
}


def algorithmic_bytes_per_slot(p: float, d: int) -> dict:
    """Bytes per generated slot (edge), per kernel and for the whole call."""
    gen = 4.0 + 4.0 * (1.0 - p)       # int32 target write + copy-source read
    row_ptr = 8.0 / d                 # int64 row offset write (fill kernel)
    return {"generate_kernel": gen, "call": gen + row_ptr}


def read_traffic(profile_dir: str, workload: str):
    """Measured HBM bytes per launch of the generation kernel, from the
    committed rocprofv3 --pmc summary for this workload (None if absent)."""
    path = os.path.join(profile_dir, f"generate_kernel_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            m = re.match(r"\s*traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", line)
            if m:
                return float(m.group(1))
    return None


def host_info():
    """The host the CPU baseline ran on."""
    info = {"nproc": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            info["cpu_model"] = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    try:
        with open("/proc/meminfo") as f:
            kb = next(int(l.split()[1]) for l in f if l.startswith("MemTotal"))
        info["ram_gb"] = round(kb / 2**20, 1)
    except (OSError, StopIteration):
        pass
    return info


def _run_ref(n: int, d: int, p: float):
    """One run of the reference serial copy model on core 0; its own timer."""
    ref = os.path.join(ROOT, "oracle", "_ref", "scm_ref")
    with tempfile.TemporaryDirectory() as tmp:
        cmd = [ref, str(n), str(d), repr(p), tmp + "/", "0"]
        cmd = ["taskset", "-c", "0"] + cmd if _has("taskset") else cmd
        out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300).stdout
        return float(re.search(r"Computation Time\s*:\s*([0-9.eE+-]+)", out).group(1))


def cpu_mix_baseline(graphs, what: str):
    """Reference serial copy model (one core) over a sample of a graph mix:
    graphs = [(n, d, p)]; edges/s = sample edges / summed run time."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "scm_ref")):
        return None
    secs = sum(_run_ref(n, d, p) for n, d, p in graphs)
    edges = sum(n * d for n, d, _ in graphs)
    return {"value": edges / secs, "unit": "edges/s", "cores": 1, "kind": "reference",
            "sample": f"reference SerialCopyModel (MT19937-64) on {len(graphs)} graphs of the {what} "
                      f"({edges:.3g} edges)", "seconds": secs, "host": host_info()}


def cpu_baseline(n: int, d: int, p: float, prefix: int):
    """Reference serial copy model (one core) over the first `prefix` nodes."""
    ref = os.path.join(ROOT, "oracle", "_ref", "scm_ref")
    sample = f"reference SerialCopyModel (MT19937-64), first {prefix} nodes of n={n}, d={d}, p={p}"
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as tmp:
            cmd = [ref, str(prefix), str(d), repr(p), tmp + "/", "0"]
            try:
                cmd = ["taskset", "-c", "0"] + cmd if _has("taskset") else cmd
                out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300).stdout
                secs = float(re.search(r"Computation Time\s*:\s*([0-9.eE+-]+)", out).group(1))
                return {"value": prefix * d / secs, "unit": "edges/s", "cores": 1, "kind": "reference",
                        "sample": sample, "seconds": secs, "host": host_info()}
            except Exception as e:  # fall through to the port
                sample += f" (reference run failed: {e})"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    T = oracle.p_threshold(p)
    t0 = time.perf_counter()
    oracle.generate(prefix, d, T, 11)
    secs = time.perf_counter() - t0
    return {"value": prefix * d / secs, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"oracle restatement, first {prefix} nodes of n={n}, d={d}, p={p}", "seconds": secs,
            "host": host_info()}


def _has(tool: str) -> bool:
    from shutil import which
    return which(tool) is not None


def copy_bandwidth_gbs(torch, nbytes: int = 1 << 32) -> float:
    a = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(5):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / 5
    del a, b
    return 2 * nbytes / (ms * 1e-3) / 1e9


def gather_ceiling(torch, table_bytes: int = 32 << 30, reads: int = 1 << 28) -> dict:
    """Random 4-byte reads per second on this GPU over an int32 table of
    `table_bytes` (each read is one 64-byte HBM request, like the generator's
    copy reads): the hand-written probe kernels of lib/libppa_probe.so (plain
    and non-temporal loads into registers or into LDS, best of 3 each; and the
    LDS form with one 2 KB block of row stores per wavefront and batch of
    eight reads, the generator's own mix of reads and publication) and, for
    comparison, torch's index_select at `reads` uniform random positions.
    requests_per_s is the best pure-read rate."""
    import ctypes
    src = torch.zeros(table_bytes // 4, dtype=torch.int32, device="cuda")
    out = {"table_bytes": table_bytes}
    probe_path = os.path.join(ROOT, "cuppa-public_amd", "lib", "libppa_probe.so")
    if os.path.exists(probe_path):
        P = ctypes.CDLL(probe_path)
        P.ppa_probe_gather.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_float)]
        P.ppa_probe_gather.restype = ctypes.c_int
        blocks, threads, iters = 2048, 256, 512
        sink = torch.empty(blocks * threads, dtype=torch.int32, device="cuda")
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for mode, name in enumerate(PROBE_MODES):
            if name == "philox_draws":
                continue
            best = float("inf")
            for _ in range(4):  # the first launch warms up
                ms = ctypes.c_float()
                rc = P.ppa_probe_gather(ctypes.c_void_p(src.data_ptr()), src.numel(), blocks, threads, iters, mode,
                                        ctypes.c_void_p(sink.data_ptr()), stream, ctypes.byref(ms))
                if rc != 0:
                    raise RuntimeError(f"ppa_probe_gather failed: {rc}")
                best = min(best, ms.value)
            out[name] = blocks * threads * iters / (best * 1e-3)
        del sink
    idx = torch.randint(0, src.numel(), (reads,), dtype=torch.int64, device="cuda")
    got = src[idx]
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(3):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        torch.index_select(src, 0, idx, out=got)
        e.record()
        torch.cuda.synchronize()
        best = min(best, s.elapsed_time(e))
    out["torch_index_select"] = reads / (best * 1e-3)
    del src, idx, got
    torch.cuda.empty_cache()
    out["requests_per_s"] = max(v for k, v in out.items() if k not in ("table_bytes", "probe_lds_nt_with_row_stores"))
    return out


# ppa_probe_gather's modes, in order (mode i = PROBE_MODES[i])
PROBE_MODES = ("probe_plain", "probe_nt", "probe_lds", "probe_lds_nt", "probe_lds_nt_with_row_stores",
               "probe_lds_sc0", "probe_lds_sc1", "probe_lds_sc0_sc1", "probe_lds_nt_sc0", "probe_lds_nt_sc1",
               "probe_lds_nt_sc0_sc1", "philox_draws")


def philox_roof(torch, blocks: int = 8192, threads: int = 256, iters: int = 4096) -> dict:
    """The integer-VALU roof of the generator's draws: keyed Philox-4x32-10
    attempts per second (coin and mulhi64 pair mapping included) with no
    memory traffic, the whole GPU busy (probe mode 11, best of 3)."""
    import ctypes
    probe_path = os.path.join(ROOT, "cuppa-public_amd", "lib", "libppa_probe.so")
    if not os.path.exists(probe_path):
        return None
    P = ctypes.CDLL(probe_path)
    P.ppa_probe_gather.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_float)]
    P.ppa_probe_gather.restype = ctypes.c_int
    sink = torch.empty(blocks * threads, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    best = float("inf")
    for _ in range(4):
        ms = ctypes.c_float()
        rc = P.ppa_probe_gather(ctypes.c_void_p(sink.data_ptr()), sink.numel(), blocks, threads, iters,
                                PROBE_MODES.index("philox_draws"), ctypes.c_void_p(sink.data_ptr()), stream,
                                ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(f"ppa_probe_gather failed: {rc}")
        best = min(best, ms.value)
    del sink
    return {"draws_per_s": blocks * threads * iters / (best * 1e-3), "lanes": blocks * threads,
            "draws_per_lane": iters}


def run_probe(args):
    """Each probe kernel once over a 32 GiB table, for counter runs
    (rocprofv3 --pmc): the gather forms and the Philox draws."""
    import torch
    out = {"gather": gather_ceiling(torch), "philox": philox_roof(torch)}
    print(json.dumps(out))


def read_requests(profile_dir: str, workload: str):
    """TCC read requests per launch of the generation kernel, from the
    committed rocprofv3 --pmc summary for this workload (None if absent)."""
    path = os.path.join(profile_dir, f"generate_kernel_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            m = re.match(r"\s*tcc_ea_read_requests\s*[:=]\s*([0-9.eE+]+)", line)
            if m:
                return float(m.group(1))
    return None


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def batch_object(args, ctx):
    """The ensemble in one captured persistent launch against the same graphs
    generated by plain ppa_generate calls, one after another. Both sides are
    the median of `runs` full passes after a warm-up. Latency is completion
    time with the whole ensemble submitted at once: per graph, from the launch
    to its completion flag (batched) or to the end of its own call (plain),
    plus the plain calls' own durations."""
    import statistics
    import torch
    import cuppa
    from cuppa import workloads
    runs = max(3, args.batch_runs)
    specs = workloads.batch_specs(args.batch_count)
    descs, tot, rot = cuppa.batch_layout(specs)
    edges = sum(n * d for n, d, _, _ in specs)
    targets = torch.empty(tot, dtype=torch.int32, device="cuda")
    row_ptr = torch.empty(rot, dtype=torch.int64, device="cuda")
    tuned = cuppa.autotune(ctx, descs, targets, row_ptr) if args.batch_autotune else ctx.get_tuning()
    b = cuppa.Batch(ctx, descs, targets, row_ptr)
    b.run()  # warm-up: also the first touch of the output buffers
    times, dones, kernel = [], [], []
    for _ in range(runs):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run()
        times.append(time.perf_counter() - t0)
        dones.append(b.completion_ms())
        kernel.append(ctx.stats()["kernel_ms"])
    st = ctx.stats()
    b.close()
    del b, targets, row_ptr
    torch.cuda.empty_cache()
    order = sorted(range(runs), key=lambda i: times[i])
    med = order[runs // 2]
    batched = times[med]
    done = dones[med]
    # plain: one ppa_generate per graph into reused buffers, with the default
    # launch rules (the single-graph tuning)
    ctx.set_tuning()
    maxn = max(n for n, _, _, _ in specs)
    scratch = torch.empty(maxn * 16, dtype=torch.int32, device="cuda")
    scratch_rp = torch.empty(maxn + 1, dtype=torch.int64, device="cuda")
    plain_descs = [cuppa.GraphDesc(n, d, 0, T, seed, 0, 0) for n, d, T, seed in specs]
    for ds in plain_descs[:64]:  # warm-up
        ctx.generate_into(ds, scratch, scratch_rp)
    passes = []
    for _ in range(runs):
        torch.cuda.synchronize()
        per_call = []
        t_start = time.perf_counter()
        for ds in plain_descs:
            t0 = time.perf_counter()
            ctx.generate_into(ds, scratch, scratch_rp)  # waits for its stream
            per_call.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        passes.append((time.perf_counter() - t_start, per_call))
    del scratch, scratch_rp
    torch.cuda.empty_cache()
    passes.sort(key=lambda x: x[0])
    plain, per_call = passes[runs // 2]
    import numpy as np
    completion_plain = np.cumsum(per_call) * 1e3
    cpu = cpu_mix_baseline([(n, d, T / 2**32) for n, d, T, _ in specs[::256]], "ensemble") \
        if args.cpu_prefix else None
    return {
        "workload": f"{len(specs)} copy-model graphs, n log-uniform [1e3,1e7], d 1..16, p [0.05,0.95], "
                    f"seed 1000+i", "graphs": len(specs), "edges": edges,
        "batched_ms": batched * 1e3, "batched_runs_ms": [t * 1e3 for t in times],
        "batched_kernel_ms": kernel[med], "edges_per_s": edges / batched,
        "plain_ms": plain * 1e3, "plain_runs_ms": sorted(p[0] * 1e3 for p in passes),
        "plain_edges_per_s": edges / plain,
        "throughput_ratio": plain / batched,
        "completion_latency_all_submitted_at_once": {
            "batched_p50_ms": _pct(done, 50), "batched_p99_ms": _pct(done, 99), "batched_max_ms": max(done),
            "plain_p50_ms": _pct(completion_plain, 50), "plain_p99_ms": _pct(completion_plain, 99),
            "plain_max_ms": float(completion_plain[-1])},
        "plain_per_call_ms": {"p50": _pct(per_call, 50) * 1e3, "p99": _pct(per_call, 99) * 1e3,
                              "mean": statistics.mean(per_call) * 1e3},
        "runs": runs, "tuning": tuned, "cpu_baseline": cpu,
        "stats": {k: st[k] for k in ("nodes", "draws", "retries", "prefetch_misses", "polls", "staged_rows",
                                     "grid_blocks", "block_threads")},
    }


def run_batch(args):
    import cuppa
    ctx = cuppa.Context(0)
    obj = batch_object(args, ctx)
    print(json.dumps({"metric": "batched ensemble edges/s (one hipGraph launch) vs plain per-graph launches",
                      "value": obj["edges_per_s"], "unit": "edges/s", "n_gpus": 1, "higher_is_better": True,
                      "dtype": "int32", "data": "synthetic (keyed Philox draws)", "batch": obj}))


def serve_object(args, ctx):
    """Served (continuous batching) against sequential plain launches at equal
    p99 latency, under Poisson arrivals, over the same `serve_count` requests
    on both sides: the plain loop (one request per launch, one stream) at
    saturation, then at 0.7x its saturation rate (the reference p50/p99); the
    served loop at multiples of that rate."""
    import numpy as np
    import torch
    import cuppa
    from cuppa import workloads
    torch.cuda.empty_cache()  # the pool below is allocated outside torch's cache
    trace = workloads.serve_trace(args.serve_count)
    edges_all = float(sum(n * d for _, n, d, _, _ in trace))
    budget = args.serve_budget_gb << 30
    sizes = np.array([n for _, n, _, _, _ in trace])

    def replay(plain, rate):
        arr = workloads.poisson_arrivals(len(trace), rate)
        srv = cuppa.Server(ctx, budget_bytes=budget, streams=1 if plain else args.serve_streams, plain=plain,
                           solo_nodes=0 if plain else args.serve_solo_nodes,
                           batch_nodes=0 if plain else args.serve_batch_nodes,
                           solo_grid_div=0 if plain else args.serve_solo_grid_div)
        lat, makespan = srv.replay(trace, arr)
        srv.close()
        lat = np.array(lat)
        by_size = {}
        for lo, hi, name in ((0, 1e5, "n<1e5"), (1e5, 1e6, "1e5<=n<1e6"), (1e6, 1e7, "1e6<=n<1e7"),
                             (1e7, 1e9, "n>=1e7")):
            m = (sizes >= lo) & (sizes < hi)
            if m.any():
                by_size[name] = {"count": int(m.sum()), "p50_ms": _pct(lat[m], 50), "p99_ms": _pct(lat[m], 99)}
        return {"rate_per_s": rate, "requests": len(trace), "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                "max_ms": float(lat.max()), "makespan_ms": makespan,
                "req_per_s": len(trace) / (makespan * 1e-3), "edges_per_s": edges_all / (makespan * 1e-3),
                "by_size": by_size}

    sat = replay(True, 0)
    plain_rate = 0.7 * sat["req_per_s"]
    plain = replay(True, plain_rate)
    served = []
    for mult in [float(m) for m in args.serve_mults.split(",")]:
        res = replay(False, plain_rate * mult)
        res["rate_multiple"] = mult
        served.append(res)
    ok = [s for s in served if s["p99_ms"] <= plain["p99_ms"]]
    best = max(ok, key=lambda s: s["edges_per_s"]) if ok else None
    return {
        "workload": "requests: n lognormal(median 1e5, sigma 1.5) clipped [1e2,1e8], d {2,4,8,16}, "
                    "p {0.1,0.5,0.9}, seed = id; Poisson arrivals", "requests": len(trace),
        "edges": edges_all, "budget_gb": args.serve_budget_gb, "streams": args.serve_streams,
        "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
        "plain_saturation": sat, "plain": plain, "served": served,
        "best_served_rate_multiple": best["rate_multiple"] if best else None,
        "throughput_ratio_at_equal_p99": (best["edges_per_s"] / plain["edges_per_s"]) if best else None,
        # the reference's serial path on the first 64 requests of the trace (one core)
        "cpu_baseline": cpu_mix_baseline([(n, d, T / 2**32) for _, n, d, T, _ in trace[:64]], "request mix")
        if args.cpu_prefix else None,
    }


def run_serve(args):
    import cuppa
    ctx = cuppa.Context(0)
    obj = serve_object(args, ctx)
    best = obj["throughput_ratio_at_equal_p99"]
    print(json.dumps({"metric": "served throughput at equal p99 latency vs sequential plain launches",
                      "value": best, "unit": "x plain edges/s", "n_gpus": 1, "higher_is_better": True,
                      "dtype": "int32", "data": "synthetic (keyed Philox draws, Poisson arrivals)", "serve": obj}))


LATENCY_GRAPHS = [(20_000, 16, 0.1), (100_000, 16, 0.1), (100_000, 16, 0.5), (1_000_000, 16, 0.1),
                  (10_000_000, 16, 0.1), (100_000_000, 16, 0.1), (100_000, 8, 0.2), (1_000_000, 8, 0.2),
                  (100_000, 6, 0.1), (100_000, 4, 0.5), (1_000_000, 2, 0.1), (100_000, 32, 0.1),
                  (1_000_000, 24, 0.5)]


def run_latency(args):
    """Wall time of one ppa_generate call (inputs resident, synchronised on
    both sides) per graph of LATENCY_GRAPHS, median of `steps` calls after
    `warmup`, with the default launch rule and LDS-resident prefix."""
    import statistics
    import torch
    import cuppa
    ctx = cuppa.Context(0)
    out = []
    for n, d, p in LATENCY_GRAPHS:
        desc = cuppa.make_desc(n, d, p, 3)
        targets = torch.empty(n * d, dtype=torch.int32, device="cuda")
        row_ptr = torch.empty(n + 1, dtype=torch.int64, device="cuda")
        ms, kernel = [], []
        for i in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.generate_into(desc, targets, row_ptr)
            if i >= args.warmup:
                ms.append((time.perf_counter() - t0) * 1e3)
                kernel.append(ctx.stats()["kernel_ms"])
        st = ctx.stats()
        out.append({"n": n, "d": d, "p": p, "call_ms": statistics.median(ms),
                    "kernel_ms": statistics.median(kernel), "grid_blocks": st["grid_blocks"]})
        del targets, row_ptr
    print(json.dumps({"metric": "single-call latency (ppa_generate wall time)", "unit": "ms",
                      "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "data": "synthetic (keyed Philox draws)", "graphs": out}))


# Out-degrees above 32 (the wavefront-stream generator), at 2^30 .. 2^31
# slots, beside d = 32 on the lane-per-node generator.
WIDE_GRAPHS = [(1 << 25, 32, 0.2), (1 << 25, 33, 0.2), (1 << 24, 64, 0.2), (1 << 23, 96, 0.2),
               (1 << 23, 128, 0.2), (1 << 22, 256, 0.2), (1 << 20, 1024, 0.2), (1 << 17, 8192, 0.5)]


def run_wide(args):
    """Wall time and kernel time of one ppa_generate call per graph of
    WIDE_GRAPHS (inputs resident, median of `steps` calls after `warmup`)."""
    import statistics
    import torch
    import cuppa
    ctx = cuppa.Context(0)
    out = []
    for n, d, p in WIDE_GRAPHS:
        desc = cuppa.make_desc(n, d, p, 7)
        targets = torch.empty(n * d, dtype=torch.int32, device="cuda")
        ms, kernel = [], []
        for i in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.generate_into(desc, targets, None)
            if i >= args.warmup:
                ms.append((time.perf_counter() - t0) * 1e3)
                kernel.append(ctx.stats()["kernel_ms"])
        st = ctx.stats()
        call = statistics.median(ms)
        out.append({"n": n, "d": d, "p": p, "call_ms": call, "kernel_ms": statistics.median(kernel),
                    "edges_per_s": n * d / (call * 1e-3), "grid_blocks": st["grid_blocks"],
                    "retries_per_node": st["retries"] / max(1, st["nodes"]),
                    "rounds_per_node": st["passes"] / max(1, st["nodes"])})
        print(f"bench: wide n={n} d={d} p={p}: {call:.2f} ms", file=sys.stderr, flush=True)
        del targets
        torch.cuda.empty_cache()
    print(json.dumps({"metric": "edges/s of single graphs by out-degree (d > 32: wavefront streams)",
                      "unit": "edges/s", "higher_is_better": True, "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "data": "synthetic (keyed Philox draws)", "graphs": out}))


def cpu_degree_csr_baselines(targets, d: int, nodes: int):
    """The oracle's degree count (the reference's host loop of
    cuPPA.cu:103-165, restated) and its in-edge counting sort, one core each,
    over the first `nodes` nodes of the generated graph: their slots name only
    nodes below them, so they form a graph of their own."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    slots = nodes * d
    G = targets[:slots].cpu().numpy()
    sample = f"oracle restatement, first {nodes} nodes ({slots} slots) of the generated graph"
    out = {}
    for key, fn in (("degree", lambda: oracle.degrees(nodes, d, G, 1 << 20)),
                    ("in_csr", lambda: oracle.in_csr(nodes, d, G))):
        t0 = time.perf_counter()
        fn()
        secs = time.perf_counter() - t0
        out[key] = {"value": slots / secs, "unit": "slots/s", "cores": 1, "kind": "port", "sample": sample,
                    "seconds": secs, "host": host_info()}
    return out


def degree_and_csr(ctx, n, d, edges, targets, row_ptr, cpu_nodes: int = 0) -> dict:
    """The degree pass and the in-edge CSR of the generated graph: the first
    call of each also allocates the context's work memory; the second is
    timed. Algorithmic bytes of the degree pass: 4 B/slot read + 8 B/slot
    degree update + 4 B/node histogram read; of the in-edge CSR: 4 B/slot
    read + 4 B/slot in_src write + 8 B/node in_ptr write. cpu_nodes > 0 adds
    the oracle's one-core rates over that many nodes as cpu_baseline."""
    import torch
    import cuppa
    out = {}
    g = cuppa.Graph(n, d, targets, row_ptr)
    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hist, mx, _ = ctx.degree_histogram(g, hist_len=1 << 26)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t1) * 1e3)
    del hist
    alg = 12.0 * edges + 4.0 * n
    out["degree"] = {"ms": times[1], "first_call_ms": times[0], "max_degree": mx,
                     "slots_per_s": edges / (times[1] * 1e-3),
                     "algorithmic_gbs": alg / (times[1] * 1e-3) / 1e9,
                     "roofline_frac": alg / (times[1] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        in_ptr, in_src = ctx.in_csr(g)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t1) * 1e3)
        del in_ptr, in_src
    alg = 8.0 * edges + 8.0 * n
    out["in_csr"] = {"ms": times[1], "first_call_ms": times[0], "edges_per_s": edges / (times[1] * 1e-3),
                     "algorithmic_gbs": alg / (times[1] * 1e-3) / 1e9,
                     "roofline_frac": alg / (times[1] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    ctx.trim()  # give the work memory back before the next measurements
    torch.cuda.empty_cache()
    if cpu_nodes:
        for key, base in cpu_degree_csr_baselines(targets, d, min(cpu_nodes, n)).items():
            out[key]["cpu_baseline"] = base
    return out


def design_floors(extra: dict, n: int, d: int, edges: int, copy_gbs: float) -> None:
    """The bytes each pass's own design moves through HBM, and the time a
    device copy's rate (copy_gbs, measured in the same run) needs for them:
    how far the degree pass and the in-edge CSR are from the bound of the way
    they work, beside their algorithmic roofline. Degree pass with two
    partition levels (n > 2^23): count 4 + level-1 scatter 4 + 4 + level-2
    scatter 4 + 2 + leaf count 2 B/slot (the hub nodes' slots, which skip
    both scatters, not taken off), 4 B/node degree write + 4 B/node histogram
    read. In-edge CSR with P = ceil(bits(n - 1) / 8) passes: P counts of 4
    B/slot, a first scatter of 4 + 8 B/slot, P - 1 scatters of 8 + 8 B/slot,
    the in_ptr read of 4 B/slot, 8 B/node in_ptr write."""
    bits = max(1, (n - 1).bit_length())
    passes = (bits + 7) // 8
    plans = {"degree": 20.0 * edges + 8.0 * n if n > (1 << 23) else None,
             "in_csr": (4.0 * passes + 12.0 + 16.0 * (passes - 1) + 4.0) * edges + 8.0 * n}
    for key, nbytes in plans.items():
        obj = extra.get(key)
        if obj is None or nbytes is None:
            continue
        floor_ms = nbytes / (copy_gbs * 1e9) * 1e3
        obj["design_bytes"] = nbytes
        obj["design_floor_ms"] = floor_ms
        obj["frac_of_design_floor"] = floor_ms / obj["ms"]


def profile_kernel_ms(profile_dir: str, workload: str):
    """The generation kernel's average duration in the committed PMC summary
    (the run its traffic figures come from), or None."""
    path = os.path.join(profile_dir, f"generate_kernel_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            m = re.match(r"\s*generate_kernel_avg_ms\s*[:=]\s*([0-9.eE+]+)", line)
            if m:
                return float(m.group(1))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="n2p30_d8_p0.2", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-prefix", type=int, default=1 << 25,
                    help="nodes of the CPU baseline sample (0 = skip every CPU baseline)")
    ap.add_argument("--no-degree", action="store_true", help="skip the degree pass and in-edge CSR timings")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble object")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-loop object")
    ap.add_argument("--no-gather-ceiling", action="store_true",
                    help="skip the random-gather ceiling measurement (needs 35 GB of free HBM)")
    ap.add_argument("--ba", action="store_true",
                    help="also time the Barabasi-Albert model (-A sba) at the workload's n and d, "
                         "and the reference's serial BA on a prefix")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB three times (counter calibration) and exit")
    ap.add_argument("--mode", default="graph", choices=("graph", "batch", "serve", "latency", "probe", "wide"),
                    help="graph: the single-graph metric with the batch and serve objects (default); "
                         "batch / serve: one of those objects alone; latency: single-call wall times of "
                         "small and mid-size graphs; wide: single graphs of out-degree 32 .. 8192")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-runs", type=int, default=3, help="timed passes of each side (median, >= 3)")
    ap.add_argument("--no-batch-autotune", dest="batch_autotune", action="store_false",
                    help="time the batch with the default launch shape instead of running the autotuner "
                         "on the ensemble first (about 25 extra launches, ~20 s)")
    ap.add_argument("--serve-count", type=int, default=10_000,
                    help="requests of each replay, plain and served alike")
    ap.add_argument("--serve-budget-gb", type=int, default=200)
    ap.add_argument("--serve-mults", default="5,6,7,8",
                    help="served arrival rates, as multiples of the plain loop's rate (0.7x its saturation)")
    ap.add_argument("--serve-solo-nodes", type=int, default=0,
                    help="requests of at least this many nodes launch alone (0 = the server's default)")
    ap.add_argument("--serve-batch-nodes", type=int, default=0,
                    help="a batch launch closes at this many nodes (0 = the server's default)")
    ap.add_argument("--serve-solo-grid-div", type=int, default=0,
                    help="lone launches get 1/this of the single-call grid (0 = the server's default)")
    ap.add_argument("--serve-streams", type=int, default=16,
                    help="concurrent launches of the served loop (HIP hardware queues are raised to match)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported; this generator runs on one GPU "
              f"(asked for {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.mode == "serve" or (args.mode == "graph" and not args.no_serve):
        # Streams beyond the runtime's hardware queues (HIP default 4) share a
        # queue and run one after another; this must be set before HIP starts.
        os.environ["GPU_MAX_HW_QUEUES"] = str(max(int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                                                  args.serve_streams))

    import torch
    import cuppa

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        a = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 3}))
        return
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "serve":
        return run_serve(args)
    if args.mode == "latency":
        return run_latency(args)
    if args.mode == "probe":
        return run_probe(args)
    if args.mode == "wide":
        return run_wide(args)

    n, d, p, seed = WORKLOADS[args.workload]
    ctx = cuppa.Context(0)
    desc = cuppa.make_desc(n, d, p, seed)
    targets = torch.empty(n * d, dtype=torch.int32, device="cuda")
    row_ptr = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        ctx.generate_into(desc, targets, row_ptr, stream)
    torch.cuda.synchronize()

    kernel_ms, call_ms, stats = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.generate_into(desc, targets, row_ptr, stream)  # waits for its stream
        st = ctx.stats()
        kernel_ms.append(st["kernel_ms"])
        call_ms.append(st["total_ms"])
        stats.append(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    ms_per_step = wall * 1e3 / args.steps
    edges = n * d
    value = edges / (ms_per_step * 1e-3)
    bpe = algorithmic_bytes_per_slot(p, d)
    k_ms = sum(kernel_ms) / len(kernel_ms)
    c_ms = sum(call_ms) / len(call_ms)
    achieved = bpe["generate_kernel"] * edges / (k_ms * 1e-3) / 1e9
    profiles = os.path.join(ROOT, "profiles")
    traffic = read_traffic(profiles, args.workload)
    prof_ms = profile_kernel_ms(profiles, args.workload)
    # the committed counters describe this kernel only if their run's kernel
    # time matches this run's (within 3 %)
    traffic_current = prof_ms is not None and abs(prof_ms - k_ms) <= 0.03 * k_ms
    last = stats[-1]

    extra = {}
    if not args.no_degree:
        extra.update(degree_and_csr(ctx, n, d, edges, targets, row_ptr,
                                    cpu_nodes=args.cpu_prefix))
    if args.ba:
        # same n, d and buffers; the first call is the warm-up
        ba_desc = cuppa.make_desc(n, d, seed=seed, model=cuppa.MODEL_BA)
        ctx.generate_into(ba_desc, targets, row_ptr, stream)
        ba_k, ba_wall = [], []
        for _ in range(max(1, args.steps)):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ctx.generate_into(ba_desc, targets, row_ptr, stream)
            ba_wall.append(time.perf_counter() - t1)
            ba_k.append(ctx.stats()["kernel_ms"])
        bk = sum(ba_k) / len(ba_k)
        bw = sum(ba_wall) / len(ba_wall)
        ba = {"edges_per_s": edges / bw, "ms_per_call": bw * 1e3, "kernel_ms": bk,
              "roofline_frac": algorithmic_bytes_per_slot(0.5, d)["generate_kernel"] * edges / (bk * 1e-3) / 1e9
              / HBM_PEAK_GBS}
        ref_ba = os.path.join(ROOT, "oracle", "_ref", "sba_ref")
        if args.cpu_prefix and os.path.exists(ref_ba):
            with tempfile.TemporaryDirectory() as tmp:
                cmd = [ref_ba, str(args.cpu_prefix), str(d), tmp + "/", "0"]
                cmd = ["taskset", "-c", "0"] + cmd if _has("taskset") else cmd
                out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300).stdout
                secs = float(re.search(r"Computation Time\s*:\s*([0-9.eE+-]+)", out).group(1))
            ba["cpu_baseline"] = {"value": args.cpu_prefix * d / secs, "unit": "edges/s", "cores": 1,
                                  "kind": "reference", "seconds": secs,
                                  "sample": f"reference SerialBA (MT19937-64), first {args.cpu_prefix} nodes, d={d}"}
        extra["ba"] = ba

    del targets, row_ptr
    torch.cuda.empty_cache()
    copy_gbs = copy_bandwidth_gbs(torch)
    design_floors(extra, n, d, edges, copy_gbs)
    requests = read_requests(profiles, args.workload)
    request_roof = None
    if requests and not args.no_gather_ceiling:
        ceil = gather_ceiling(torch)
        rate = requests / (k_ms * 1e-3)
        request_roof = {"read_requests_per_launch": requests, "achieved_requests_per_s": rate,
                        "gather_ceiling": ceil, "frac": rate / ceil["requests_per_s"],
                        "unit": "64-byte read requests/s", "requests_from_committed_profile": True}
        if "probe_lds_nt_with_row_stores" in ceil:
            # against reads that share the memory with row stores, as the generator's do
            request_roof["frac_of_read_write_mix"] = rate / ceil["probe_lds_nt_with_row_stores"]
    roof = philox_roof(torch)
    if roof:
        # the generator's own draws per second against that roof
        roof["generator_draws_per_s"] = last["draws"] / (k_ms * 1e-3)
        roof["frac"] = roof["generator_draws_per_s"] / roof["draws_per_s"]
    cpu = cpu_baseline(n, d, p, args.cpu_prefix) if args.cpu_prefix else None
    if not args.no_batch:
        extra["batch"] = batch_object(args, ctx)
        ctx.set_tuning()
    if not args.no_serve:
        extra["serve"] = serve_object(args, ctx)

    line = {
        "metric": "edges/s at n=2^30 d=8 + HBM-roofline %; batched-vs-plain p50/p99 latency",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (keyed Philox draws)",
        "config": {"workload": f"copy-model single graph n={n} d={d} p={p} seed={seed}",
                   "n": n, "d": d, "p": p, "seed": seed, "edges": edges},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic if traffic_current else None,
                     "traffic_profile": {"file": f"profiles/generate_kernel_{args.workload}_pmc.txt",
                                         "bytes_per_launch": traffic, "kernel_ms": prof_ms,
                                         "matches_this_run": traffic_current},
                     "kernel": f"generate_kernel<{d}>",
                     "algorithmic_bytes_per_edge": bpe["generate_kernel"],
                     "kernel_ms": k_ms},
        "cpu_baseline": cpu,
        # the kernel's 64-byte read requests (committed PMC profile) per second
        # against a plain random gather's on the same GPU
        "request_roofline": request_roof,
        # the integer-VALU roof of the draws (Philox alone on the whole GPU)
        "philox_roof": roof,
        "call_roofline_frac": bpe["call"] * edges / (c_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "call_ms": c_ms,
        "measured_copy_gbs": copy_gbs,
        "stats": {**{k: last[k] for k in ("draws", "retries", "prefetch_misses", "polls", "stalls", "claims",
                                          "sleeps", "staged_rows", "passes", "grid_blocks", "block_threads")},
                  # nodes a wavefront takes on per loop pass (64 lanes)
                  "nodes_per_pass": last["nodes"] / max(1, last["passes"]),
                  # share of slots whose first copy read found the source not yet published
                  "stall_fraction": last["prefetch_misses"] / max(1, last["slots"])},
        **extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
