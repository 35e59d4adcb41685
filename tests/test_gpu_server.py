"""Continuous-batching server: each request's result equals the same graph
generated alone (the oracle's), whatever else is in flight; admission never
exceeds the memory budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix(oracle_mod, count, seed=0, nmax=150_000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.clip(rng.lognormal(np.log(2e4), 1.2), 20, nmax))
        d = int(rng.choice([2, 4, 8, 16]))
        n = max(n, d + 1)
        p = float(rng.choice([0.1, 0.5, 0.9]))
        out.append((i, n, d, oracle_mod.p_threshold(p), i))
    return out


def _expect(oracle_mod, spec):
    _, n, d, T, seed = spec
    G, _, _ = oracle_mod.generate(n, d, T, seed)
    return G


def _drain(srv, specs, check_budget=None):
    import cuppa
    pending = {s[0] for s in specs}
    results = {}
    for _ in range(2_000_000):
        srv.step()
        if check_budget:
            u = srv.usage()
            assert u["used"] <= u["budget"]
        for rid in list(pending):
            r = srv.poll(rid)
            if r.state in (cuppa.REQ_DONE, cuppa.REQ_FAILED):
                assert r.state == cuppa.REQ_DONE, (rid, r.status)
                results[rid] = (cuppa.read_device(r.targets, r.n * r.d, np.int32),
                                cuppa.read_device(r.row_ptr, r.n + 1, np.int64))
                assert srv.release(rid) == 0
                pending.discard(rid)
        if not pending:
            return results
    raise AssertionError("server did not drain")


@pytest.mark.parametrize("plain", [False, True])
def test_mix_matches_alone(ctx, oracle_mod, plain):
    import cuppa
    specs = _mix(oracle_mod, 80 if not plain else 20, seed=1 if not plain else 2)
    srv = cuppa.Server(ctx, budget_bytes=1 << 30, streams=4, batch_nodes=1 << 18, solo_nodes=1 << 17,
                       plain=plain)
    for s in specs:
        assert srv.submit(*s) == 0
    res = _drain(srv, specs)
    srv.close()
    for s in specs:
        got, rp = res[s[0]]
        assert np.array_equal(got, _expect(oracle_mod, s)), s
        assert np.array_equal(rp, np.arange(s[1] + 1, dtype=np.int64) * s[2])


def test_budget_admission(ctx, oracle_mod):
    import cuppa
    specs = [(i, 200_000, 8, oracle_mod.p_threshold(0.5), 100 + i) for i in range(12)]
    one = 200_000 * 8 * 4 + 200_001 * 8
    srv = cuppa.Server(ctx, budget_bytes=3 * one + 4096, streams=4)
    assert srv.submit(99, 10_000_000, 16, 0, 1) == -6  # PPA_ERR_BUDGET: can never fit
    for s in specs:
        assert srv.submit(*s) == 0
    srv.step()
    u = srv.usage()
    assert u["running"] + u["queued"] == len(specs) and u["running"] <= 3
    res = _drain(srv, specs, check_budget=True)
    srv.close()
    for s in specs[:3]:
        assert np.array_equal(res[s[0]][0], _expect(oracle_mod, s))


def test_cancel_and_errors(ctx, oracle_mod):
    import cuppa
    srv = cuppa.Server(ctx, budget_bytes=1 << 26)
    T = oracle_mod.p_threshold(0.5)
    assert srv.submit(1, 1000, 4, T, 1) == 0
    assert srv.submit(1, 1000, 4, T, 1) == -1       # duplicate id
    assert srv.submit(2, 3, 4, T, 1) == -1          # n < d + 1
    assert srv.cancel(1) == 0
    assert srv.poll(1).state == cuppa.REQ_CANCELLED
    assert srv.release(1) == 0
    assert srv.release(1) == -7                     # unknown now
    srv.close()


def test_replay_trace(ctx, oracle_mod):
    import cuppa
    specs = _mix(oracle_mod, 200, seed=7, nmax=50_000)
    rng = np.random.default_rng(3)
    arrivals = np.cumsum(rng.exponential(0.2, len(specs))).tolist()  # ~5000 requests/s
    srv = cuppa.Server(ctx, budget_bytes=1 << 30, streams=4)
    lat, makespan = srv.replay(specs, arrivals)
    srv.close()
    assert all(x > 0 for x in lat)
    assert makespan >= arrivals[-1]
