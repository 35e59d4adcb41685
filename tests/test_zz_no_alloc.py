"""ppa_generate_noalloc (the entry point that allocates nothing: its launch
record lives in a static device buffer) against ppa_generate through the C
ABI, bit for bit, on two calls in a row, the second with a smaller graph, so
that the second call must not see the first call's record."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(desc, targets, row_ptr):
    import cuppa
    st = cuppa.Stats()
    rc = cuppa.lib().ppa_generate_noalloc(ctypes.byref(desc), ctypes.c_void_p(targets.data_ptr()),
                                          ctypes.c_void_p(row_ptr.data_ptr()), ctypes.byref(st),
                                          cuppa._stream_handle(None))
    return rc, st


def test_noalloc_matches_generate_twice(ctx, oracle_mod):
    import torch
    import cuppa
    for n, d, p, seed in ((400_000, 8, 0.2, 71), (60_000, 16, 0.1, 72)):
        desc = cuppa.make_desc(n, d, p, seed)
        want = torch.full((n * d,), -7, dtype=torch.int32, device="cuda")
        want_rp = torch.full((n + 1,), -7, dtype=torch.int64, device="cuda")
        ctx.generate_into(desc, want, want_rp)
        got = torch.full((n * d,), -9, dtype=torch.int32, device="cuda")
        got_rp = torch.full((n + 1,), -9, dtype=torch.int64, device="cuda")
        rc, st = _noalloc(desc, got, got_rp)
        assert rc == 0, cuppa.strerror(rc)
        assert torch.equal(got, want), (n, d)
        assert torch.equal(got_rp, want_rp), (n, d)
        assert st.nodes == n - d - 1 and st.slots == (n - d - 1) * d
        assert st.retries == ctx.stats()["retries"]
    # the smaller, second graph is also the oracle's
    T = oracle_mod.p_threshold(0.1)
    ref, _, _ = oracle_mod.generate(60_000, 16, T, 72)
    assert np.array_equal(got.cpu().numpy(), ref)


def test_noalloc_wavefront_d_twice(ctx, oracle_mod):
    """Out-degrees above 32 (one wavefront per node) through the same record."""
    import torch
    import cuppa
    for n, d, p, seed in ((20_000, 100, 0.3, 73), (3_000, 40, 0.5, 74)):
        desc = cuppa.make_desc(n, d, p, seed)
        want = torch.full((n * d,), -7, dtype=torch.int32, device="cuda")
        ctx.generate_into(desc, want, None)
        got = torch.full((n * d,), -9, dtype=torch.int32, device="cuda")
        got_rp = torch.full((n + 1,), -9, dtype=torch.int64, device="cuda")
        rc, st = _noalloc(desc, got, got_rp)
        assert rc == 0, cuppa.strerror(rc)
        assert torch.equal(got, want), (n, d)
        assert st.nodes == n - d - 1 and st.slots == (n - d - 1) * d
        assert st.retries == ctx.stats()["retries"]
    ref, _, _ = oracle_mod.generate(3_000, 40, oracle_mod.p_threshold(0.5), 74)
    assert np.array_equal(got.cpu().numpy(), ref)


def test_noalloc_rejects_bad_descriptors(ctx):
    import torch
    import cuppa
    t = torch.empty(1024, dtype=torch.int32, device="cuda")
    rp = torch.empty(65, dtype=torch.int64, device="cuda")
    for desc in (cuppa.make_desc(9000, 8193, 0.5, 1), cuppa.make_desc(8, 8, 0.5, 1)):
        rc, _ = _noalloc(desc, t, rp)
        assert rc == -1  # PPA_ERR_INVALID
    rc = cuppa.lib().ppa_generate_noalloc(ctypes.byref(cuppa.make_desc(64, 8, 0.5, 1)), None, None, None,
                                          cuppa._stream_handle(None))
    assert rc == -1
