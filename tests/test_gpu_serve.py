"""Serving loop on the GPU: requests arriving while others run are admitted, batched
and decoded; every result equals the original symbols, which is also what each stream
decodes to alone."""
import numpy as np
import pytest


@pytest.mark.gpu
def test_served_results_match_alone_and_mix():
    import torch
    from multians_amd import serve
    from multians_amd.decoder import DeviceStream
    import bench

    specs = serve.request_mix(40, seed=5, max_bytes=1 << 20)
    items = bench.make_streams(specs)
    pool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items])
    srv = serve.Server(pool, 1 << 30, max_inflight=3)
    arr = serve.poisson_arrivals(3000.0, 0.4, seed=3)
    ids = np.random.default_rng(4).integers(0, pool.n, len(arr))
    print("serve: %d requests over %d streams" % (len(arr), pool.n), flush=True)
    st = srv.run(arr, ids, check_every=1)
    s = st.summary()
    print(s, flush=True)
    assert s["requests"] == len(arr)
    assert st.checked == len(arr) and st.mismatches == 0
    assert s["batches"] >= 2                      # arrivals spread over several launches
    assert srv.arena.used == 0
    # the same streams decoded alone
    for i in range(0, pool.n, 7):
        es, sym = items[i]
        got = DeviceStream(es).decode()[:es.n_symbols]
        assert torch.equal(got, pool.expected(i))
        assert np.array_equal(got.cpu().numpy()[::-1], sym)


@pytest.mark.gpu
def test_served_budget_smaller_than_backlog():
    """A budget that holds only a few requests at once still finishes every request."""
    from multians_amd import serve
    import bench

    specs = serve.request_mix(12, seed=9, min_bytes=200 << 10, max_bytes=1 << 20)
    items = bench.make_streams(specs)
    pool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items])
    srv = serve.Server(pool, 3 << 20, max_inflight=2)
    st = srv.run(np.zeros(60), np.arange(60) % pool.n, check_every=1)
    assert st.summary()["requests"] == 60 and st.mismatches == 0
    assert srv.arena.peak <= 3 << 20 and st.n_batches >= 10


@pytest.mark.gpu
def test_served_from_item_records():
    """use_records: a stream's first decode writes its item records, later requests for
    it decode from them (trusted write pass), in the same batches as first decodes.
    Every result is checked; the records equal a fresh decode's."""
    import torch
    from multians_amd import _lib, serve
    from multians_amd.decoder import DeviceStream
    import bench

    specs = serve.request_mix(30, seed=6, max_bytes=2 << 20)
    items = bench.make_streams(specs)
    pool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items], records=True)
    srv = serve.Server(pool, 1 << 30, max_inflight=3, use_records=True)
    arr = serve.poisson_arrivals(4000.0, 0.5, seed=8)
    ids = np.random.default_rng(9).integers(0, pool.n, len(arr))
    print("serve records: %d requests over %d streams" % (len(arr), pool.n), flush=True)
    st = srv.run(arr, ids, check_every=1)
    assert st.checked == len(arr) and st.mismatches == 0
    assert pool.have_rec.sum() == len(set(ids.tolist()))
    assert srv.arena.used == 0
    # the records a stream holds are those a decode of it alone writes
    for i in sorted(set(ids.tolist()))[::5]:
        es, _ = items[i]
        ds = DeviceStream(es)
        rec = ds.new_index(32)
        ds.decode(items_out=rec)
        torch.cuda.synchronize()
        n = _lib.sync_point_count(es.n_units, 32)
        a = int((pool.rptr[i] - pool.rec.data_ptr()) // 16)
        assert torch.equal(pool.rec[16 * a:16 * (a + n)], rec[:16 * n])


@pytest.mark.gpu
def test_served_failed_stream_is_reported():
    """A request whose stream has a rejected table (an entry with 16 bits) is reported as
    failed from its own status word, and never gets item records; the requests batched
    with it still succeed and are checked."""
    import torch
    from multians_amd import serve
    import bench

    specs = serve.request_mix(20, seed=12, max_bytes=1 << 20)
    items = bench.make_streams(specs)
    pool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items], records=True)
    bad = 5
    t0 = int((pool.tptr[bad] - pool.tables.data_ptr()) // 4)
    pool.tables[t0 + 3] = (16 << 24) | (int(pool.tables[t0 + 3]) & 0xFFFFFF)
    torch.cuda.synchronize()
    srv = serve.Server(pool, 1 << 30, max_inflight=2, use_records=True)
    ids = np.arange(60) % pool.n
    st = srv.run(np.zeros(len(ids)), ids, check_every=1)
    s = st.summary()
    print(s, flush=True)
    n_bad = int(np.count_nonzero(ids == bad))
    assert s["failed"] == n_bad
    assert st.checked == len(ids) - n_bad and st.mismatches == 0
    assert not pool.have_rec[bad]
    assert pool.have_rec.sum() == pool.n - 1


@pytest.mark.gpu
def test_served_host_inputs_match_alone():
    """Requests that bring their compressed stream and table from (pinned) host memory:
    copied to the device on admission, decoded in the mix, every result equal to the
    symbols and to the stream decoded alone; nothing stays on the device afterwards."""
    import torch
    from multians_amd import serve
    from multians_amd.decoder import DeviceStream
    import bench

    specs = serve.request_mix(40, seed=15, max_bytes=1 << 20)
    items = bench.make_streams(specs)
    pool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items], host_inputs=True)
    srv = serve.Server(pool, 1 << 30, max_inflight=3)
    arr = serve.poisson_arrivals(2000.0, 0.4, seed=13)
    ids = np.random.default_rng(14).integers(0, pool.n, len(arr))
    print("serve host inputs: %d requests" % len(arr), flush=True)
    st = srv.run(arr, ids, check_every=1)
    s = st.summary()
    print(s, flush=True)
    assert s["requests"] == len(arr) and s["failed"] == 0
    assert st.checked == len(arr) and st.mismatches == 0
    assert srv.arena.used == 0
    for i in range(0, pool.n, 9):
        es, sym = items[i]
        got = DeviceStream(es).decode()[:es.n_symbols]
        assert torch.equal(got, pool.expected(i))


@pytest.mark.gpu
def test_served_shape_budget_binds():
    """The served configuration's request shape -- 1024 or 4096 states, uint8 sizes
    log-uniform up to 64 MiB, lambda ~ LogU(0.005, 2.0), inputs from the host -- under a
    budget that holds only a few requests at once: the budget binds (several batches,
    peak use within it) and every request decodes to its symbols."""
    from multians_amd import serve
    import bench

    specs = serve.request_mix(16, seed=21, min_bytes=1 << 20, max_bytes=64 << 20)
    items = bench.make_streams(specs)
    pool = serve.StreamPool([es for es, _ in items], [sym for _, sym in items], host_inputs=True)
    budget = 400 << 20
    srv = serve.Server(pool, budget, max_inflight=3)
    ids = np.arange(48) % pool.n
    print("serve shape: %d requests, %.0f MB of symbols" %
          (len(ids), pool.n_sym[ids].sum() / 1e6), flush=True)
    st = srv.run(np.zeros(len(ids)), ids, check_every=1)
    s = st.summary()
    print(s, flush=True)
    assert s["requests"] == len(ids) and s["failed"] == 0 and st.mismatches == 0
    assert st.checked == len(ids)
    assert srv.arena.peak <= budget and st.n_batches >= 6
    assert srv.arena.used == 0
