"""GPU parity of sync points (index_out / index_in of mans_decode and mans_decode_batch)
and of partial decodes (mans_stream_range + mans_decode), through the C ABI.

- The records a decode writes equal the oracle's (mo_decode_seq at the chunk length:
  entry of each chunk's first owned step, symbols before it).
- Decodes that start from those records, from records planted wrong on the host, and
  from the records they overwrite in the same call, are bit-exact.
- Windows decoded alone equal the same slice of the full decode.
Bar: bit-exact.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _truth(st, S):
    from multians_amd import _lib
    _, entries, counts, _, _ = ol.decode_seq(st, S=S)
    rec = np.zeros(len(entries), _lib.SYNC_POINT_DTYPE)
    rec["state"] = entries & 0xFFFF
    rec["bit"] = (entries >> 16) & 63
    rec["symbols_before"] = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]])
    return rec


def _stream(seed, lam, n, L=4096):
    from multians_amd.decoder import encode_symbols
    sym = ol.gen_symbols(seed, lam, n)
    es = encode_symbols(sym, L)
    st = ol.Stream(es.units, es.first_bit, es.first_state, es.table, L, n)
    return sym, es, st


def _host(t):
    from multians_amd import _lib
    return t.cpu().numpy().view(_lib.SYNC_POINT_DTYPE)


@pytest.mark.parametrize("lam,S", [(0.01, 1024), (0.2, 256), (1.0, 64), (3.0, 16)])
def test_single_index_out_in(lam, S):
    import torch
    from multians_amd.decoder import DeviceStream
    print(f"case single index lam={lam} S={S}", flush=True)
    sym, es, st = _stream(81, lam, 2_000_000)
    ds = DeviceStream(es, chunk_units=S, checkpoint_units=16)
    idx = ds.new_index(S)
    out = ds.decode(chunk_units=S, index_out=idx)
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(out[:es.n_symbols].cpu().numpy()[::-1], sym)
    truth = _truth(st, S)
    assert np.array_equal(_host(idx), truth)
    # from the records, writing them again in the same call
    ds.out.zero_()
    out = ds.decode(chunk_units=S, index_in=idx, index_out=idx)
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(out[:es.n_symbols].cpu().numpy()[::-1], sym)
    assert np.array_equal(_host(idx), truth)


@pytest.mark.parametrize("how", ["random", "shifted", "out_of_range"])
def test_single_wrong_records(how):
    """Wrong records (planted on the host, masked into range by the decoder) cost walks
    and fixup rounds, never correctness."""
    import torch
    from multians_amd.decoder import DeviceStream
    print(f"case wrong records {how}", flush=True)
    sym, es, st = _stream(82, 0.1, 1_500_000)
    S = 256
    truth = _truth(st, S)
    bad = truth.copy()
    rng = np.random.default_rng(3)
    if how == "random":
        bad["state"] = rng.integers(0, es.L, len(bad))
        bad["bit"] = rng.integers(0, 32, len(bad))
    elif how == "shifted":
        bad[1:] = truth[:-1]
    else:
        bad["state"] = 0xFFFFFFFF
        bad["bit"] = 0xFFFFFFFF
    ds = DeviceStream(es, chunk_units=S, checkpoint_units=16)
    idx_in = torch.from_numpy(bad.view(np.uint8).copy()).cuda()
    idx_out = ds.new_index(S)
    out = ds.decode(chunk_units=S, index_in=idx_in, index_out=idx_out)
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(out[:es.n_symbols].cpu().numpy()[::-1], sym)
    assert np.array_equal(_host(idx_out), truth)


def test_single_rejects_bad_chunk_units():
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream
    print("case index chunk_units", flush=True)
    _, es, _ = _stream(83, 0.5, 10_000)
    ds = DeviceStream(es, chunk_units=256, checkpoint_units=16)
    idx = ds.new_index(256)
    for cu in (0, 8, 100):
        with pytest.raises(_lib.MansError):
            ds.decode(chunk_units=cu, index_out=idx)


@pytest.mark.parametrize("n_streams", [300, 700])
def test_batch_index(n_streams):
    """A batch (700 streams: split into concurrent parts) writes every stream's records;
    decodes from them, and from them while writing them again, are bit-exact."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceBatch, EncodedStream
    print(f"case batch index n={n_streams}", flush=True)
    rng = np.random.default_rng(n_streams)
    items = []
    for i in range(n_streams):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(400_000))))
        lam = float(np.exp(rng.uniform(np.log(0.005), np.log(2.0))))
        units, fb, fs, dec, sym = _lib.make_stream(500 + i, lam, n, 4096, True)
        items.append((EncodedStream(units, fb, fs, dec, 4096, n), sym))
    b = DeviceBatch([e for e, _ in items], sync_points=True)

    def check():
        torch.cuda.synchronize()
        assert b.status() == 0
        for i, (_, sym) in enumerate(items):
            assert np.array_equal(b.output(i).cpu().numpy()[::-1], sym), i

    b.decode(index="out")
    check()
    recs = [b.sync_points(i).copy() for i in range(n_streams)]
    for i in range(0, n_streams, 37):            # against the oracle
        es = items[i][0]
        st = ol.Stream(es.units, es.first_bit, es.first_state, es.table, 4096, es.n_symbols)
        assert np.array_equal(recs[i], _truth(st, int(b.chunk_units[i]))), i
    b.out.zero_()
    b.decode(index="in")
    check()
    b.out.zero_()
    b.decode(index="both")
    check()
    for i in range(n_streams):
        assert np.array_equal(b.sync_points(i), recs[i]), i


@pytest.mark.parametrize("lam,S", [(0.02, 1024), (0.5, 256), (2.5, 64)])
def test_range_decode(lam, S):
    """Windows decoded alone (mans_stream_range + mans_decode from the first chunk's
    record) equal the full decode's slice: both ends, one symbol, chunk boundaries,
    random windows, the whole stream."""
    import torch
    from multians_amd.decoder import DeviceStream
    print(f"case range lam={lam} S={S}", flush=True)
    sym, es, st = _stream(84, lam, 3_000_000)
    ds = DeviceStream(es, chunk_units=S, checkpoint_units=16)
    idx = ds.new_index(S)
    ds.decode(chunk_units=S, index_out=idx)
    torch.cuda.synchronize()
    full = sym[::-1]
    host = _host(idx)
    n = es.n_symbols
    b = host["symbols_before"]
    k = len(b) // 3
    wins = [(0, 1), (0, 4096), (n - 1, n), (n - 100_000, n), (0, n),
            (int(b[k]), int(b[k]) + 1), (int(b[k]) - 1, int(b[k]) + 1), (int(b[k]), int(b[k + 5]))]
    rng = np.random.default_rng(11)
    for _ in range(20):
        a = int(rng.integers(0, n - 1))
        wins.append((a, int(rng.integers(a + 1, min(n, a + 200_000) + 1))))
    for first, last in wins:
        got = ds.decode_range(first, last, idx, S, index_host=host)
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), full[first:last]), (first, last)


def test_graph_from_sync_points():
    """A captured graph of a batch (600 streams: concurrent parts) that decodes from its
    sync points and writes them again replays bit-exact, with the program's own eager
    decodes between the replays."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceBatch, EncodedStream
    print("case graph sync points", flush=True)
    rng = np.random.default_rng(44)
    items = []
    for i in range(600):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(200_000))))
        units, fb, fs, dec, sym = _lib.make_stream(7000 + i, float(rng.uniform(0.01, 2.0)), n,
                                                   4096, True)
        items.append((EncodedStream(units, fb, fs, dec, 4096, n), sym))
    b = DeviceBatch([e for e, _ in items], sync_points=True)
    b.decode(index="out")
    torch.cuda.synchronize()
    rec = b.index.clone()
    g = b.graph(index="both")
    for rep in range(3):
        b.out.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert b.status() == 0
        assert torch.equal(b.index, rec), rep
        for i in range(0, 600, 7):
            assert np.array_equal(b.output(i).cpu().numpy()[::-1], items[i][1]), (rep, i)
        b.decode()                                  # eager work between the replays
        torch.cuda.synchronize()
    g.close()


@pytest.mark.parametrize("lam,S", [(0.01, 1024), (0.3, 256), (2.0, 64), (3.0, 16)])
def test_items_out_and_trusted(lam, S):
    """A decode at chunk length S writes a record per 32 units (items_out): they equal
    the oracle's at subsequence length 32.  The write pass alone from them
    (MANS_F_TRUST_SYNC_POINTS) is bit-exact."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream
    print(f"case items lam={lam} S={S}", flush=True)
    sym, es, st = _stream(85, lam, 2_500_003)
    ds = DeviceStream(es, chunk_units=S, checkpoint_units=16)
    items = ds.new_index(32)
    ds.decode(chunk_units=S, items_out=items)
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(_host(items), _truth(st, 32))
    ds.out.zero_()
    ds2 = DeviceStream(es, chunk_units=32, checkpoint_units=16)
    out = ds2.decode(chunk_units=32, index_in=items, flags=_lib.MANS_F_TRUST_SYNC_POINTS)
    torch.cuda.synchronize()
    assert ds2.status() == 0
    assert np.array_equal(out[:es.n_symbols].cpu().numpy()[::-1], sym)


def test_trusted_edge_streams():
    """One-symbol (zero-bit) stream, tiny streams, a stream shorter than one item."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, encode_symbols
    print("case trusted edges", flush=True)
    cases = [np.full(100_000, 7, np.uint8), ol.gen_symbols(86, 0.5, 1),
             ol.gen_symbols(87, 0.5, 37), ol.gen_symbols(88, 0.5, 5000)]
    for sym in cases:
        es = encode_symbols(sym, 1024)
        st = ol.Stream(es.units, es.first_bit, es.first_state, es.table, 1024, len(sym))
        ds = DeviceStream(es, chunk_units=32, checkpoint_units=16)
        idx = torch.from_numpy(_truth(st, 32).view(np.uint8).copy()).cuda()
        out = ds.decode(chunk_units=32, index_in=idx, flags=_lib.MANS_F_TRUST_SYNC_POINTS)
        torch.cuda.synchronize()
        assert ds.status() == 0
        assert np.array_equal(out[:len(sym)].cpu().numpy()[::-1], sym), len(sym)
    with pytest.raises(_lib.MansError):                 # trusted records must be per 32 units
        ds = DeviceStream(es, chunk_units=64, checkpoint_units=16)
        ds.decode(chunk_units=64, index_in=idx, flags=_lib.MANS_F_TRUST_SYNC_POINTS)


def test_trusted_wrong_records_stay_in_bounds():
    """Unproven records planted wrong (random states, bits and offsets): the decode ends,
    and nothing is written past n_out (guard bytes after the output)."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream
    print("case trusted wrong records", flush=True)
    sym, es, st = _stream(89, 0.2, 1_000_000)
    bad = _truth(st, 32)
    rng = np.random.default_rng(6)
    bad["state"] = rng.integers(0, 1 << 32, len(bad), dtype=np.uint64).astype(np.uint32)
    bad["bit"] = rng.integers(0, 1 << 32, len(bad), dtype=np.uint64).astype(np.uint32)
    bad["symbols_before"][1:] = np.sort(rng.integers(0, 3 * es.n_symbols, len(bad) - 1))
    ds = DeviceStream(es, chunk_units=32, checkpoint_units=16)
    n = es.n_symbols
    ds.out = torch.full((n + 65536,), 0xAB, dtype=torch.uint8, device="cuda")
    idx = torch.from_numpy(bad.view(np.uint8).copy()).cuda()
    ds.decode(chunk_units=32, index_in=idx, flags=_lib.MANS_F_TRUST_SYNC_POINTS)
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert bool((ds.out[n:] == 0xAB).all())


def test_trusted_batch_and_windows():
    """A 700-stream batch (concurrent parts) writes its item records in a decode at the
    tuned chunk lengths; the write pass alone from them, and 2,000 windows from them,
    are bit-exact."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceBatch, EncodedStream, WindowDecoder
    print("case trusted batch", flush=True)
    rng = np.random.default_rng(91)
    items = []
    for i in range(700):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(400_000))))
        units, fb, fs, dec, sym = _lib.make_stream(9100 + i, float(rng.uniform(0.01, 2.5)), n,
                                                   4096, True)
        items.append((EncodedStream(units, fb, fs, dec, 4096, n), sym))
    b = DeviceBatch([e for e, _ in items])
    b.use_items()
    b.decode(index="items_out")
    torch.cuda.synchronize()
    b.out.zero_()
    b.decode(index="trusted", trust=True)
    torch.cuda.synchronize()
    assert b.status() == 0
    for i, (_, sym) in enumerate(items):
        assert np.array_equal(b.output(i).cpu().numpy()[::-1], sym), i
    wins = []
    for _ in range(2000):
        i = int(rng.integers(0, 700))
        nn = items[i][0].n_symbols
        a = int(rng.integers(0, nn))
        wins.append((i, a, int(min(nn, a + 1 + rng.integers(0, 50_000)))))
    wd = WindowDecoder()
    arena, spans = wd.decode(b.plan_windows(wins, items=True), wins, trust=True)
    torch.cuda.synchronize()
    assert wd.status() == 0
    host = arena.cpu().numpy()
    for (i, a, z), (o, ln) in zip(wins, spans):
        assert np.array_equal(host[o:o + ln], items[i][1][::-1][a:z]), (i, a, z)
