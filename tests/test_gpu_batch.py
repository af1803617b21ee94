"""GPU parity of the batched decoder: every stream of a mixed batch equals its
oracle decode and its original symbols, and the same as when decoded alone."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _mixed(nstreams, seed, max_log2=20):
    from multians_amd.decoder import encode_symbols
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nstreams):
        n = int(np.exp(rng.uniform(np.log(1), np.log(1 << max_log2))))
        lam = float(np.exp(rng.uniform(np.log(0.005), np.log(3.0))))
        L = int(rng.choice([1024, 4096]))
        sym = ol.gen_symbols(seed * 1000 + i, lam, n)
        out.append((encode_symbols(sym, L), sym))
    return out


@pytest.mark.parametrize("nstreams,seed", [(1, 1), (7, 2), (300, 3)])
def test_batch_matches_oracle_and_solo(nstreams, seed):
    import torch
    from multians_amd.decoder import DeviceBatch, DeviceStream
    print(f"case batch n={nstreams} seed={seed}", flush=True)
    items = _mixed(nstreams, seed)
    b = DeviceBatch([es for es, _ in items])
    stats = b.decode()
    torch.cuda.synchronize()
    assert b.status() == 0
    assert stats.n_streams == nstreams
    for i, (es, sym) in enumerate(items):
        got = b.output(i).cpu().numpy()
        assert np.array_equal(got[::-1], sym), "stream %d" % i
        if i % 50 == 0:
            ref, *_ = ol.decode_seq(ol.Stream(es.units, es.first_bit, es.first_state, es.table,
                                              es.L, es.n_symbols))
            assert np.array_equal(got, ref)
            solo = DeviceStream(es).decode()
            torch.cuda.synchronize()
            assert np.array_equal(solo[:es.n_symbols].cpu().numpy(), got)


def test_batch_repeat_is_stable():
    """Decoding the same batch twice (workspace reuse) gives identical results."""
    import torch
    from multians_amd.decoder import DeviceBatch
    items = _mixed(40, 9, max_log2=22)
    b = DeviceBatch([es for es, _ in items])
    b.decode()
    first = b.out.clone()
    b.decode()
    torch.cuda.synchronize()
    assert torch.equal(first, b.out)
    for i, (es, sym) in enumerate(items):
        assert np.array_equal(b.output(i).cpu().numpy()[::-1], sym)


def test_batch_graph_replays_match_eager():
    """A captured batch decode, replayed several times with eager work in between
    (output cleared, another decode on the same device), gives the eager result on
    every replay."""
    import torch
    from multians_amd.decoder import DeviceBatch, DeviceStream
    print("case graph replays", flush=True)
    items = _mixed(60, 21, max_log2=21)
    b = DeviceBatch([es for es, _ in items])
    b.decode()
    torch.cuda.synchronize()
    eager = b.out.clone()
    g = b.graph()
    other = _mixed(3, 22)
    for rep in range(4):
        b.out.zero_()
        ds = DeviceStream(other[rep % 3][0])      # eager work between replays
        ds.decode()
        g.launch()
        torch.cuda.synchronize()
        assert b.status() == 0
        assert torch.equal(b.out, eager), "replay %d" % rep
    for i, (es, sym) in enumerate(items):
        assert np.array_equal(b.output(i).cpu().numpy()[::-1], sym)
    g.close()


def test_large_batch_runs_as_parts():
    """A batch of 800 streams decodes as concurrent parts (one pipeline per share of
    the streams); every stream still equals its symbols, and a timed decode (one
    pipeline, kernel events) gives the same bytes."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceBatch
    print("case parts n=800", flush=True)
    items = _mixed(800, 31, max_log2=17)
    b = DeviceBatch([es for es, _ in items])
    b.decode()
    torch.cuda.synchronize()
    assert b.status() == 0
    split = b.out.clone()
    for i, (es, sym) in enumerate(items):
        assert np.array_equal(b.output(i).cpu().numpy()[::-1], sym), "stream %d" % i
    b.out.zero_()
    kt = _lib.KernelTimer()
    b.decode(kernel_events=kt.ptr)
    torch.cuda.synchronize()
    kt.close()
    assert b.status() == 0
    assert torch.equal(b.out, split)


def test_parts_report_table_errors():
    """A rejected table in any part of a split batch reaches mans_decode_status: the
    second and third largest streams go to the second and third parts."""
    import torch
    from multians_amd.decoder import DeviceBatch
    print("case parts status", flush=True)
    items = _mixed(800, 32, max_log2=16)
    order = np.argsort([-es.n_symbols for es, _ in items], kind="stable")
    streams = [es for es, _ in items]
    for j in (order[1], order[2]):
        es = streams[j]
        bad = es.table.copy()
        bad[0] = (bad[0] & 0x00FFFFFF) | (20 << 24)       # nbits 20: rejected on device
        streams[j] = type(es)(es.units, es.first_bit, es.first_state, bad, es.L, es.n_symbols,
                              es.counts)
    b = DeviceBatch(streams)
    b.decode()
    torch.cuda.synchronize()
    assert b.status() != 0
    good = [i for i in range(len(items)) if i not in (order[1], order[2])]
    for i in good[::40]:
        assert np.array_equal(b.output(i).cpu().numpy()[::-1], items[i][1]), "stream %d" % i


def test_parts_graph_replays_match_eager():
    """The graph of a split batch (parts forked onto private streams inside the
    capture) replays to the eager result."""
    import torch
    from multians_amd.decoder import DeviceBatch
    print("case parts graph", flush=True)
    items = _mixed(600, 33, max_log2=15)
    b = DeviceBatch([es for es, _ in items])
    b.decode()
    torch.cuda.synchronize()
    eager = b.out.clone()
    g = b.graph()
    for rep in range(3):
        b.out.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert b.status() == 0
        assert torch.equal(b.out, eager), "replay %d" % rep
    g.close()


def test_batch_with_small_stream_part():
    """A split batch with at least 256 streams of at most 64 sync chunks decodes those in
    their own part, whose sync kernel runs 64-lane workgroups; every stream bit-exact."""
    import torch
    from multians_amd.decoder import DeviceBatch, encode_symbols
    print("case batch small-stream part: 260 small + 260 larger streams", flush=True)
    rng = np.random.default_rng(11)
    items = []
    for i in range(520):
        if i % 2 == 0:     # small: a few thousand units, a handful of chunks
            n = int(rng.integers(4096, 16384))
            lam = float(np.exp(rng.uniform(np.log(0.005), np.log(2.0))))
        else:              # larger: 256 Ki symbols of high entropy, 256 chunks or more
            n = 1 << 18
            lam = float(rng.uniform(0.005, 0.05))
        sym = ol.gen_symbols(11_000 + i, lam, n)
        items.append((encode_symbols(sym, 4096), sym))
    b = DeviceBatch([es for es, _ in items])
    b.decode()
    torch.cuda.synchronize()
    assert b.status() == 0
    for i, (es, sym) in enumerate(items):
        assert np.array_equal(b.output(i).cpu().numpy()[::-1], sym), "stream %d" % i
    # the same streams as one pipeline (256-lane sync workgroups throughout)
    first = b.out.clone()
    b.out.zero_()
    b.decode(one_pipeline=True)
    torch.cuda.synchronize()
    assert b.status() == 0
    assert torch.equal(first, b.out)
