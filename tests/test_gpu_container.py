"""GPU parity of the batch container (include/mans_container.h): a packed image copied
to the device once decodes every stream with one mans_decode_batch call (with and
without its stored sync points), and batches of windows decode alone from the stored
sync points.  Bar: bit-exact against the original symbols.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make(n_streams, seed):
    """Streams with sync points recorded by a decode (index_out), packed."""
    import torch
    from multians_amd import _lib, container
    from multians_amd.decoder import DeviceBatch, EncodedStream
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n_streams):
        n = int(np.exp(rng.uniform(np.log(500), np.log(600_000))))
        lam = float(np.exp(rng.uniform(np.log(0.005), np.log(2.5))))
        L = 1024 if i % 4 == 0 else 4096
        units, fb, fs, dec, sym = _lib.make_stream(seed * 1000 + i, lam, n, L, True)
        items.append((EncodedStream(units, fb, fs, dec, L, n), sym))
    b = DeviceBatch([e for e, _ in items], sync_points=True)
    b.decode(index="out")
    torch.cuda.synchronize()
    assert b.status() == 0
    sps = [b.sync_points(i).copy() if i % 5 else None for i in range(n_streams)]
    img = container.pack([e for e, _ in items], sps, [int(c) for c in b.chunk_units])
    return items, sps, container.Container(img)


@pytest.mark.parametrize("n_streams", [250, 640])
def test_container_decode(n_streams):
    import torch
    from multians_amd import container
    print(f"case container n={n_streams}", flush=True)
    items, sps, c = _make(n_streams, 7 + n_streams)
    assert c.info.n_with_sync_points == sum(s is not None for s in sps)
    dc = container.DeviceContainer(c)
    for use in (True, False):
        dc.out.zero_()
        dc.decode(use_sync_points=use)
        torch.cuda.synchronize()
        assert dc.status(("all", use)) == 0
        for i, (_, sym) in enumerate(items):
            assert np.array_equal(dc.output(i).cpu().numpy()[::-1], sym), (use, i)


def test_container_windows():
    """1,500 windows of the streams with sync points, in one batch call, twice (the
    second call with fewer, shorter windows reuses the arena)."""
    import torch
    from multians_amd import container
    print("case container windows", flush=True)
    items, sps, c = _make(120, 3)
    dc = container.DeviceContainer(c)
    have = [i for i, s in enumerate(sps) if s is not None]
    rng = np.random.default_rng(17)
    for count, span in ((1500, 200_000), (300, 2_000)):
        wins = []
        for _ in range(count):
            i = int(rng.choice(have))
            n = items[i][0].n_symbols
            a = int(rng.integers(0, n))
            wins.append((i, a, int(min(n, a + 1 + rng.integers(0, span)))))
        wins += [(have[0], 0, 1), (have[1], items[have[1]][0].n_symbols - 1,
                                   items[have[1]][0].n_symbols)]
        arena, spans = dc.decode_windows(wins)
        torch.cuda.synchronize()
        assert dc._win.status() == 0
        host = arena.cpu().numpy()
        for (i, a, z), (o, ln) in zip(wins, spans):
            assert ln == z - a
            assert np.array_equal(host[o:o + ln], items[i][1][::-1][a:z]), (i, a, z)


def test_container_item_records_trusted():
    """Item records (one per 32 units) written by a decode, stored in a container, decode
    the container's streams by the write pass alone (MANS_BATCH_TRUST_SYNC_POINTS)."""
    import torch
    from multians_amd import _lib, container
    from multians_amd.decoder import DeviceBatch, EncodedStream
    print("case container item records", flush=True)
    rng = np.random.default_rng(12)
    items = []
    for i in range(300):
        n = int(np.exp(rng.uniform(np.log(100), np.log(500_000))))
        units, fb, fs, dec, sym = _lib.make_stream(12000 + i, float(rng.uniform(0.01, 2.5)), n,
                                                   4096 if i % 3 else 1024, True)
        items.append((EncodedStream(units, fb, fs, dec, 4096 if i % 3 else 1024, n), sym))
    b = DeviceBatch([e for e, _ in items])
    b.use_items()
    b.decode(index="items_out")
    torch.cuda.synchronize()
    host = b.items.cpu().numpy().view(_lib.SYNC_POINT_DTYPE)
    recs = [host[b.item_off[i]:b.item_off[i + 1]].copy() for i in range(len(items))]
    img = container.pack([e for e, _ in items], recs, [32] * len(items))
    dc = container.DeviceContainer(container.Container(img))
    dc.decode(trust=True)
    torch.cuda.synchronize()
    assert dc.status(("all", True)) == 0
    for i, (_, sym) in enumerate(items):
        assert np.array_equal(dc.output(i).cpu().numpy()[::-1], sym), i
