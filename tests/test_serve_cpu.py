"""Host logic of the serving loop (no GPU): the device arena's allocator, the request
mix and the arrival process."""
import numpy as np
import pytest

from multians_amd import serve


def test_arena_first_fit_and_coalescing():
    a = serve.DeviceArena(1 << 20, device="cpu")
    blocks = [a.alloc(n) for n in (1000, 5000, 256, 70000)]
    assert all(b is not None for b in blocks)
    offs = [b[0] for b in blocks]
    assert offs == sorted(offs) and all(o % 256 == 0 for o in offs)
    # no overlap
    for (o1, n1), (o2, n2) in zip(blocks, blocks[1:]):
        assert o1 + n1 <= o2
    a.free(*blocks[1])
    a.free(*blocks[0])                       # merges with the following free block
    b = a.alloc(1024 + 5120)
    assert b[0] == 0
    for blk in (b, blocks[2], blocks[3]):
        a.free(*blk)
    assert a.used == 0 and a.free_off == [0] and a.free_len[0] == a.size
    assert a.alloc(a.size + 1) is None
    assert a.peak >= 1024 + 5120 + 256 + 70144


def test_arena_random_alloc_free_keeps_invariants():
    rng = np.random.default_rng(3)
    a = serve.DeviceArena(1 << 22, device="cpu")
    live = []
    for _ in range(2000):
        if live and rng.random() < 0.45:
            a.free(*live.pop(int(rng.integers(len(live)))))
        else:
            b = a.alloc(int(rng.integers(1, 100000)))
            if b is not None:
                live.append(b)
        spans = sorted(live) + sorted((o, a.free_len[o]) for o in a.free_off)
        spans.sort()
        pos = 0
        for o, n in spans:                   # live and free blocks tile the arena
            assert o == pos
            pos += n
        assert pos == a.size
        assert a.used == sum(n for _, n in live)


def test_request_mix_and_arrivals():
    mix = serve.request_mix(4000)
    sizes = np.array([m[2] for m in mix])
    lams = np.array([m[1] for m in mix])
    Ls = np.array([m[3] for m in mix])
    assert sizes.min() >= 1 << 10 and sizes.max() <= 64 << 20
    assert 0.005 <= lams.min() and lams.max() <= 2.0
    assert set(Ls) == {1024, 4096} and abs((Ls == 1024).mean() - 0.5) < 0.05
    # log-uniform: the median size is near the geometric mean of the range
    assert abs(np.log(np.median(sizes)) - np.log(np.sqrt((1 << 10) * (64 << 20)))) < 0.5
    t = serve.poisson_arrivals(1000.0, 5.0)
    assert np.all(np.diff(t) > 0) and t[-1] < 5.0
    assert abs(len(t) - 5000) < 5 * np.sqrt(5000)
