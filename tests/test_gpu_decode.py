"""GPU parity: the HIP decoder (through the C ABI) against the oracle.

Bar: bit-exact.  Each case decodes a seeded stream and compares with the oracle's
sequential decoder and with the original symbols (reversed: the decoder emits in
decode order, src/main.cc:123).
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _case(lam, L, n, seed=11, **kw):
    import torch
    from multians_amd.decoder import DeviceStream, encode_symbols
    sym = ol.gen_symbols(seed, lam, n)
    es = encode_symbols(sym, L)
    ds = DeviceStream(es, chunk_units=kw.get("chunk_units", 0),
                      checkpoint_units=kw.get("checkpoint_units", 0))
    out = ds.decode(**kw)
    torch.cuda.synchronize()
    got = out[:n].cpu().numpy()
    st = ol.Stream(es.units, es.first_bit, es.first_state, es.table, L, n)
    ref, _, _, _, _ = ol.decode_seq(st)
    assert ds.status() == 0
    return got, ref, sym


@pytest.mark.parametrize("lam", [0.002, 0.01, 0.1, 1.0, 3.0])
@pytest.mark.parametrize("L", [1024, 4096])
def test_decode_matches_oracle(lam, L):
    print(f"case lam={lam} L={L}", flush=True)
    got, ref, sym = _case(lam, L, 3_000_000)
    assert np.array_equal(got, ref)
    assert np.array_equal(got[::-1], sym)


@pytest.mark.parametrize("S,K", [(64, 8), (256, 16), (1024, 16), (4096, 32)])
def test_decode_geometries(S, K):
    """Small chunks force many walk failures and fixup rounds."""
    print(f"case S={S} K={K}", flush=True)
    got, ref, sym = _case(0.1, 4096, 2_000_000, chunk_units=S, checkpoint_units=K)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 1000, 4095, 65537])
def test_small_streams(n):
    """n=1 and other single-symbol streams have no coded bits and a table whose only
    symbol owns every state (zero-bit self-loops; the reference decoder never
    terminates on them)."""
    print(f"case n={n}", flush=True)
    got, ref, sym = _case(0.5, 1024, n)
    assert np.array_equal(got, ref)
    assert np.array_equal(got[::-1], sym)


@pytest.mark.parametrize("n", [1, 50, 100000])
def test_constant_stream(n):
    print(f"case constant n={n}", flush=True)
    import torch
    from multians_amd.decoder import DeviceStream, encode_symbols
    sym = np.full(n, 77, np.uint8)
    es = encode_symbols(sym, 1024)
    assert es.n_units == 1 and es.first_bit == 0
    ds = DeviceStream(es)
    out = ds.decode()
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(out[:n].cpu().numpy(), sym)


@pytest.mark.parametrize("polls", [0, 1])
def test_range_wait_runs_out(polls):
    """A range whose wait for its predecessor's exit runs out leaves its first chunk to
    the fixup rounds; the result must not change.  Forced from the host by a tiny poll
    limit (every wait still ends), on a stream of many 256-chunk ranges."""
    from multians_amd import _lib
    print(f"case polls={polls}", flush=True)
    _lib.set_wait_polls(polls)
    try:
        got, ref, sym = _case(0.1, 4096, 2_000_000, chunk_units=64, checkpoint_units=8)
        got2, ref2, _ = _case(1.0, 1024, 3_000_000, seed=5, chunk_units=128, checkpoint_units=16)
    finally:
        _lib.set_wait_polls()
    assert np.array_equal(got, ref)
    assert np.array_equal(got[::-1], sym)
    assert np.array_equal(got2, ref2)


def test_fixup_round_limit_reports():
    """The fixup rounds stop at their bound and report a stream that is still
    inconsistent (status MANS_ECORRUPT) instead of looping on.  Forced from the host by
    a one-round limit on a stream with 16-unit chunks, whose walks fail for many chunks
    in a row (the merge distance is hundreds of units); with the default bound the same
    stream decodes bit-exact."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, encode_symbols
    print("case round limit", flush=True)
    sym = ol.gen_symbols(17, 1.0, 400_000)
    es = encode_symbols(sym, 4096)
    ds = DeviceStream(es, chunk_units=16, checkpoint_units=16)
    _lib.set_sync_round_limit(1)
    try:
        ds.decode(chunk_units=16, checkpoint_units=16)
        torch.cuda.synchronize()
        limited = ds.status()
    finally:
        _lib.set_sync_round_limit()
    assert limited == _lib.MANS_ECORRUPT
    out = ds.decode(chunk_units=16, checkpoint_units=16)
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(out[:len(sym)].cpu().numpy()[::-1], sym)


@pytest.mark.parametrize("spec", [(2026181085674, 0.25656187489628735, 1_105_978),
                                  (2026181087171, 0.17296609080379266, 1_971_802)])
def test_walk_enters_chunk_as_another_leaves(spec):
    """Streams of one 256-chunk range whose spec warm-up puts most chunk starts on the
    true path: a walk that did not merge in its own chunk enters the next one and merges
    at its first checkpoint in the same round in which that chunk's own walk closes it.
    Both write the chunk's record; the closing one must land first.  Decoded four times
    (the order of the two writes is timing-dependent if it is not enforced)."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, EncodedStream
    seed, lam, n = spec
    print(f"case walk-enter seed={seed}", flush=True)
    units, fb, fs, dec, sym = _lib.make_stream(seed, lam, n, 4096, True)
    ds = DeviceStream(EncodedStream(units, fb, fs, dec, 4096, n))
    st = ol.Stream(units, fb, fs, dec, 4096, n)
    ref, _, _, _, _ = ol.decode_seq(st)
    for _ in range(4):
        ds.ws.zero_()
        out = ds.decode()
        torch.cuda.synchronize()
        assert ds.status() == 0
        got = out[:n].cpu().numpy()
        assert np.array_equal(got, ref)
        assert np.array_equal(got[::-1], sym)
