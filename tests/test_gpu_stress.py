"""Stress: one stream of more than 2^32 uint8 symbols (64-bit output offsets), 4096
states, lambda = 3.0 (skewed: symbol 0 holds 95% of the mass, so the first-encoded
symbols decode as zero-bit steps after the last unit).  Bit-exact against the original
symbols, compared on the device."""
import pytest


@pytest.mark.gpu
def test_stream_beyond_4gib_symbols():
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, EncodedStream
    n = (1 << 32) + 4097
    print("case stress n=%d lambda=3.0" % n, flush=True)
    sym = _lib.gen_symbols(91, 3.0, n, threads=16)
    counts = _lib.normalize_counts(_lib.histogram(sym, threads=16), 4096)
    enc, dec = _lib.build_tables(counts, 4096)
    units, fb, fs = _lib.encode(sym, enc, 4096)
    del enc
    es = EncodedStream(units, fb, fs, dec, 4096, n, counts)
    print("encoded: %d units" % es.n_units, flush=True)
    ds = DeviceStream(es)
    out = ds.decode()
    torch.cuda.synchronize()
    assert ds.status() == 0
    ref = torch.from_numpy(sym).cuda()
    del sym
    assert torch.equal(out[:n], ref.flip(0))
