"""GPU parity of the device encoder (mans_encode_device): units, first_bit and
first_state equal the host codec's mans_encode (pinned against the reference's
ANSEncoder in test_oracle.py / test_dropin_cpu.py), for sizes around the chunk and range
boundaries, several entropies and table sizes, batches of mixed jobs, one-symbol
tables, and the error statuses; and a device round trip (encode, then decode).
Bar: bit-exact.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _host(sym, L):
    from multians_amd import _lib
    counts = _lib.normalize_counts(_lib.histogram(sym), L)
    enc, dec = _lib.build_tables(counts, L)
    units, fb, fs = _lib.encode(sym, enc, L)
    return units, fb, fs, dec


def _job(sym, dec, L, cap=None):
    import torch
    from multians_amd import _lib
    cap = _lib.lib().mans_max_encoded_units(len(sym)) if cap is None else cap
    return (torch.from_numpy(np.ascontiguousarray(sym)).cuda(),
            torch.from_numpy(dec.view(np.int32).copy()).cuda(), L,
            torch.full((max(cap, 1),), -1, dtype=torch.int32, device="cuda"))


def _check(sym, L, res, job):
    units, fb, fs, _ = _host(sym, L)
    nu, gfb, gfs, status = res
    assert status == 0
    assert (nu, gfb, gfs) == (len(units), fb, fs)
    got = job[3][:nu].cpu().numpy().view(np.uint32)
    assert np.array_equal(got, units)


@pytest.mark.parametrize("lam", [0.002, 0.05, 0.5, 3.0])
@pytest.mark.parametrize("L", [1024, 4096])
def test_encode_matches_host(lam, L):
    from multians_amd import _lib
    print(f"case encode lam={lam} L={L}", flush=True)
    sym = ol.gen_symbols(21, lam, 3_000_001)
    dec = _host(sym, L)[3]
    job = _job(sym, dec, L)
    [res] = _lib.encode_device([job])
    _check(sym, L, res, job)


@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 4095, 4096, 4097, 8197, 256 * 4096,
                               256 * 4096 + 1, 600 * 4096 + 77])
def test_encode_sizes(n):
    """Chunk (4,096 symbols) and range (256 chunks) boundaries, and the empty stream."""
    from multians_amd import _lib
    print(f"case encode n={n}", flush=True)
    sym = ol.gen_symbols(22, 0.3, n) if n else np.zeros(0, np.uint8)
    L = 4096
    dec = _host(sym if n else np.array([1], np.uint8), L)[3]
    job = _job(sym, dec, L)
    [res] = _lib.encode_device([job])
    if n == 0:
        assert res == (1, 0, L, 0)
        assert int(job[3][0]) == 0
    else:
        _check(sym, L, res, job)


def test_encode_batch_and_tables():
    """One call, 40 jobs: L from 64 to 16384, one-symbol and two-symbol inputs, unaligned
    symbol pointers (a slice starting at an odd offset)."""
    import torch
    from multians_amd import _lib
    print("case encode batch", flush=True)
    rng = np.random.default_rng(8)
    cases, jobs = [], []
    for k in range(40):
        L = [64, 1024, 4096, 16384][k % 4]
        n = int(rng.integers(1, 700_000))
        if k % 10 == 3:
            sym = np.full(n, 9, np.uint8)
        elif k % 10 == 5:
            sym = rng.integers(0, 2, n).astype(np.uint8)
        else:
            sym = ol.gen_symbols(100 + k, float(rng.uniform(0.01, 2.0)), n)
        if L == 64:
            sym = np.minimum(sym, 40).astype(np.uint8)
        dec = _host(sym, L)[3]
        job = _job(sym, dec, L)
        if k % 7 == 2:                      # symbols at an odd address
            big = torch.zeros(n + 1, dtype=torch.uint8, device="cuda")
            big[1:] = job[0]
            job = (big[1:], job[1], job[2], job[3])
        cases.append((sym, L))
        jobs.append(job)
    res = _lib.encode_device(jobs)
    for (sym, L), r, j in zip(cases, res, jobs):
        _check(sym, L, r, j)


def test_encode_errors():
    from multians_amd import _lib
    print("case encode errors", flush=True)
    sym = ol.gen_symbols(23, 0.5, 100_000)
    L = 1024
    dec = _host(sym, L)[3]
    other = sym.copy()
    other[500] = 255 if 255 not in sym else 254       # a symbol the table does not hold
    assert not np.any(sym == other[500])
    bad = dec.copy()
    bad[7] = (bad[7] & 0xFFFF0000) | 0xFFFF          # a `next` outside its symbol's range
    jobs = [_job(other, dec, L), _job(sym, dec, L, cap=10), _job(sym, bad, L), _job(sym, dec, L)]
    res = _lib.encode_device(jobs)
    assert res[0][3] == 2
    assert res[1][3] == 4
    assert res[2][3] == 1
    _check(sym, L, res[3], jobs[3])


def test_encode_decode_round_trip():
    """2^27 symbols encoded on the device, decoded on the device."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, EncodedStream
    print("case encode round trip", flush=True)
    n = 1 << 27
    sym = _lib.gen_symbols(24, 0.25, n, threads=16)
    L = 4096
    counts = _lib.normalize_counts(_lib.histogram(sym), L)
    _, dec = _lib.build_tables(counts, L, want_enc=False)
    job = _job(sym, dec, L)
    [(nu, fb, fs, status)] = _lib.encode_device([job])
    assert status == 0
    units = job[3][:nu].cpu().numpy().view(np.uint32).copy()
    ds = DeviceStream(EncodedStream(units, fb, fs, dec, L, n))
    out = ds.decode()
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(out[:n].cpu().numpy()[::-1], sym)


@pytest.mark.parametrize("lam,L,n", [(0.05, 4096, 3_000_001), (1.5, 1024, 2_000_000),
                                     (0.3, 4096, 33 * 1024 * 4), (0.5, 1024, 1), (0.5, 1024, 0)])
def test_encode_writes_decoder_item_records(lam, L, n):
    """The encoder's item records (items_out) equal those a decode of the stream writes
    (for every item that holds a step; trailing items only by symbols_before), and the
    stream decodes bit-exact from them by the write pass alone."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, EncodedStream
    print(f"case encode items lam={lam} L={L} n={n}", flush=True)
    sym = ol.gen_symbols(31, lam, n) if n else np.zeros(0, np.uint8)
    dec = _host(sym if n else np.array([3], np.uint8), L)[3]
    job = _job(sym, dec, L)
    cap = job[3].numel()
    items = torch.zeros(_lib.sync_point_count(cap, 32) * 16, dtype=torch.uint8, device="cuda")
    [(nu, fb, fs, status)] = _lib.encode_device([job + (items,)])
    assert status == 0
    units = job[3][:nu].cpu().numpy().view(np.uint32).copy()
    es = EncodedStream(units, fb, fs, dec, L, n)
    nrec = _lib.sync_point_count(nu, 32)
    enc_rec = items[:16 * nrec].cpu().numpy().view(_lib.SYNC_POINT_DTYPE)
    if n:
        ds = DeviceStream(es)
        drec = ds.new_index(32)
        ds.decode(items_out=drec)
        torch.cuda.synchronize()
        dec_rec = drec[:16 * nrec].cpu().numpy().view(_lib.SYNC_POINT_DTYPE)
        assert np.array_equal(enc_rec["symbols_before"], dec_rec["symbols_before"])
        live = enc_rec["symbols_before"] < n
        assert np.array_equal(enc_rec[live], dec_rec[live])
        ds2 = DeviceStream(es, chunk_units=32, checkpoint_units=16)
        out = ds2.decode(chunk_units=32, index_in=items[:16 * nrec],
                         flags=_lib.MANS_F_TRUST_SYNC_POINTS)
        torch.cuda.synchronize()
        assert ds2.status() == 0
        assert np.array_equal(out[:n].cpu().numpy()[::-1], sym)
    else:
        assert (nrec, int(enc_rec["state"][0]), int(enc_rec["bit"][0])) == (1, 0, 32)


def test_encode_round_limit_reports():
    """The encoder's fixup rounds stop at their bound and report a stream still
    inconsistent then (result status 8) instead of looping on.  Forced from the host by
    a one-round limit on low-entropy streams of 4,096 chunks (encoding from a guessed
    state meets the true path only after 1,000+ symbols there, so failed walks chain over
    consecutive chunks); with the default bound every stream is unit-exact again."""
    from multians_amd import _lib
    print("case encode round limit", flush=True)
    syms = [ol.gen_symbols(60 + i, lam, 1 << 24) for i, lam in enumerate([1.5, 1.9, 2.5, 3.0])]
    jobs = [_job(s, _host(s, 4096)[3], 4096) for s in syms]
    _lib.set_sync_round_limit(1)
    try:
        limited = _lib.encode_device(jobs)
    finally:
        _lib.set_sync_round_limit()
    assert any(r[3] == 8 for r in limited)
    assert all(r[3] in (0, 8) for r in limited)
    for s, res, job in zip(syms, _lib.encode_device(jobs), jobs):
        _check(s, 4096, res, job)
