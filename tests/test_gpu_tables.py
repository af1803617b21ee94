"""GPU parity of the device table builder (mans_build_tables_device): counts and decoder
tables equal the host codec's (mans_histogram -> mans_normalize_counts ->
mans_build_tables, pinned against the reference's ANSTableGenerator in
test_oracle.py), and a stream decodes bit-exact with the device-built table.  Bar:
bit-exact."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _host(sym, L):
    from multians_amd import _lib
    counts = _lib.normalize_counts(_lib.histogram(sym), L)
    _, dec = _lib.build_tables(counts, L, want_enc=False)
    return counts, dec


def _device(syms_L):
    import torch
    from multians_amd import _lib
    jobs = []
    for sym, L in syms_L:
        t = torch.from_numpy(np.ascontiguousarray(sym)).cuda()
        jobs.append((t, L, torch.zeros(L, dtype=torch.int32, device="cuda"),
                     torch.zeros(256, dtype=torch.int32, device="cuda")))
    status = _lib.build_tables_device(jobs)
    torch.cuda.synchronize()
    return status, [(j[3].cpu().numpy().view(np.uint32), j[2].cpu().numpy().view(np.uint32))
                    for j in jobs]


@pytest.mark.parametrize("lam", [0.002, 0.01, 0.1, 1.0, 3.0])
@pytest.mark.parametrize("L", [1024, 4096])
def test_tables_match_host(lam, L):
    print(f"case tables lam={lam} L={L}", flush=True)
    sym = ol.gen_symbols(61, lam, 1_000_003)
    status, [(counts, dec)] = _device([(sym, L)])
    assert status == [0]
    hc, hd = _host(sym, L)
    assert np.array_equal(counts, hc)
    assert np.array_equal(dec, hd)


def test_tables_edge_distributions():
    """One symbol, two symbols, tiny inputs, unaligned lengths, a 64-state table and a
    batch of all of them in one call."""
    print("case tables edges", flush=True)
    rng = np.random.default_rng(5)
    cases = [
        (np.full(1, 7, np.uint8), 1024),
        (np.full(100_000, 255, np.uint8), 4096),
        (np.array([0, 1] * 50_001 + [1], np.uint8), 1024),
        (rng.integers(0, 256, 17, dtype=np.uint8), 4096),
        (rng.integers(0, 40, 1_234_567, dtype=np.uint8), 64),
        (ol.gen_symbols(3, 0.05, 4_194_305), 4096),
        (np.concatenate([np.zeros(10_000_000, np.uint8), np.arange(256, dtype=np.uint8)]), 4096),
    ]
    status, got = _device(cases)
    assert status == [0] * len(cases)
    for (sym, L), (counts, dec) in zip(cases, got):
        hc, hd = _host(sym, L)
        assert np.array_equal(counts, hc), L
        assert np.array_equal(dec, hd), L


def test_tables_reject_too_many_symbols():
    """More distinct symbols than states cannot be normalised: status 2, table untouched."""
    print("case tables reject", flush=True)
    sym = np.arange(256, dtype=np.uint8).repeat(10)
    status, [(counts, dec)] = _device([(sym, 64)])
    assert status == [2]
    assert not dec.any()


def test_device_table_decodes_stream():
    """Encode with the host codec's tables, decode with the device-built decoder table:
    bit-exact against the symbols."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, EncodedStream
    print("case tables round trip", flush=True)
    sym = ol.gen_symbols(8, 0.2, 3_000_000)
    counts = _lib.normalize_counts(_lib.histogram(sym), 4096)
    enc, _ = _lib.build_tables(counts, 4096)
    units, fb, fs = _lib.encode(sym, enc, 4096)
    status, [(dcounts, ddec)] = _device([(sym, 4096)])
    assert status == [0]
    ds = DeviceStream(EncodedStream(units, fb, fs, ddec.copy(), 4096, len(sym), dcounts))
    out = ds.decode()
    torch.cuda.synchronize()
    assert ds.status() == 0
    assert np.array_equal(out[:len(sym)].cpu().numpy()[::-1], sym)


def test_tables_large_input_histogram():
    """A 2^30-symbol input spans many histogram workgroups."""
    import torch
    from multians_amd import _lib
    print("case tables 2^30", flush=True)
    n = 1 << 30
    sym = _lib.gen_symbols(77, 0.3, n, threads=16)
    t = torch.from_numpy(sym).cuda()
    table = torch.zeros(4096, dtype=torch.int32, device="cuda")
    counts = torch.zeros(256, dtype=torch.int32, device="cuda")
    status = _lib.build_tables_device([(t, 4096, table, counts)])
    assert status == [0]
    hc, hd = _host(sym, 4096)
    assert np.array_equal(counts.cpu().numpy().view(np.uint32), hc)
    assert np.array_equal(table.cpu().numpy().view(np.uint32), hd)


@pytest.mark.gpu
@pytest.mark.parametrize("lam", [0.002, 0.01, 0.3, 1.0, 3.0])
def test_device_symbols_equal_host(lam):
    """mans_gen_symbols_device writes the host generator's bytes (whole blocks, a ragged
    tail, and an output that is not 16-byte aligned)."""
    import torch
    from multians_amd import _lib
    print(f"case gen lam={lam}", flush=True)
    for n, skew in [(1, 0), (17, 0), (1_000_003, 0), (5_000_000, 3)]:
        host = _lib.gen_symbols(77, lam, n, threads=4)
        buf = torch.zeros(n + 32, dtype=torch.uint8, device="cuda")
        _lib.gen_symbols_device(77, lam, buf[skew:skew + n])
        torch.cuda.synchronize()
        got = buf.cpu().numpy()
        assert np.array_equal(got[skew:skew + n], host)
        assert not got[:skew].any() and not got[skew + n:].any()
