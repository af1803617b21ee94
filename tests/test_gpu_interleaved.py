"""GPU parity for two-state interleaved streams: mans_decode_interleaved (through the C
ABI) against the oracle's interleaved decoder (mo_decode_interleaved) and the original
symbols (reversed: decode order).

The format is the library's own (the reference names multiple states / tables as future
work, README.md:11); the oracle restating it is pinned in tests/test_oracle.py, the host
encoder and the decoder's per-lane code against it in tests/test_interleaved.py.
Bar: bit-exact; every interval ends on its next record (status 0).
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _oracle(ie):
    st = ol.ILStream(ie.units, ie.first_bit, ie.first_states, list(ie.tables), ie.L, ie.n_symbols)
    out, *_ = ol.decode_interleaved(st)
    return out


def _run(sym, L, R, per_slot, check_oracle=True):
    import torch
    from multians_amd.decoder import DeviceInterleaved, encode_interleaved
    ie = encode_interleaved(sym, L, record_units=R, per_slot_tables=per_slot)
    d = DeviceInterleaved(ie)
    out = d.decode()
    torch.cuda.synchronize()
    assert d.status() == 0
    got = out[:len(sym)].cpu().numpy()
    assert np.array_equal(got[::-1], sym)
    if check_oracle:
        assert np.array_equal(got, _oracle(ie))
    return d, ie


@pytest.mark.parametrize("lam,L", [(0.002, 1024), (0.05, 4096), (0.3, 4096), (1.0, 1024),
                                   (3.0, 64), (0.5, 16384)])
@pytest.mark.parametrize("per_slot", [False, True])
def test_decode_matches_oracle(lam, L, per_slot):
    print(f"case il lam={lam} L={L} per_slot={per_slot}", flush=True)
    _run(ol.gen_symbols(31, lam, 2_000_000), L, 256, per_slot)


@pytest.mark.parametrize("R", [16, 17, 64, 1000, 1 << 16])
def test_record_lengths(R):
    """Short records (many lanes, intervals shorter than the ring) to one interval."""
    print(f"case il R={R}", flush=True)
    _run(ol.gen_symbols(32, 0.2, 1_000_000), 4096, R, True)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 100, 4095, 65537])
def test_small_streams(n):
    print(f"case il n={n}", flush=True)
    for per_slot in (False, True):
        _run(ol.gen_symbols(33, 0.5, n), 1024, 16, per_slot)


def test_constant_and_zero_bit_streams():
    """A constant stream (every step reads zero bits, all in the interval after record 0)
    and a stream whose trailing decoded symbols take zero bits."""
    print("case il constant", flush=True)
    _run(np.full(300_000, 9, np.uint8), 1024, 16, False)
    _run(ol.gen_symbols(34, 0.002, 500_000), 1024, 32, True)


def test_corrupt_record_reported():
    """A record with a wrong state, and one out of order, set MANS_ECORRUPT; the decode
    stays inside the output (its trailing guard bytes are untouched)."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceInterleaved, encode_interleaved
    print("case il corrupt", flush=True)
    sym = ol.gen_symbols(35, 0.4, 400_000)
    ie = encode_interleaved(sym, 1024, record_units=64, per_slot_tables=True)
    d = DeviceInterleaved(ie)
    guard = torch.full((1024,), 0xA5, dtype=torch.uint8, device="cuda")
    big = torch.cat([torch.zeros(ie.n_symbols, dtype=torch.uint8, device="cuda"), guard])
    d.out = big
    M = len(ie.records)
    for mut in ("state", "order"):
        bad = ie.records.copy()
        if mut == "state":
            bad[M // 2]["state"][1] ^= 3
        else:
            bad[M // 3]["before"] = bad[M // 3 + 1]["before"] - 1
        rec = torch.from_numpy(bad.view(np.uint8).copy()).cuda()
        d.decode(records=rec)
        torch.cuda.synchronize()
        assert d.status() == _lib.MANS_ECORRUPT, mut
        assert torch.equal(big[ie.n_symbols:], guard)
    d.decode()
    torch.cuda.synchronize()
    assert d.status() == 0
    assert np.array_equal(big[:ie.n_symbols].cpu().numpy()[::-1], sym)


def test_large_stream_2_28():
    """2^28 symbols at lambda = 1.0, L = 4096, per-slot tables: against the symbols on the
    device, and the oracle on a prefix."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceInterleaved, encode_interleaved
    print("case il 2^28", flush=True)
    n = 1 << 28
    sym = _lib.gen_symbols(36, 1.0, n)
    ie = encode_interleaved(sym, 4096, record_units=256, per_slot_tables=True)
    d = DeviceInterleaved(ie)
    out = d.decode()
    torch.cuda.synchronize()
    assert d.status() == 0
    ref = torch.from_numpy(sym).cuda().flip(0)
    assert torch.equal(out, ref)
    st = ol.ILStream(ie.units, ie.first_bit, ie.first_states, list(ie.tables), ie.L, n)
    head, *_ = ol.decode_interleaved(st, n_out=1 << 22)
    assert np.array_equal(head, out[:1 << 22].cpu().numpy())
