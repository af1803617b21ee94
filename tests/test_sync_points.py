"""Sync points (mans_sync_point) on the CPU: the emulation build's records against the
oracle's sequential decode, decodes that start from recorded sync points (right, and
deliberately wrong records planted on the host), and the partial-decode helper
mans_stream_range (host-only logic of the product library).

The truth for chunk c of length S is the oracle's entry of subsequence c
(mo_decode_seq with S: state x - L in bits 0-15, bit offset in bits 16-21) and the sum
of the step counts of subsequences 0 .. c-1.  Bar: bit-exact.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def emu():
    lib = C.CDLL(os.path.join(ROOT, "multians_amd", "lib", "libmans_emu.so"))
    lib.mans_emu_decode_index.argtypes = [_u32p, C.c_uint64, _u8p, C.c_uint64, _u32p, C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_void_p, C.c_void_p, _u64p]
    return lib


def _truth(st, S):
    """Oracle sync points at chunk length S (SYNC_POINT_DTYPE records)."""
    from multians_amd import _lib
    _, entries, counts, _, _ = ol.decode_seq(st, S=S)
    rec = np.zeros(len(entries), _lib.SYNC_POINT_DTYPE)
    rec["state"] = entries & 0xFFFF
    rec["bit"] = (entries >> 16) & 63
    rec["symbols_before"] = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]])
    return rec


def _emu(emu, st, S, index_in=None, want_out=True):
    from multians_amd import _lib
    n = -(-len(st.units) // S)
    out = np.zeros(st.n, np.uint8)
    stats = np.zeros(8, np.uint64)
    rec = np.zeros(n, _lib.SYNC_POINT_DTYPE) if want_out else None
    rc = emu.mans_emu_decode_index(st.units, len(st.units), out, st.n, st.table, st.L,
                                   st.first_state, st.first_bit, S, 16,
                                   index_in.ctypes.data if index_in is not None else None,
                                   rec.ctypes.data if want_out else None, stats)
    assert rc == 0
    return out, rec, stats


def _stream(seed, lam, n, L=4096):
    sym = ol.gen_symbols(seed, lam, n)
    st, _ = ol.encode(sym, L=L)
    return sym, st


@pytest.mark.parametrize("lam,S", [(0.01, 256), (0.2, 512), (1.0, 128), (3.0, 64)])
def test_index_out_is_oracle_truth(emu, lam, S):
    sym, st = _stream(71, lam, 400_000)
    out, rec, _ = _emu(emu, st, S)
    assert np.array_equal(out[::-1], sym)
    assert np.array_equal(rec, _truth(st, S))


@pytest.mark.parametrize("lam,S", [(0.05, 256), (1.0, 64)])
def test_index_in_decodes_without_failed_walks(emu, lam, S):
    """From the true records every chunk's speculative decode is the true path: no walk
    fails, and the records written again are the same."""
    sym, st = _stream(72, lam, 500_000)
    truth = _truth(st, S)
    out, rec, stats = _emu(emu, st, S, index_in=truth)
    assert np.array_equal(out[::-1], sym)
    assert np.array_equal(rec, truth)
    assert stats[2] == 0 and stats[3] == 0           # no failed walks, no fixup rounds


@pytest.mark.parametrize("how", ["random", "shifted", "zero", "out_of_range"])
def test_wrong_records_cost_walks_not_correctness(emu, how):
    """Records planted wrong on the host (random, shifted by one chunk, all zero, states
    and bits beyond their ranges): the decode still proves every chunk start."""
    sym, st = _stream(73, 0.2, 300_000)
    S = 128
    truth = _truth(st, S)
    bad = truth.copy()
    rng = np.random.default_rng(5)
    if how == "random":
        bad["state"] = rng.integers(0, st.L, len(bad))
        bad["bit"] = rng.integers(0, 32, len(bad))
    elif how == "shifted":
        bad[1:] = truth[:-1]
    elif how == "zero":
        bad[:] = 0
    else:
        bad["state"] = 0xFFFFFFFF
        bad["bit"] = 999
    out, rec, _ = _emu(emu, st, S, index_in=bad)
    assert np.array_equal(out[::-1], sym)
    assert np.array_equal(rec, truth)


def _full_desc(st, S, units_base):
    from multians_amd import _lib
    d = _lib.StreamDesc()
    d.units, d.n_units, d.out, d.n_out = units_base, len(st.units), 0, st.n
    d.table, d.L, d.first_state, d.first_bit = 8, st.L, st.first_state, st.first_bit
    d.chunk_units = S
    return d


@pytest.mark.parametrize("lam,S", [(0.02, 256), (2.0, 64)])
def test_stream_range_windows(emu, lam, S):
    """mans_stream_range's sub-stream, decoded by the emulation build from its first
    chunk's record, holds the window at `skip`: windows at both ends, inside one chunk,
    across many chunks, one symbol, the whole stream."""
    from multians_amd import _lib
    sym, st = _stream(74, lam, 250_000)
    full_out = sym[::-1]
    truth = _truth(st, S)
    ubase = st.units.ctypes.data
    ibase = truth.ctypes.data
    d = _full_desc(st, S, ubase)
    n = st.n
    mid = int(truth["symbols_before"][len(truth) // 2])
    rng = np.random.default_rng(9)
    wins = [(0, 1), (0, 1000), (n - 1, n), (n - 5000, n), (0, n), (mid, mid + 1), (mid - 1, mid + 1)]
    for _ in range(12):
        a = int(rng.integers(0, n - 1))
        wins.append((a, int(rng.integers(a + 1, min(n, a + 40_000) + 1))))
    for first, last in wins:
        sub, skip = _lib.stream_range(d, truth, ibase, first, last)
        u0 = (sub.units - ubase) // 4
        k0 = (sub.index_in - ibase) // 16
        assert u0 == k0 * S and skip <= first
        units = np.ascontiguousarray(st.units[u0:u0 + sub.n_units])
        sst = ol.Stream(units, sub.first_bit, sub.first_state, st.table, st.L, sub.n_out)
        rin = np.ascontiguousarray(truth[k0:])
        out, _, _ = _emu(emu, sst, S, index_in=rin, want_out=False)
        assert np.array_equal(out[skip:skip + last - first], full_out[first:last]), (first, last)


def test_stream_range_rejects_bad_input():
    from multians_amd import _lib
    sym, st = _stream(75, 0.3, 20_000)
    S = 64
    truth = _truth(st, S)
    d = _full_desc(st, S, st.units.ctypes.data)
    ib = truth.ctypes.data
    for first, last in [(5, 5), (10, 3), (0, st.n + 1)]:
        with pytest.raises(_lib.MansError):
            _lib.stream_range(d, truth, ib, first, last)
    d2 = _full_desc(st, 100, st.units.ctypes.data)              # not a power of two
    with pytest.raises(_lib.MansError):
        _lib.stream_range(d2, truth, ib, 0, 10)
    bad = truth.copy()
    bad["symbols_before"][0] = 7                                # must start at 0
    with pytest.raises(_lib.MansError):
        _lib.stream_range(d, bad, ib, 0, 10)
    bad = truth.copy()
    bad["symbols_before"][3:] = 0                               # not ascending
    with pytest.raises(_lib.MansError):
        _lib.sync_points_check(bad, st.L)
    _lib.sync_points_check(truth, st.L)
    bad = truth.copy()
    bad["bit"][2] = 32                                          # 32 only in record 0
    with pytest.raises(_lib.MansError):
        _lib.sync_points_check(bad, st.L)
    bad = truth.copy()
    bad["state"][4] = st.L                                      # state out of range
    with pytest.raises(_lib.MansError):
        w = int(truth["symbols_before"][4])
        _lib.stream_range(d, bad, ib, w, w + 1)


def test_sync_point_count():
    from multians_amd import _lib
    assert _lib.sync_point_count(1, 16) == 1
    assert _lib.sync_point_count(4096, 256) == 16
    assert _lib.sync_point_count(4097, 256) == 17
    assert _lib.sync_point_count(0, 256) == 0
    for bad in (0, 8, 100, 1 << 25):
        with pytest.raises(_lib.MansError):
            _lib.sync_point_count(100, bad)
