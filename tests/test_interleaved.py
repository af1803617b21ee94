"""Two-state interleaved streams on the CPU: the product's host encoder
(mans_encode_interleaved) against the oracle's encoder and its records against the
oracle's decoder, and the device decoder's per-lane code (mans_il_lane.h, run lane by lane
by the emulation build) against the oracle's decode.

The format is the library's own (the reference names multiple states / tables as future
work, README.md:11); the oracle restating it is pinned in tests/test_oracle.py (N = 1 is
the reference's single-state format; N = 2 against a pure-Python brute force).
Bar: bit-exact units, states, records and symbols.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
from multians_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def emu():
    lib = C.CDLL(os.path.join(ROOT, "multians_amd", "lib", "libmans_emu.so"))
    lib.mans_emu_il_decode.argtypes = [_u32p, C.c_uint64, _u32p, _u32p, C.c_uint32, _u32p,
                                       C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p, _u8p,
                                       _u32p]
    return lib


def _tables(sym, L, two):
    """Per-slot tables: slot 0 from the histogram, slot 1 (two=True) from the same support
    with the present symbols' counts rotated."""
    hist = np.bincount(sym, minlength=256).astype(np.uint64)
    c0 = ol.normalize(hist, L)
    if not two:
        return [c0, c0]
    nz = np.nonzero(hist)[0]
    h = np.zeros_like(hist)
    h[nz] = np.roll(hist[nz], 1)
    return [c0, ol.normalize(h, L)]


def _make(seed, lam, n, L, R, two):
    sym = ol.gen_symbols(seed, lam, n) if n else np.zeros(0, np.uint8)
    if n == 0:
        sym = np.zeros(0, np.uint8)
    counts = _tables(sym if n else np.array([3], np.uint8), L, two)
    il, _ = ol.encode_interleaved(sym, 2, counts, L)
    built = [ol.build_tables(c, L) for c in counts]
    enc = [b[1] for b in built]
    units, fb, fs, rec = _lib.encode_interleaved(sym, enc[0], enc[1] if two else enc[0], L, R)
    return sym, il, units, fb, fs, rec


def _emu_decode(emu, il, units, fb, fs, rec, R, tables=None):
    tables = il.tables if tables is None else tables
    n = il.n
    out = np.zeros(n + 64, np.uint8)         # 64-byte aligned start, like a device buffer
    M = len(rec)
    lane = np.zeros(M + 1, np.uint32)
    rec = np.ascontiguousarray(rec)
    rc = emu.mans_emu_il_decode(units, len(units), tables[0], tables[1], il.L,
                                np.ascontiguousarray(fs, np.uint32), fb, n, R,
                                rec.ctypes.data, out, lane)
    return rc, out[:n], lane


CASES = [(0.002, 1024), (0.05, 4096), (0.3, 4096), (1.0, 1024), (3.0, 64), (0.5, 16384)]


@pytest.mark.parametrize("lam,L", CASES)
@pytest.mark.parametrize("two", [False, True])
def test_encoder_matches_oracle(lam, L, two):
    """Units, first bit and both final states as the oracle's encoder; records as the
    oracle's decoder derives them (end-anchored, R = 64 and a prime R = 37)."""
    for R in (64, 37):
        sym, il, units, fb, fs, rec = _make(5, lam, 60_000, L, R, two)
        assert np.array_equal(units, il.units)
        assert fb == il.first_bit and list(fs) == list(il.first_states)
        out, orec, es, eb = ol.decode_interleaved(il, R=R)
        assert np.array_equal(out[::-1], sym)
        assert len(rec) == len(orec)
        assert np.array_equal(rec["state"], orec["state"])
        assert np.array_equal(rec["bit"], orec["bit"])
        assert np.array_equal(rec["before"], orec["before"])


@pytest.mark.parametrize("lam,L", CASES)
@pytest.mark.parametrize("R", [16, 23, 64, 1000])
def test_lane_code_matches_oracle(emu, lam, L, R):
    """Every lane of the device decoder (ring blocks, guard word, pair and single steps,
    head and tail stores) through the emulation build: the oracle's symbols, and every
    interval ends on its next record."""
    for two in (False, True):
        sym, il, units, fb, fs, rec = _make(9, lam, 50_000, L, R, two)
        ref, *_ = ol.decode_interleaved(il)
        rc, out, lane = _emu_decode(emu, il, units, fb, fs, rec, R)
        assert rc == 0 and not lane.any()
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 17, 100])
def test_tiny_streams(emu, n):
    for two in (False, True):
        sym, il, units, fb, fs, rec = _make(3, 0.7, n, 64, 16, two)
        assert np.array_equal(units, il.units)
        rc, out, lane = _emu_decode(emu, il, units, fb, fs, rec, 16)
        assert rc == 0 and not lane.any()
        assert np.array_equal(out[::-1], sym)


def test_constant_stream(emu):
    """One symbol owning every state: every step reads zero bits, one unit, and the whole
    stream falls into the interval after record 0."""
    sym = np.full(5000, 42, np.uint8)
    counts = np.zeros(256, np.uint32)
    counts[42] = 256
    il, _ = ol.encode_interleaved(sym, 2, [counts, counts], 256)
    enc = ol.build_tables(counts, 256)[1]
    units, fb, fs, rec = _lib.encode_interleaved(sym, enc, enc, 256, 16)
    assert len(units) == 1 and len(rec) == 1 and int(rec[0]["before"]) == 0
    rc, out, lane = _emu_decode(emu, il, units, fb, fs, rec, 16)
    assert rc == 0 and np.array_equal(out, sym)


def test_corrupt_records_are_reported(emu):
    """A record with a wrong state, a wrong bit or out-of-order counts makes the lanes
    next to it report (status bit 2); the others still decode their intervals."""
    sym, il, units, fb, fs, rec = _make(21, 0.4, 80_000, 1024, 32, True)
    ref, *_ = ol.decode_interleaved(il)
    M = len(rec)
    for field, m in (("state", M // 2), ("bit", M // 3), ("before", M // 4)):
        bad = rec.copy()
        if field == "state":
            bad[m]["state"][0] ^= 1
        elif field == "bit":
            bad[m]["bit"] += 1
        else:
            bad[m]["before"] = bad[m + 1]["before"] - 1 if m + 1 < M else 0
        rc, out, lane = _emu_decode(emu, il, units, fb, fs, bad, 32)
        assert rc == 2, field
        flagged = set(np.nonzero(lane)[0])
        # lanes M-m-1 (ends at record m) and M-m (starts there)
        assert flagged and flagged <= {M - m - 1, M - m}, (field, flagged)
        good = [q for q in range(M + 1) if q not in (M - m - 1, M - m)]
        starts = [0] + [int(rec[M - q]["before"]) for q in range(1, M + 1)]
        ends = [int(rec[M - q - 1]["before"]) for q in range(M)] + [len(sym)]
        for q in good[:50]:
            assert np.array_equal(out[starts[q]:ends[q]], ref[starts[q]:ends[q]])


def test_record_count_and_arguments():
    assert _lib.il_record_count(0, 16) == 0
    assert _lib.il_record_count(1, 16) == 1
    assert _lib.il_record_count(33, 16) == 3
    with pytest.raises(_lib.MansError):
        _lib.il_record_count(10, 15)
    sym = ol.gen_symbols(1, 0.5, 1000)
    c = _tables(sym, 256, False)[0]
    enc = ol.build_tables(c, 256)[1]
    with pytest.raises(_lib.MansError):
        _lib.encode_interleaved(sym, enc, enc, 256, 8)          # record_units below 16
    bad = sym.copy()
    bad[7] = 255 if c[255] == 0 else bad[7]
    if c[255] == 0:
        with pytest.raises(_lib.MansError):
            _lib.encode_interleaved(bad, enc, enc, 256, 16)      # symbol without a code
