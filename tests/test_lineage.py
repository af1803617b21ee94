"""Lineage synchronisation of a single stream (multians_amd/csrc/mans_lineage.h), run on
the CPU by the emulation build (libmans_emu.so: lineages in the kernel's ticket order,
then the shared map / resolve / scan / trace lane code and the trusted write pass),
against the oracle.

The item records the lineage path produces must be the true path's (the oracle's
sequential decoder: entry of every 32-unit subsequence and the steps before it), and the
output the original symbols.  The failure paths (slot lists full, a lineage cap) must
report failure rather than produce records: the decode then runs the chunked pipeline.
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
REC = np.dtype([("state", np.uint32), ("bit", np.uint32), ("before", np.uint64)])


@pytest.fixture(scope="module")
def emu():
    lib = C.CDLL(os.path.join(ROOT, "multians_amd", "lib", "libmans_emu.so"))
    lib.mans_emu_lineage_decode.argtypes = [_u32p, C.c_uint64, _u8p, C.c_uint64, _u32p,
                                            C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_uint32, _u64p, C.c_void_p]
    return lib


def _run(emu, st, D, n_out=None, slot_limit=0, cap=0):
    n_out = st.n if n_out is None else n_out
    out = np.zeros(max(n_out, 1), np.uint8)
    stats = np.zeros(8, np.uint64)
    nq = -(-len(st.units) // 32)
    rec = np.zeros(nq, REC)
    rc = emu.mans_emu_lineage_decode(st.units, len(st.units), out, n_out, st.table, st.L,
                                     st.first_state, st.first_bit, D, slot_limit, cap, stats,
                                     rec.ctypes.data)
    return rc, out[:n_out], stats, rec


def _check_records(st, rec):
    _, entries, counts, _, _ = ol.decode_seq(st, S=32)
    assert np.array_equal(rec["state"], entries & 0xFFFF)
    assert np.array_equal(rec["bit"], entries >> 16)
    before = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]])
    assert np.array_equal(rec["before"], before)


@pytest.mark.parametrize("lam", [0.002, 0.05, 0.2, 1.0, 3.0])
@pytest.mark.parametrize("D", [32, 128, 512])
def test_lineage_matches_oracle(emu, lam, D):
    sym = ol.gen_symbols(41, lam, 600_000)
    st, _ = ol.encode(sym, L=4096)
    rc, out, stats, rec = _run(emu, st, D)
    assert rc == 0, "lineage records did not hold (reason %d)" % stats[2]
    assert np.array_equal(out[::-1], sym)
    _check_records(st, rec)
    # every lineage start at or after a checkpoint boundary, plus lineage 0
    assert stats[0] >= 1 and stats[4] <= 16


@pytest.mark.parametrize("L", [64, 1024])
@pytest.mark.parametrize("n", [1, 7, 300, 4000, 40_000])
def test_lineage_small_streams(emu, L, n):
    sym = ol.gen_symbols(1000 + n, 0.3, n)
    st, _ = ol.encode(sym, L=L)
    rc, out, stats, rec = _run(emu, st, 32)
    assert rc == 0
    assert np.array_equal(out[::-1], sym)
    _check_records(st, rec)


def test_lineage_zero_bit_tail(emu):
    # the first-encoded symbol dominates: its trailing zero-bit steps decode after the last
    # unit (the reference loses them when n_units % S == 0)
    for seed in range(6):
        sym = ol.gen_symbols(seed, 3.0, 50_000 + 977 * seed)
        st, _ = ol.encode(sym, L=4096)
        rc, out, _, rec = _run(emu, st, 64)
        assert rc == 0
        assert np.array_equal(out[::-1], sym)


def test_lineage_failure_paths(emu):
    sym = ol.gen_symbols(5, 0.2, 400_000)
    st, _ = ol.encode(sym, L=4096)
    # one slot per checkpoint: the second lineage to pass a checkpoint finds it full
    rc, _, stats, _ = _run(emu, st, 32, slot_limit=1)
    assert rc == 1 and (int(stats[2]) & 1)
    # lineages capped at one checkpoint: they stop before merging
    rc, _, stats, _ = _run(emu, st, 256, cap=1)
    assert rc == 1 and (int(stats[2]) & 2)
    # a cap far above the merge distances changes nothing
    rc, out, _, _ = _run(emu, st, 256, cap=4096)
    assert rc == 0 and np.array_equal(out[::-1], sym)


def test_lineage_work(emu):
    # the model behind the method: lineages every 128 units merge after a few hundred
    # units, so all lineages together decode a small multiple of the stream
    sym = ol.gen_symbols(9, 0.2, 2_000_000)
    st, _ = ol.encode(sym, L=4096)
    rc, out, stats, _ = _run(emu, st, 128)
    assert rc == 0
    n_q = -(-len(st.units) // 32)
    work = stats[1] / n_q
    assert 1.5 < work < 8.0, work
