"""The decode pipeline's per-lane logic (multians_amd/csrc/mans_core.h), run on the CPU
by the emulation build (libmans_emu.so, phases in kernel order), against the oracle.

This checks the algorithm -- speculative chunks, merge walks, fixup rounds, scan,
count-driven writes -- for many geometries, including tiny chunks that force long
chains of failed walks, and the checkpoint table (entries + output offsets) against the
oracle's true path.  The GPU tests run the same logic on the device.
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
    lib.mans_emu_decode.argtypes = [_u32p, C.c_uint64, _u8p, C.c_uint64, _u32p, C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u64p,
                                    C.c_void_p, C.c_void_p]
    return lib


def _run(emu, st, S, K, n_out=None, want_ck=False):
    n_out = st.n if n_out is None else n_out
    out = np.zeros(max(n_out, 1), np.uint8)
    stats = np.zeros(8, np.uint64)
    nq = -(-len(st.units) // K)
    ent = np.zeros(nq, np.uint32)
    off = np.zeros(nq, np.uint64)
    rc = emu.mans_emu_decode(st.units, len(st.units), out, n_out, st.table, st.L, st.first_state,
                             st.first_bit, S, K, stats, ent.ctypes.data, off.ctypes.data)
    assert rc == 0
    return out[:n_out], stats, ent, off


@pytest.mark.parametrize("lam", [0.002, 0.01, 0.1, 1.0, 3.0])
@pytest.mark.parametrize("S,K", [(16, 4), (64, 8), (256, 16), (1024, 16), (2048, 64)])
def test_pipeline_matches_oracle(emu, lam, S, K):
    sym = ol.gen_symbols(31, lam, 1_000_000)
    st, _ = ol.encode(sym, L=4096)
    out, stats, ent, off = _run(emu, st, S, K)
    assert np.array_equal(out[::-1], sym)
    ref, entries, counts, _, _ = ol.decode_seq(st, S=K)
    assert np.array_equal(out, ref)
    # checkpoint entries are the true path's; offsets are the exclusive scan of
    # per-K-unit step counts
    assert np.array_equal(ent, entries)
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]]))


@pytest.mark.parametrize("seed", range(12))
def test_random_small_streams(emu, seed):
    rng = np.random.default_rng(seed)
    L = int(rng.choice([64, 256, 1024, 4096]))
    n = int(rng.integers(1, 20000))
    lam = float(rng.choice([0.01, 0.3, 1.0, 4.0]))
    sym = ol.gen_symbols(seed, lam, n)
    if len(np.unique(sym)) > L:
        L = 4096
    st, _ = ol.encode(sym, L=L)
    K = int(rng.choice([1, 2, 4, 8]))
    S = K * int(rng.choice([1, 2, 4, 16]))
    out, *_ = _run(emu, st, S, K)
    assert np.array_equal(out[::-1], sym)


def test_zero_bit_tail_not_lost(emu):
    """Trailing zero-bit symbols at bit 32*n_units are produced for any chunk size
    (the reference loses them when n_units % S == 0)."""
    L = 1024
    counts = np.ones(256, np.uint32)
    counts[7] = L - 255
    rng = np.random.default_rng(1)
    for n in range(3000, 3040):
        data = np.where(rng.random(n) < 0.9, 7, rng.integers(0, 256, n)).astype(np.uint8)
        data[:5] = 7
        st, _ = ol.encode(data, counts, L)
        for S, K in [(4, 4), (4, 1), (8, 2)]:
            out, *_ = _run(emu, st, S, K)
            assert np.array_equal(out[::-1], data)


def test_partial_output(emu):
    sym = ol.gen_symbols(5, 0.2, 50000)
    st, _ = ol.encode(sym, L=1024)
    full, *_ = _run(emu, st, 64, 8)
    for n_out in [1, 17, 4096, 49999]:
        part, *_ = _run(emu, st, 64, 8, n_out=n_out)
        assert np.array_equal(part, full[:n_out])


@pytest.mark.parametrize("lam", [0.05, 0.3, 1.0, 4.0])
@pytest.mark.parametrize("S,K", [(8, 4), (16, 8), (32, 16)])
def test_chained_walks_many_ranges(emu, lam, S, K):
    """Chunks far shorter than the merge distance: walks chain through many chunks,
    ranges of 256 chunks meet at seams, and the consistency check + fixup rounds must
    still yield the true path at every checkpoint."""
    sym = ol.gen_symbols(77, lam, 400_000 if lam < 1.0 else 3_000_000)
    st, _ = ol.encode(sym, L=4096)
    out, stats, ent, off = _run(emu, st, S, K)
    assert stats[0] > 256                        # several ranges
    assert np.array_equal(out[::-1], sym)
    _, entries, counts, _, _ = ol.decode_seq(st, S=K)
    assert np.array_equal(ent, entries)
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]]))


@pytest.mark.parametrize("warm", [16, 256, 2048])
@pytest.mark.parametrize("lam,S", [(0.01, 256), (1.0, 512), (0.1, 64)])
def test_spec_warmup(emu, lam, S, warm):
    """A spec that starts `warm` units before each chunk (its first checkpoint is then the
    warm-up path's state) gives the same output."""
    emu.mans_emu_decode_warm.argtypes = [_u32p, C.c_uint64, _u8p, C.c_uint64, _u32p, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, _u64p]
    sym = ol.gen_symbols(44, lam, 600_000)
    st, _ = ol.encode(sym, L=4096)
    out = np.zeros(len(sym), np.uint8)
    stats = np.zeros(8, np.uint64)
    rc = emu.mans_emu_decode_warm(st.units, len(st.units), out, len(sym), st.table, st.L,
                                  st.first_state, st.first_bit, S, 16, warm, stats)
    assert rc == 0
    assert np.array_equal(out[::-1], sym)


def test_fixup_round_bound(emu):
    """The fixup rounds stop after n_chunks + 1 rounds (or a smaller limit) and report a
    stream still inconsistent then (-5, the kernel's status bit 1) instead of looping on.
    16-unit chunks at lambda = 1 need many rounds (the merge distance is hundreds of
    units): one round is too few, the default bound is never reached, and the number of
    rounds the default run takes is a limit that suffices exactly."""
    emu.mans_emu_decode_limited.argtypes = [_u32p, C.c_uint64, _u8p, C.c_uint64, _u32p,
                                            C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_uint32, _u64p]
    sym = ol.gen_symbols(17, 1.0, 400_000)
    st, _ = ol.encode(sym, L=4096)

    def run(limit):
        out = np.zeros(len(sym), np.uint8)
        stats = np.zeros(8, np.uint64)
        rc = emu.mans_emu_decode_limited(st.units, len(st.units), out, len(sym), st.table, st.L,
                                         st.first_state, st.first_bit, 16, 16, limit, stats)
        return rc, out, stats

    rc, out, stats = run(0)
    assert rc == 0 and np.array_equal(out[::-1], sym)
    rounds = int(stats[3])
    assert 2 <= rounds <= int(stats[0])
    assert run(1)[0] == -5
    assert run(rounds - 1)[0] == -5
    rc, out, _ = run(rounds)
    assert rc == 0 and np.array_equal(out[::-1], sym)


@pytest.mark.parametrize("lam,L", [(0.5, 4096), (1.0, 4096), (2.0, 1024), (3.0, 4096),
                                   (6.0, 4096), (1.0, 64)])
def test_zero_bit_run_folding(emu, lam, L):
    """The sync kernels fold zero-bit runs into their staged tables (one lookup = 1 + z
    steps).  The kernel's in-place pointer jumping, run here sequentially, equals following
    each run step by step, and decoding with the folded words meets every unit boundary in
    the same state, at the same bit, after the same number of steps."""
    emu.mans_emu_fold_check.argtypes = [_u32p, C.c_uint64, _u32p, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint64, _u64p]
    sym = ol.gen_symbols(8, lam, 300_000)
    st, _ = ol.encode(sym, L=L)
    stats = np.zeros(4, np.uint64)
    rc = emu.mans_emu_fold_check(st.units, len(st.units), st.table, st.L, st.first_state,
                                 st.first_bit, len(sym), stats)
    assert rc == 0
    if lam >= 1.0:
        assert stats[0] > 0 and stats[2] < stats[3]     # runs exist and are folded
