"""Pin the oracle (oracle/libmans_oracle.so) before trusting it.

Checks, each chosen so a plausible mistake fails one of them:
  * against the reference's own CPU code compiled from its sources (oracle/_ref):
    tables of the demo distributions (src/main.cc:55-73), encoder output
    (units, first_bit, first_state), MulticoreDecoder output;
  * against committed golden vectors (tests/golden/) made by that same build,
    so the pinning holds where the reference is absent;
  * against invariants: decode(encode(x)) == x, the decoder ends at state L at bit
    32*n_units, brute-force re-derivation of subsequence entries and counts;
  * closed forms: symbol frequencies of mo_gen_symbols vs exp(-lambda k).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
need_ref = pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref not built")


def _ref_demo(lam, L=1024, nsym=256, seed=5):
    R = ol.ref()
    counts = np.zeros(nsym, np.uint32)
    dec = np.zeros(L, np.uint32)
    en, ec, el = (np.zeros(nsym * L, np.uint32) for _ in range(3))
    R.ref_demo_tables(lam, nsym, L, seed, counts, dec, en, ec, el)
    return counts, dec, en, ec, el


@need_ref
@pytest.mark.parametrize("lam", [0.1, 0.26, 0.42, 1.06, 2.34])
@pytest.mark.parametrize("L", [1024, 4096])
def test_tables_match_reference(lam, L):
    counts, dec, en, ec, el = _ref_demo(lam, L)
    _, enc, dec_o = ol.build_tables(counts, L)
    assert np.array_equal(dec, dec_o)
    assert np.array_equal(en, enc[:, :, 0].ravel())
    assert np.array_equal(ec, enc[:, :, 1].ravel())
    assert np.array_equal(el, enc[:, :, 2].ravel())


def _ref_encode(counts, L, data):
    R = ol.ref()
    n = len(data)
    buf = np.zeros(n + 8, np.uint8)
    buf[:n] = data
    units = np.zeros(n // 2 + 16, np.uint32)
    fb, fs = C.c_uint32(), C.c_uint32()
    nu = R.ref_encode(counts, len(counts), L, buf, n, units, len(units), C.byref(fb), C.byref(fs))
    assert nu > 0
    return units[:nu], fb.value, fs.value


def _units_equal(a, b, first_bit):
    """Equal except the don't-care low bits of unit 0 (left-aligned first_bit valid bits)."""
    if len(a) != len(b):
        return False
    mask = np.uint32((0xFFFFFFFF << (32 - first_bit)) & 0xFFFFFFFF) if first_bit else np.uint32(0)
    return bool(np.array_equal(a[1:], b[1:]) and (a[0] & mask) == (b[0] & mask))


@need_ref
@pytest.mark.parametrize("lam", [0.1, 0.58, 2.34])
@pytest.mark.parametrize("n", [1, 2, 31, 1000, 150000])
def test_encoder_matches_reference(lam, n):
    L = 1024
    counts, *_ = _ref_demo(lam, L)
    data = np.zeros(n + 8, np.uint8)
    ol.ref().ref_gen_test_data(counts, 256, n, L, 5, data)
    data = data[:n]
    ru, rfb, rfs = _ref_encode(counts, L, data)
    st, _ = ol.encode(data, counts, L)
    assert (rfb, rfs) == (st.first_bit, st.first_state)
    assert _units_equal(ru, st.units, rfb)


@need_ref
@pytest.mark.parametrize("lam", [0.1, 1.0, 2.34])
@pytest.mark.parametrize("S,T", [(4, 1), (4, 8), (8, 3), (32, 16)])
@pytest.mark.parametrize("L", [1024, 4096])
def test_multicore_matches_reference(lam, S, T, L):
    counts, *_ = _ref_demo(lam, L)
    n = 120000
    data = np.zeros(n + 8, np.uint8)
    ol.ref().ref_gen_test_data(counts, 256, n, L, 5, data)
    st, _ = ol.encode(data[:n], counts, L)
    ours = ol.decode_multicore(st, S=S, T=T)
    ref = np.zeros(n, np.uint8)
    t = ol.ref().ref_multicore(S, T, np.concatenate([st.units, np.zeros(S + 8, np.uint32)]),
                               len(st.units), st.table, L, st.first_state, st.first_bit, ref, n)
    assert t >= 0
    assert np.array_equal(ours, ref)


def _q1_stream(S):
    """A stream whose zero-bit trailing symbols fall at unit n_units with n_units % S == 0:
    symbol 7 takes > L/2 of the states, so the first encoded symbols cost 0 bits."""
    L = 1024
    counts = np.ones(256, np.uint32)
    counts[7] = L - 255
    for seed in range(200):
        rng = np.random.default_rng(seed)
        n = 5000 + seed
        data = np.where(rng.random(n) < 0.9, 7, rng.integers(0, 256, n)).astype(np.uint8)
        data[:3] = 7
        st, _ = ol.encode(data, counts, L)
        if len(st.units) % S == 0:
            return st, data, counts
    raise AssertionError("no Q1 stream found")


@need_ref
def test_multicore_drops_zero_bit_tail_like_reference():
    """Where the reference loses trailing zero-bit symbols (n_units % S == 0), the
    restatement loses exactly the same ones, and the sequential oracle does not."""
    S, T = 4, 4
    st, data, counts = _q1_stream(S)
    n = len(data)
    seq, *_ = ol.decode_seq(st)
    assert np.array_equal(seq[::-1], data)
    ours = ol.decode_multicore(st, S=S, T=T)
    ref = np.zeros(n, np.uint8)
    ol.ref().ref_multicore(S, T, np.concatenate([st.units, np.zeros(S + 8, np.uint32)]),
                           len(st.units), st.table, st.L, st.first_state, st.first_bit, ref, n)
    assert np.array_equal(ours, ref)
    assert not np.array_equal(ref, seq)          # the reference really drops symbols
    diff = np.nonzero(ref != seq)[0]
    assert diff.min() >= n - 8 and np.all(ref[diff] == 0)


@pytest.mark.parametrize("lam,L", [(0.002, 4096), (0.1, 1024), (1.0, 4096), (3.0, 32),
                                   (3.0, 4096), (0.05, 512), (1.5, 16384)])
def test_roundtrip_and_end_invariant(lam, L):
    sym = ol.gen_symbols(9, lam, 200000)
    st, _ = ol.encode(sym, L=L)
    out, _, _, es, eb = ol.decode_seq(st)
    assert np.array_equal(out[::-1], sym)
    assert es == L                       # encoder starts at state index 0
    assert eb == 32 * len(st.units)      # every bit consumed


def test_entries_counts_bruteforce():
    """Entries/counts per subsequence re-derived by an independent pure-Python walk."""
    L = 64
    sym = ol.gen_symbols(3, 0.4, 700)
    st, _ = ol.encode(sym, L=L)
    bits = []
    for u in st.units:
        bits += [(int(u) >> i) & 1 for i in range(32)]
    bits += [0] * 64
    for S in (1, 2, 3, 5):
        out, ent, cnt, _, _ = ol.decode_seq(st, S=S)
        nsub = -(-len(st.units) // S)
        p, x = 32 - st.first_bit, st.first_state
        e_bf, c_bf = [None] * nsub, [0] * nsub
        for _ in range(len(sym)):
            for j in range(nsub):
                if e_bf[j] is None and p >= 32 * S * j:
                    e_bf[j] = (x - L) | ((p - 32 * S * j) << 16)
            c_bf[min(p // (32 * S), nsub - 1)] += 1
            e = int(st.table[x - L])
            nb = e >> 24
            v = sum(bits[p + i] << i for i in range(nb))
            x = ((e & 0xFFFF) << nb) | v
            p += nb
        assert list(ent) == e_bf and list(cnt) == c_bf


def test_subseq_last_bruteforce():
    """Last step of each subsequence re-derived by an independent pure-Python walk, and
    one decoding step from each record (but the last) lands exactly on the next
    subsequence's entry as mo_decode_seq reports it."""
    L = 64
    sym = ol.gen_symbols(5, 0.7, 900)
    st, _ = ol.encode(sym, L=L)
    bits = []
    for u in st.units:
        bits += [(int(u) >> i) & 1 for i in range(32)]
    bits += [0] * 64

    def step(x, p):
        e = int(st.table[x - L])
        nb = e >> 24
        v = sum(bits[p + i] << i for i in range(nb))
        return ((e & 0xFFFF) << nb) | v, p + nb

    for S in (1, 2, 3, 4, 7):
        ls, lb = ol.subseq_last(st, S)
        nsub = -(-len(st.units) // S)
        p, x = 32 - st.first_bit, st.first_state
        s_bf, b_bf = [0] * nsub, [0] * nsub
        for _ in range(len(sym)):
            j = min(p // (32 * S), nsub - 1)
            s_bf[j], b_bf[j] = x, p
            x, p = step(x, p)
        assert list(ls) == s_bf and list(lb) == b_bf
        _, ent, cnt, _, _ = ol.decode_seq(st, S=S)
        for j in range(nsub - 1):
            if cnt[j] == 0 or cnt[j + 1] == 0:
                continue
            x2, p2 = step(int(ls[j]), int(lb[j]))
            assert ent[j + 1] == (x2 - L) | ((p2 - 32 * S * (j + 1)) << 16)
            assert lb[j] < 32 * S * (j + 1) <= p2


@pytest.mark.parametrize("lam", [0.05, 0.5, 2.0])
def test_symbol_generator_closed_form(lam):
    n = 2_000_000
    s = ol.gen_symbols(123, lam, n)
    k = np.arange(256)
    p = np.exp(-lam * k) - np.exp(-lam * (k + 1))
    p[255] = np.exp(-lam * 255)
    f = np.bincount(s, minlength=256) / n
    sd = np.sqrt(p * (1 - p) / n)
    assert np.all(np.abs(f - p) <= 6 * sd + 1e-9)
    # element i depends only on (seed, i)
    assert np.array_equal(ol.gen_symbols(123, lam, 1000), s[:1000])


def test_golden_vectors():
    """Vectors written by tests/golden/make_golden.py from the compiled reference."""
    path = os.path.join(GOLDEN, "reference_vectors.json")
    g = json.load(open(path))
    for case in g["cases"]:
        L = case["L"]
        counts = np.array(case["counts"], np.uint32)
        _, _, dec = ol.build_tables(counts, L)
        assert np.array_equal(dec, np.array(case["dec_table"], np.uint32))
        data = np.frombuffer(bytes.fromhex(case["symbols_hex"]), np.uint8)
        st, _ = ol.encode(data, counts, L)
        assert st.first_bit == case["first_bit"] and st.first_state == case["first_state"]
        assert _units_equal(np.array(case["units"], np.uint32), st.units, st.first_bit)
        mc = ol.decode_multicore(st, S=case["S"], T=case["T"])
        assert mc.tobytes().hex() == case["multicore_hex"]


@pytest.mark.parametrize("seed", range(6))
def test_normalize_invariants(seed):
    """mo_normalize (the build's own normaliser; the reference's breaks on absent
    symbols, src/ans_table_generator.cc:84-85, 274-283): counts sum to L, absent
    symbols get 0, present ones >= 1, exact multiples are reproduced exactly,
    scaling the histogram changes nothing, and every symbol but the one that takes
    the rounding surplus is within one count of its share."""
    rng = np.random.default_rng(seed)
    L = [256, 1024, 4096][seed % 3]
    hist = rng.integers(0, 1 << 20, 256).astype(np.uint64)
    hist[rng.random(256) < 0.3] = 0
    hist[rng.integers(0, 256, 3)] = rng.integers(1, 4, 3)     # rare symbols
    c = ol.normalize(hist, L)
    assert int(c.sum()) == L
    assert np.array_equal(c == 0, hist == 0)
    assert np.array_equal(ol.normalize(hist * np.uint64(3), L), c)
    share = hist.astype(np.float64) * L / float(hist.sum())
    big = int(np.argmax(c))
    others = np.arange(256) != big
    assert np.all(np.abs(c[others] - np.maximum(share[others], (hist[others] > 0))) < 1.0 + 1e-9)
    # exact multiples: counts proportional to the histogram come back unchanged
    exact = np.zeros(256, np.uint64)
    exact[:16] = np.arange(1, 17, dtype=np.uint64)
    c2 = ol.normalize(exact * np.uint64(7), 136)
    assert np.array_equal(c2[:16], np.arange(1, 17)) and int(c2.sum()) == 136
    # a single present symbol takes every state; too many present symbols is an error
    one = np.zeros(256, np.uint64)
    one[200] = 5
    assert ol.normalize(one, L)[200] == L
    with pytest.raises(ValueError):
        ol.normalize(np.ones(256, np.uint64), 128)


def _py_bits(units):
    bits = []
    for u in units:
        bits += [(int(u) >> i) & 1 for i in range(32)]
    return bits + [0] * 64


@pytest.mark.parametrize("lam", [0.3, 2.0])
def test_interleaved_one_slot_is_the_single_format(lam):
    """N = 1: the interleaved encoder and decoder are the single-state ones (pinned against
    the reference's ANSEncoder and the sequential decoder)."""
    sym = ol.gen_symbols(12, lam, 20000)
    st1, counts = ol.encode(sym, L=1024)
    il, _ = ol.encode_interleaved(sym, 1, [counts], 1024)
    assert np.array_equal(il.units, st1.units)
    assert (il.first_bit, int(il.first_states[0])) == (st1.first_bit, st1.first_state)
    out, _, es, eb = ol.decode_interleaved(il)
    ref, _, _, e1, b1 = ol.decode_seq(st1)
    assert np.array_equal(out, ref) and (int(es[0]), eb) == (e1, b1)


@pytest.mark.parametrize("N,shared", [(2, True), (2, False), (3, False)])
def test_interleaved_bruteforce(N, shared):
    """Interleaved encode and decode re-derived by an independent pure-Python encoder
    (codes appended MSB-first, slot i mod N, every slot from L) and decoder (decode order,
    slot (n-1-j) mod N) on tiny streams; round trip; every slot ends at L and the bit
    cursor at 32 n_units; records (every R units) by brute force."""
    L = 64
    sym = ol.gen_symbols(40 + N, 0.5, 700)
    hist = np.bincount(sym, minlength=256).astype(np.uint64)
    c0 = ol.normalize(hist, L)
    # other slots: the same support, the counts of the present symbols rotated
    nz = np.nonzero(hist)[0]
    others = []
    for k in range(1, N):
        h = np.zeros_like(hist)
        h[nz] = np.roll(hist[nz], k)
        others.append(ol.normalize(h, L))
    counts = [c0] * N if shared else [c0] + others
    il, encs = ol.encode_interleaved(sym, N, counts, L)
    # Python encoder
    earr = [np.frombuffer(e, dtype=np.uint32).reshape(256, L, 3) for e in encs]
    x = [L] * N
    bits = []
    for i, s in enumerate(sym):
        k = i % N
        nx, code, ln = (int(v) for v in earr[k][s, x[k] - L])
        bits += [(code >> (ln - 1 - b)) & 1 for b in range(ln)]
        x[k] = nx
    fb = len(bits) % 32
    bits += [0] * ((32 - fb) % 32)
    fwd = [int("".join(map(str, bits[32 * u:32 * u + 32])), 2) for u in range(len(bits) // 32)]
    if fb == 0:
        fwd.append(0)
    assert il.first_bit == fb and list(il.first_states) == x
    assert list(il.units) == fwd[::-1]
    # Python decoder; positions and states before every step (j = n: the end)
    ub = _py_bits(il.units)
    p, xs, out = 32 - il.first_bit, list(il.first_states), []
    trace = []
    for j in range(len(sym)):
        trace.append((p, list(xs)))
        k = (len(sym) - 1 - j) % N
        e = int(il.tables[k][xs[k] - L])
        nb = e >> 24
        out.append((e >> 16) & 0xFF)
        xs[k] = ((e & 0xFFFF) << nb) | sum(ub[p + b] << b for b in range(nb))
        p += nb
    trace.append((p, list(xs)))
    assert out == list(sym[::-1])
    assert xs == [L] * N and p == 32 * len(il.units)
    R = 3 if N <= 2 else 0          # records hold two slots
    got, rec, es, eb = ol.decode_interleaved(il, R=R)
    assert np.array_equal(got, np.array(out, np.uint8))
    assert list(es) == [L] * N and eb == 32 * len(il.units)
    # records anchored at the end: record m at unit n_units - R m, first step at/after it
    nu = len(il.units)
    assert len(rec) == (-(-nu // R) if R else 0)
    for m in range(len(rec)):
        T = 32 * (nu - R * m)
        j = next(j for j, (pj, _) in enumerate(trace) if pj >= T)
        pj, xj = trace[j]
        assert [int(v) for v in rec[m]["state"][:min(N, 2)]] == [v - L for v in xj[:2]]
        assert (int(rec[m]["bit"]), int(rec[m]["before"])) == (pj - T, j)
