"""The device encoder's table rule (multians_amd/csrc/mans_encode.hip, stage_enc and
enc_step), restated in numpy from a decoder table, against the host codec's full
encoder table (mans_build_tables, pinned against the reference's ANSTableGenerator in
test_oracle.py): for every symbol s and state x in [L, 2L), the bits emitted
nb = m_s - (x < f_s << m_s), m_s = log2 L - floor(log2 f_s), the code x mod 2^nb, and the
next state (the state y of s whose decoder entry has next = x >> nb).  Bar: exact.
"""
import numpy as np
import pytest

import oracle_lib as ol


def _device_rule(dec, L):
    """(next_state, code, nbits) for every (symbol, x - L), from the decoder table."""
    R = L.bit_length() - 1
    sym = (dec >> 16) & 0xFF
    nxt = dec & 0xFFFF
    f = np.bincount(sym, minlength=256)
    base = np.concatenate([[0], np.cumsum(f)[:-1]])
    ys = np.full(L, -1, np.int64)
    for y in range(L):
        s, i = sym[y], nxt[y]
        assert f[s] <= i < 2 * f[s]
        ys[base[s] + i - f[s]] = y
    assert (ys >= 0).all()                    # every (symbol, next) pair exactly once
    x = np.arange(L, 2 * L, dtype=np.int64)
    out = np.zeros((256, L, 3), np.int64)
    for s in np.nonzero(f)[0]:
        m = R - (int(f[s]).bit_length() - 1)
        nb = m - (x < (int(f[s]) << m)).astype(np.int64)
        out[s, :, 0] = L + ys[base[s] + (x >> nb) - f[s]]
        out[s, :, 1] = x & ((1 << nb) - 1)
        out[s, :, 2] = nb
    return out, f


@pytest.mark.parametrize("lam", [0.002, 0.1, 1.0, 3.0])
@pytest.mark.parametrize("L", [64, 1024, 4096, 16384])
def test_encoder_rule_matches_host_table(lam, L):
    from multians_amd import _lib
    sym = ol.gen_symbols(17, lam, 200_000)
    if L == 64:
        sym = np.minimum(sym, 40).astype(np.uint8)
    counts = _lib.normalize_counts(_lib.histogram(sym), L)
    enc, dec = _lib.build_tables(counts, L)
    enc = enc.reshape(256, L, 3).astype(np.int64)
    rule, f = _device_rule(dec, L)
    present = f > 0
    assert np.array_equal(rule[present], enc[present])


def test_encoder_rule_one_symbol():
    """A symbol owning every state: m = 0, no bits, next = x itself."""
    from multians_amd import _lib
    L = 1024
    counts = np.zeros(256, np.uint32)
    counts[9] = L
    enc, dec = _lib.build_tables(counts, L)
    rule, _ = _device_rule(dec, L)
    assert (rule[9, :, 2] == 0).all()
    assert np.array_equal(rule[9], enc.reshape(256, L, 3)[9].astype(np.int64))
