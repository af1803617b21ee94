"""GPU parity of lineage decodes (mans_decode with a lineage rule, MANS_RULE_LINEAGE)
against the oracle: bit-exact output, lineage records that hold (no fallback), and the
fallback to the chunked pipeline (forced with a one-checkpoint lineage cap) giving the same
output.  The CPU emulation of the same algorithm is tests/test_lineage.py.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture
def lineage_rule():
    from multians_amd import _lib

    def use(D, L_max=16384):
        _lib.set_geometry_rules([(L_max, 1e30, 1 << 40, D, 0, _lib.RULE_LINEAGE)])
    yield use
    _lib.set_geometry_rules([])
    _lib.check(_lib.lib().mans_set_lineage_cap(0))


def _decode(sym, L):
    import torch
    from multians_amd.decoder import DeviceStream, encode_symbols
    es = encode_symbols(sym, L)
    ds = DeviceStream(es)
    out = ds.decode()
    torch.cuda.synchronize()
    got = out[:len(sym)].cpu().numpy()
    return es, ds, got


@pytest.mark.parametrize("lam", [0.002, 0.05, 0.2, 1.0, 3.0])
@pytest.mark.parametrize("D", [64, 256])
def test_lineage_matches_oracle(lineage_rule, lam, D):
    print(f"case lam={lam} D={D}", flush=True)
    lineage_rule(D)
    sym = ol.gen_symbols(71, lam, 3_000_000)
    es, ds, got = _decode(sym, 4096)
    fell_back, reason, n_lin, n_ck = ds.lineage_stats()
    assert ds.status() == 0
    assert fell_back == 0 and reason == 0, (fell_back, reason)
    assert n_lin > 0 and n_ck > 0
    st = ol.Stream(es.units, es.first_bit, es.first_state, es.table, 4096, len(sym))
    ref, _, _, _, _ = ol.decode_seq(st)
    assert np.array_equal(got, ref)
    assert np.array_equal(got[::-1], sym)


@pytest.mark.parametrize("n", [40, 1000, 4095, 65537, 1_000_003])
@pytest.mark.parametrize("L", [64, 1024])
def test_lineage_small_streams(lineage_rule, n, L):
    print(f"case n={n} L={L}", flush=True)
    lineage_rule(32)
    sym = ol.gen_symbols(500 + n, 0.4, n)
    es, ds, got = _decode(sym, L)
    assert ds.status() == 0
    assert np.array_equal(got[::-1], sym)


def test_lineage_zero_bit_tail(lineage_rule):
    print("case zero-bit tail", flush=True)
    lineage_rule(64)
    for seed in range(4):
        sym = ol.gen_symbols(seed, 3.0, 200_000 + 1009 * seed)
        es, ds, got = _decode(sym, 4096)
        assert ds.status() == 0 and ds.lineage_stats()[0] == 0
        assert np.array_equal(got[::-1], sym)


def test_lineage_fallback_runs_chunked_pipeline(lineage_rule):
    """A one-checkpoint cap makes every lineage give up: the chunked pipeline enqueued
    behind the lineage kernels runs (gated on the device) and decodes the same bytes."""
    from multians_amd import _lib
    print("case fallback", flush=True)
    lineage_rule(128)
    _lib.check(_lib.lib().mans_set_lineage_cap(1))
    sym = ol.gen_symbols(17, 0.2, 2_000_000)
    es, ds, got = _decode(sym, 4096)
    fell_back, reason, _, _ = ds.lineage_stats()
    assert fell_back == 1 and (reason & 2)
    assert ds.status() == 0
    assert np.array_equal(got[::-1], sym)
    # and back: the default cap decodes by lineages again, same bytes
    _lib.check(_lib.lib().mans_set_lineage_cap(0))
    out = ds.decode()
    import torch
    torch.cuda.synchronize()
    assert ds.lineage_stats()[0] == 0
    assert np.array_equal(out[:len(sym)].cpu().numpy()[::-1], sym)


def test_lineage_repeat_is_stable(lineage_rule):
    """Repeated decodes on one workspace (stale slot values from the previous decode must
    never match: k_lin_init rewrites them) give the same bytes every time."""
    import torch
    print("case repeat", flush=True)
    lineage_rule(96)
    sym = ol.gen_symbols(23, 1.0, 2_500_000)
    es, ds, got = _decode(sym, 4096)
    for _ in range(4):
        out = ds.decode()
        torch.cuda.synchronize()
        assert ds.lineage_stats()[0] == 0
        assert np.array_equal(out[:len(sym)].cpu().numpy(), got)
    assert np.array_equal(got[::-1], sym)
