"""GPU parity at the full sizes of the benchmark configurations.

- single stream, 2^30 uint8 symbols, 4096 states, lambda in {0.01, 0.05, 0.2, 1.0};
- the batch of 10,000 streams (sizes log-uniform 4 KiB-16 MiB, lambda ~ LogU(0.005, 2.0),
  a 4096-state table each), the benchmark's own stream list;
- stress: one stream of 2^32 symbols at lambda = 0.002 (symbol 255 holds 60% of the
  mass, the first-encoded symbols decode as zero-bit steps after the last unit).

Streams this large are made on the device (symbols, tables and units: the host
generator's and host codec's bytes, see tests/test_gpu_tables.py and
tests/test_gpu_encode.py).  Every decoded stream is compared on the device with its
original symbols (reversed: the decoder emits in decode order, src/main.cc:123).  The
oracle (mo_decode_seq) checks a prefix of each large stream from its true start and
windows that start at the sync points the decode wrote (so those records are checked
too); sampled batch streams are checked whole against mo_decode_seq and against the
reference's MulticoreDecoder (src/multicore_decoder.cc, built into oracle/_ref when the
reference was present, else the oracle's restatement of it).  Bar: bit-exact.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

SINGLE_N = 1 << 30


def _oracle_windows(es, index, chunk_units, gpu_out, rng, n_windows=8, win=1 << 20,
                    prefix=1 << 23):
    """The oracle's decode of the stream's first `prefix` symbols, and of windows that start
    at random sync points, against the device output (host copy)."""
    n = es.n_symbols
    head = min(prefix, n)
    ref, *_ = ol.decode_seq(ol.Stream(es.units, es.first_bit, es.first_state, es.table, es.L, n),
                            n_out=head)
    assert np.array_equal(ref, gpu_out[:head]), "prefix"
    for c in rng.integers(1, len(index), n_windows):
        rec = index[int(c)]
        before = int(rec["symbols_before"])
        m = min(win, n - before)
        if m <= 0:
            continue
        sub = ol.Stream(es.units[int(c) * chunk_units:], 32 - int(rec["bit"]),
                        int(rec["state"]) + es.L, es.table, es.L, m)
        got, *_ = ol.decode_seq(sub)
        assert np.array_equal(got, gpu_out[before:before + m]), "window at chunk %d" % c


def _decode_large(gb, i, seed):
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import workspace_bytes
    n = int(gb.n_symbols[i])
    nu = int(gb.n_units[i])
    S = _lib.chunk_units(nu, n, int(gb.L[i]))
    out = torch.empty(n, dtype=torch.uint8, device="cuda")
    ws = torch.zeros(workspace_bytes(nu, int(gb.L[i]), S, 16), dtype=torch.uint8, device="cuda")
    idx = torch.zeros(16 * _lib.sync_point_count(nu, S), dtype=torch.uint8, device="cuda")
    gb.decode_one(i, out, ws, chunk_units=S, index_out=idx)
    torch.cuda.synchronize()
    assert gb.status(ws) == 0
    assert torch.equal(out, gb.symbols(i).flip(0)), "stream differs from its symbols"
    index = idx.cpu().numpy().view(_lib.SYNC_POINT_DTYPE)
    es = gb.encoded(i)
    _oracle_windows(es, index, S, out.cpu().numpy(), np.random.default_rng(seed))


@pytest.mark.parametrize("lam", [0.01, 0.05, 0.2, 1.0])
def test_single_stream_full_size(lam):
    import torch
    from multians_amd.decoder import GeneratedBatch
    print(f"case single n=2^30 lambda={lam}", flush=True)
    gb = GeneratedBatch([(1000 + int(lam * 1000), lam, SINGLE_N, 4096)])
    print("units %d" % gb.n_units[0], flush=True)
    _decode_large(gb, 0, 7)
    del gb
    torch.cuda.empty_cache()


def test_stress_4gib_near_uniform():
    """lambda = 0.002 at 2^32 symbols: 64-bit offsets, the longest streams, and zero-bit
    trailing steps of symbol 255."""
    import torch
    from multians_amd.decoder import GeneratedBatch
    print("case stress n=2^32 lambda=0.002", flush=True)
    gb = GeneratedBatch([(4000, 0.002, 1 << 32, 4096)])
    print("units %d" % gb.n_units[0], flush=True)
    _decode_large(gb, 0, 8)
    del gb
    torch.cuda.empty_cache()


def test_batch_10k_full_size():
    """The benchmark's 10,000-stream batch in one mans_decode_batch call (split into its
    concurrent parts), every stream against its symbols; 64 sampled streams against the
    oracle and the multicore decoder."""
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import GeneratedBatch
    import bench
    print("case batch 10k", flush=True)
    specs = bench.batch_specs(10000)
    gb = GeneratedBatch(specs)
    ooff = np.concatenate([[0], np.cumsum(gb.n_symbols)])
    out = torch.empty(int(ooff[-1]), dtype=torch.uint8, device="cuda")
    descs = gb.descs(out, ooff)
    b = C_u64()
    _lib.check(_lib.lib().mans_batch_workspace_bytes(descs, gb.n, _byref(b)))
    ws = torch.zeros(b.value, dtype=torch.uint8, device="cuda")
    status = torch.full((gb.n,), -1, dtype=torch.int32, device="cuda")
    o = _lib.BatchOpts()
    o.workspace, o.workspace_bytes = ws.data_ptr(), ws.numel()
    o.stream = torch.cuda.current_stream().cuda_stream
    o.stream_status = status.data_ptr()
    _lib.check(_lib.lib().mans_decode_batch(descs, gb.n, _byref(o), None))
    torch.cuda.synchronize()
    assert int(status.abs().sum()) == 0
    print("decoded %d symbols" % int(ooff[-1]), flush=True)
    bad = [i for i in range(gb.n)
           if not torch.equal(out[int(ooff[i]):int(ooff[i + 1])], gb.symbols(i).flip(0))]
    assert not bad, "streams %s" % bad[:10]
    rng = np.random.default_rng(12)
    T = 4
    for i in sorted(rng.choice(gb.n, 64, replace=False)):
        es = gb.encoded(int(i))
        got = out[int(ooff[i]):int(ooff[i + 1])].cpu().numpy()
        st = ol.Stream(es.units, es.first_bit, es.first_state, es.table, es.L, es.n_symbols)
        ref, *_ = ol.decode_seq(st)
        assert np.array_equal(got, ref), "stream %d vs oracle" % i
        # the reference's multicore decoder is correct when every thread's range is long
        # against the self-synchronisation distance (see bench.cpu_baseline)
        if es.n_units >= 8192 * T:
            units = np.concatenate([es.units, np.zeros(16, np.uint32)])
            mc = np.zeros(es.n_symbols, np.uint8)
            if ol.ref_available():
                ol.ref().ref_multicore(4, T, units, es.n_units, es.table, es.L, es.first_state,
                                       es.first_bit, mc, es.n_symbols)
            else:
                mc = ol.decode_multicore(st, S=4, T=T)
            diff = np.nonzero(mc != got)[0]
            # the reference drops trailing zero-bit symbols when n_units % 4 == 0 (they
            # are the last ones in decode order); nothing else may differ
            assert len(diff) == 0 or diff.min() >= es.n_symbols - 64, "stream %d vs multicore" % i


def C_u64():
    import ctypes
    return ctypes.c_uint64()


def _byref(x):
    import ctypes
    return ctypes.byref(x)
