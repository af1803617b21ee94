"""Stress: one stream of more than 2^32 uint8 symbols (64-bit output offsets), 4096
states, lambda = 3.0 (skewed: symbol 0 holds 95% of the mass, so the first-encoded
symbols decode as zero-bit steps after the last unit).  The stream is made on the device
(symbols, table and units: the host generator's and host codec's bytes); the decode is
compared on the device with the original symbols, and a prefix with the oracle."""
import numpy as np
import pytest

import oracle_lib as ol


@pytest.mark.gpu
def test_stream_beyond_4gib_symbols():
    import torch
    from multians_amd.decoder import GeneratedBatch, workspace_bytes
    n = (1 << 32) + 4097
    print("case stress n=%d lambda=3.0" % n, flush=True)
    gb = GeneratedBatch([(91, 3.0, n, 4096)])
    print("encoded: %d units" % gb.n_units[0], flush=True)
    out = torch.empty(n, dtype=torch.uint8, device="cuda")
    ws = torch.zeros(workspace_bytes(int(gb.n_units[0]), 4096), dtype=torch.uint8, device="cuda")
    gb.decode_one(0, out, ws)
    torch.cuda.synchronize()
    assert gb.status(ws) == 0
    assert torch.equal(out, gb.symbols(0).flip(0))
    es = gb.encoded(0)
    head = 1 << 24
    ref, *_ = ol.decode_seq(ol.Stream(es.units, es.first_bit, es.first_state, es.table, 4096, n),
                            n_out=head)
    assert np.array_equal(ref, out[:head].cpu().numpy())
