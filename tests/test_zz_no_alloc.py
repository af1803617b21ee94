"""mans_decode_noalloc (the decode with the library's static scratch instead of a
workspace) against mans_decode on the same streams: bit-exact, on two calls in a row,
the second with a smaller stream; a stream too large for the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol


@pytest.mark.gpu
def test_noalloc_matches_decode():
    import torch
    from multians_amd import _lib
    from multians_amd.decoder import DeviceStream, encode_symbols
    lib = _lib.lib()
    for lam, n, L in [(0.1, 3_000_000, 4096), (1.0, 700_001, 1024)]:
        print("case noalloc lam=%g n=%d" % (lam, n), flush=True)
        sym = ol.gen_symbols(101, lam, n)
        ds = DeviceStream(encode_symbols(sym, L))
        ref = ds.decode().clone()
        torch.cuda.synchronize()
        assert ds.status() == 0
        out = torch.full((n + 64,), 0xA5, dtype=torch.uint8, device="cuda")
        a = ds.args(flags=_lib.MANS_F_SYNC_ERRORS)
        a.out = out.data_ptr()
        a.workspace = None
        a.workspace_bytes = 0
        _lib.check(lib.mans_decode_noalloc(C.byref(a)))
        torch.cuda.synchronize()
        assert torch.equal(out[:n], ref[:n])
        assert bool((out[n:] == 0xA5).all())              # nothing past n_out
        assert np.array_equal(out[:n].cpu().numpy()[::-1], sym)


@pytest.mark.gpu
def test_noalloc_refuses_what_does_not_fit():
    import torch
    from multians_amd import _lib
    lib = _lib.lib()
    units = torch.zeros(64, dtype=torch.int32, device="cuda")
    out = torch.zeros(64, dtype=torch.uint8, device="cuda")
    table = torch.zeros(4096, dtype=torch.int32, device="cuda")
    a = _lib.DecodeArgs()
    a.units, a.n_units = units.data_ptr(), 400_000_000      # about 400 MB of workspace
    a.out, a.n_out = out.data_ptr(), 64
    a.table, a.L = table.data_ptr(), 4096
    a.first_state, a.first_bit = 4096, 32
    assert lib.mans_decode_noalloc(C.byref(a)) == _lib.MANS_ENOMEM
    assert b"static scratch" in lib.mans_last_error()
    a.n_units = 64
    a.workspace, a.workspace_bytes = units.data_ptr(), 256
    assert lib.mans_decode_noalloc(C.byref(a)) == _lib.MANS_EINVAL
