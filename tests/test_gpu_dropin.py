"""The reference's own demo program (src/main.cc) on the drop-in layer.

oracle/_ref/ref_demo_dropin is src/main.cc compiled unchanged against include/ and
linked with libcuhd_amd.so + libmultians_amd.so (oracle/Makefile, target `dropin`).
For lambda = 0.1, 0.26, ..., 2.34 it builds the reference demo's distribution, tables
and stream (src/main.cc:55-80), decodes with cuhd::CUHDGPUDecoder::decode, copies back,
reverses, and prints "mismatch" when the output differs from its input
(src/main.cc:123-128).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_demo_dropin")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="reference demo not built (needs the reference sources)")
@pytest.mark.parametrize("mb", [1, 16])
def test_reference_demo_roundtrip_on_dropin(mb):
    print("reference demo, %d MB" % mb, flush=True)
    r = subprocess.run([EXE, "0", str(mb), "1"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0
    assert "mismatch" not in r.stdout and "error" not in r.stdout
    rows = [l for l in r.stdout.splitlines()[1:] if l.strip()]
    assert len(rows) == 15            # lambda = 0.1 + 0.16 k < 2.5
