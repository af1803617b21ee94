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


DEMO = os.path.join(ROOT, "multians_amd", "lib", "mans_demo")


@pytest.mark.gpu
def test_decoder_memory_holds_reference_sync_state(tmp_path):
    """The kept cuhd::CUHDGPUDecoderMemory accessors return what the reference's decoder
    leaves there (src/multians_gpu_decoder.cu:335-353, 659-670): per subsequence of S = 4
    units the unit, bit and state of its last step and the symbols before it
    (get_sync_info), the same scan (get_output_sizes), and a set flag per block
    (retrieve_sync_data + get_sequence_synced_host).  The drop-in demo (mans_demo, the
    reference demo's flow on the drop-in classes) dumps them after its decode; they are
    checked against the oracle's entries, counts and last steps (mo_decode_seq,
    mo_subseq_last)."""
    import numpy as np
    import oracle_lib as ol
    print("case decoder memory sync state", flush=True)
    r = subprocess.run([DEMO, "0", "1", "--dump-sync", str(tmp_path)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "mismatch" not in r.stdout
    for k in range(2):
        raw = (tmp_path / ("sync_%d.bin" % k)).read_bytes()
        hdr = np.frombuffer(raw[:72], np.uint64)
        nu, fb, fs, L, n, S, threads, nsub, nblk = (int(v) for v in hdr)
        o = 72
        units = np.frombuffer(raw[o:o + 4 * nu], np.uint32).copy(); o += 4 * nu
        table = np.frombuffer(raw[o:o + 4 * L], np.uint32).copy(); o += 4 * L
        sync = np.frombuffer(raw[o:o + 16 * nsub], np.uint32).reshape(nsub, 4); o += 16 * nsub
        sizes = np.frombuffer(raw[o:o + 4 * nsub], np.uint32); o += 4 * nsub
        flags = np.frombuffer(raw[o:o + nblk], np.uint8)
        assert nsub == -(-nu // S) and nblk == -(-nsub // threads)
        st = ol.Stream(units, fb, fs, table, L, n)
        _, _, cnt, _, _ = ol.decode_seq(st, S=S)
        ls, lb = ol.subseq_last(st, S)
        scan = np.concatenate([np.zeros(1, np.uint64),
                               np.cumsum(cnt.astype(np.uint64))[:-1]]) & np.uint64(0xFFFFFFFF)
        own = cnt > 0
        assert own.sum() >= nsub - 1
        assert np.array_equal(sizes, scan.astype(np.uint32))
        assert np.array_equal(sync[own, 2], scan[own].astype(np.uint32))
        j = np.nonzero(own)[0]
        assert np.array_equal(sync[own, 3], ls[own])
        assert np.array_equal(sync[own, 0].astype(np.uint64),
                              (lb[own] >> np.uint64(5)) - j.astype(np.uint64) * np.uint64(S))
        assert np.array_equal(sync[own, 1].astype(np.uint64), lb[own] & np.uint64(31))
        assert flags.all()
