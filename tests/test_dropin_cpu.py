"""CPU checks of the drop-in C++ layer and the C ABI (no GPU calls).

* libmultians_amd.so loads and exports every function include/mans_c_api.h declares.
* The drop-in ANSTableGenerator / ANSEncoder (libcuhd_amd.so, driven by mans_demo, a
  program shaped like the reference's src/main.cc) produce the reference's counts,
  decoder tables, test data and encoded units for every demo lambda (pinned against
  the reference's own code compiled into oracle/_ref).
* The product host codec (mans_* C functions) equals the oracle's.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "multians_amd", "lib")


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mans_[a-z0-9_]+)\s*\(", src)))


def test_c_abi_exports_everything_declared():
    from multians_amd import _lib
    lib = _lib.lib()
    names = _declared("mans_c_api.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from multians_amd import _lib
    assert sorted(_lib.EXPORTS) == names


def test_c_abi_errors_without_gpu():
    from multians_amd import _lib
    L = _lib.lib()
    assert L.mans_decode(None) == _lib.MANS_EINVAL
    assert b"NULL" in L.mans_last_error()
    assert L.mans_error_string(_lib.MANS_ETABLE) == b"decoder table rejected"
    # the state count must be a power of two (argument checks come before any GPU call)
    units = np.zeros(16, np.uint32)
    table = np.zeros(1000, np.uint32)
    out = np.zeros(16, np.uint8)
    args = _lib.DecodeArgs()
    args.units = units.ctypes.data
    args.n_units = 16
    args.out = out.ctypes.data
    args.n_out = 16
    args.table = table.ctypes.data
    args.L = 1000
    assert L.mans_decode(C.byref(args)) == _lib.MANS_EINVAL
    assert b"power of two" in L.mans_last_error()


def _fnv(b):
    h = 1469598103934665603
    for x in b:
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref not built")
def test_dropin_codec_matches_reference():
    out = subprocess.run([os.path.join(LIB, "mans_demo"), "0", "0", "--no-gpu", "--checksums"],
                         check=True, stdout=subprocess.PIPE, text=True).stdout
    rows = [l.split() for l in out.splitlines() if l.startswith("CHK")]
    assert len(rows) == 15
    R = ol.ref()
    n, L = 100000, 1024
    for r in rows:
        kv = dict(x.split("=") for x in r[2:])
        lam = float(np.float32(r[1]))
        counts = np.zeros(256, np.uint32)
        dec = np.zeros(L, np.uint32)
        en, ec, el = (np.zeros(256 * L, np.uint32) for _ in range(3))
        R.ref_demo_tables(lam, 256, L, 5, counts, dec, en, ec, el)
        assert int(kv["counts"]) == _fnv(counts.tobytes())
        assert int(kv["dec"]) == _fnv(dec.tobytes())
        data = np.zeros(n + 8, np.uint8)
        R.ref_gen_test_data(counts, 256, n, L, 5, data)
        assert int(kv["data"]) == _fnv(data[:n].tobytes())
        units = np.zeros(n, np.uint32)
        fb, fs = C.c_uint32(), C.c_uint32()
        nu = R.ref_encode(counts, 256, L, data, n, units, n, C.byref(fb), C.byref(fs))
        units = units[:nu].copy()
        units[0] &= np.uint32((0xFFFFFFFF << (32 - fb.value)) & 0xFFFFFFFF) if fb.value else 0
        assert (int(kv["nunits"]), int(kv["first_bit"]), int(kv["first_state"])) == (nu, fb.value, fs.value)
        assert int(kv["units"]) == _fnv(units.tobytes())


@pytest.mark.parametrize("lam,L", [(0.002, 4096), (0.1, 1024), (1.0, 4096), (3.0, 256)])
def test_product_codec_matches_oracle(lam, L):
    from multians_amd import _lib
    from multians_amd.decoder import encode_symbols
    sym = ol.gen_symbols(77, lam, 400000)
    assert np.array_equal(_lib.gen_symbols(77, lam, 400000), sym)
    es = encode_symbols(sym, L)
    st, counts = ol.encode(sym, L=L)
    assert np.array_equal(es.counts, counts)
    assert np.array_equal(es.table, st.table)
    assert np.array_equal(es.units, st.units)
    assert (es.first_bit, es.first_state) == (st.first_bit, st.first_state)
    enc, _ = _lib.build_tables(counts, L)
    _, encarr, _ = ol.build_tables(counts, L)
    assert np.array_equal(enc.reshape(256, L, 3), encarr)


def test_geometry_rules_pick_chunk_length():
    """Autotuner rules: first matching rule wins, lengths round down to a power of two,
    short lengths are refused, and clearing the rules restores the built-in choice."""
    from multians_amd import _lib
    builtin = _lib.chunk_units(1 << 20, 4 << 20, 4096)
    try:
        _lib.set_geometry_rules([(1024, 100.0, 1 << 30, 300),
                                 (4096, 5.0, 1 << 20, 1000),
                                 (4096, 100.0, 1 << 30, 4096)])
        assert _lib.chunk_units(1 << 20, 4 << 20, 4096) == 512      # second rule
        assert _lib.chunk_units(1 << 20, 4 << 20, 1024) == 256      # first rule
        assert _lib.chunk_units(1 << 20, 20 << 20, 4096) == 4096    # spu 20 -> third
        assert _lib.chunk_units((1 << 20) + 1, 4 << 20, 4096) == 4096
        with pytest.raises(_lib.MansError):
            _lib.set_geometry_rules([(4096, 5.0, 1 << 20, 16)])
        assert _lib.chunk_units(1 << 20, 4 << 20, 4096) == 512      # unchanged on error
    finally:
        _lib.set_geometry_rules([])
    assert _lib.chunk_units(1 << 20, 4 << 20, 4096) == builtin


def test_tuner_rule_table_matches_builtin_choice():
    """The library's built-in geometry is the autotuner's table: loading the rules recorded
    in multians_amd/tuning/gfx950.json picks the same chunk length for every stream shape
    (the batch scope, which mans_chunk_units reports), the batch rules are the tuner's
    shape rebuilt from the recorded parameters, and shapes no rule covers fall back to the
    same heuristic."""
    import json
    from multians_amd import _lib, tune
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "multians_amd", "tuning", "gfx950.json")
    meta = json.load(open(path))
    rules = [tuple(r[k] for k in ("L_max", "spu_max", "units_max", "chunk_units", "warm_units",
                                  "scope")) for r in meta["rules"]]
    batch = [r for r in rules if r[5] == _lib.RULE_BATCH]
    assert batch == tune.batch_rules(meta["floor"], meta["safe"], meta["warm1"], meta["warm2"])
    assert all(r[5] in (_lib.RULE_SINGLE, _lib.RULE_BATCH, _lib.RULE_LINEAGE) for r in rules)
    # lineage rules: a spacing (a multiple of 32) or 0 (the chunked pipeline) per cell
    assert all(r[3] % 32 == 0 for r in rules if r[5] == _lib.RULE_LINEAGE)
    rng = np.random.default_rng(4)
    shapes = [(int(np.exp(rng.uniform(0, np.log(1 << 32)))), float(np.exp(rng.uniform(0, 5))),
               int(rng.choice([256, 1024, 2048, 4096, 8192]))) for _ in range(3000)]
    builtin = [_lib.chunk_units(n, int(n * spu), L) for n, spu, L in shapes]
    builtin_lin = [_lib.lineage_spacing(n, int(n * spu), L) for n, spu, L in shapes]
    try:
        _lib.set_geometry_rules(rules)
        tuned = [_lib.chunk_units(n, int(n * spu), L) for n, spu, L in shapes]
        tuned_lin = [_lib.lineage_spacing(n, int(n * spu), L) for n, spu, L in shapes]
    finally:
        _lib.set_geometry_rules([])
    assert tuned == builtin
    assert tuned_lin == builtin_lin
    assert any(builtin_lin) and not all(builtin_lin)


def test_histogram_matches_bincount():
    from multians_amd import _lib
    rng = np.random.default_rng(8)
    for n in (0, 1, 1000, (1 << 22) + 77):
        sym = rng.integers(0, 256, n).astype(np.uint8)
        ref = np.bincount(sym, minlength=256).astype(np.uint64)
        for t in (1, 5):
            assert np.array_equal(_lib.histogram(sym, threads=t), ref)
