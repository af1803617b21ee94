"""Batch container (include/mans_container.h) on the CPU: exports, pack/open round trip,
the stream descriptors it gives the batch decoder, decoding every stored stream from
the image (emulation build, stored sync points), and rejection of damaged images.
Bar: bit-exact; every damaged image is refused with MANS_ECORRUPT.
"""
import ctypes as C
import os
import re
import struct

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
ECORRUPT = -5


def _truth(units, fb, fs, table, L, n, S):
    from multians_amd import _lib
    st = ol.Stream(units, fb, fs, table, L, n)
    _, entries, counts, _, _ = ol.decode_seq(st, S=S)
    rec = np.zeros(len(entries), _lib.SYNC_POINT_DTYPE)
    rec["state"] = entries & 0xFFFF
    rec["bit"] = (entries >> 16) & 63
    rec["symbols_before"] = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]])
    return rec


@pytest.fixture(scope="module")
def streams():
    """12 streams (two share a table, one is empty) with sync points on every other."""
    from multians_amd import _lib
    from multians_amd.decoder import EncodedStream
    out = []
    for i in range(12):
        n = [0, 1, 5000, 70_000, 123_457, 40_000, 300_000, 9_999, 64, 200_000, 31, 77_777][i]
        lam = [0.5, 0.5, 0.01, 0.2, 1.0, 3.0, 0.05, 0.5, 0.5, 0.3, 2.0, 0.1][i]
        L = 1024 if i % 3 == 0 else 4096
        if n == 0:
            sym = np.zeros(0, np.uint8)
            es = EncodedStream(np.zeros(0, np.uint32), 0, L, np.full(L, (1 << 24), np.uint32)
                               | (np.arange(L, dtype=np.uint32) + L), L, 0)
        else:
            units, fb, fs, dec, sym = _lib.make_stream(900 + i, lam, n, L, True)
            es = EncodedStream(units, fb, fs, dec, L, n)
        out.append((es, sym))
    # stream 7: the first 3,000 symbols of stream 2 under stream 2's table (same bytes)
    es2, sym2 = out[2]
    enc, _ = _lib.build_tables(_lib.normalize_counts(_lib.histogram(sym2), es2.L), es2.L)
    units, fb, fs = _lib.encode(sym2[:3000], enc, es2.L)
    out[7] = (EncodedStream(units, fb, fs, es2.table, es2.L, 3000), sym2[:3000].copy())
    return out


def _sync(streams):
    sps, cus = [], []
    for i, (es, _) in enumerate(streams):
        if i % 2 == 0 or es.n_symbols == 0:
            sps.append(None)
            cus.append(0)
            continue
        S = [16, 64, 256, 1024][i % 4]
        sps.append(_truth(es.units, es.first_bit, es.first_state, es.table, es.L, es.n_symbols, S))
        cus.append(S)
    return sps, cus


def test_container_exports_everything_declared():
    from multians_amd import container
    src = open(os.path.join(ROOT, "include", "mans_container.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(mans_[a-z0-9_]+)\s*\(", src)))
    from multians_amd import _lib
    lib = _lib.lib()
    assert [n for n in names if not hasattr(lib, n)] == []
    assert sorted(container.EXPORTS) == names


def test_pack_open_descs(streams):
    from multians_amd import container
    sps, cus = _sync(streams)
    img = container.pack([e for e, _ in streams], sps, cus)
    c = container.Container(img)
    assert c.info.n_streams == 12
    assert c.info.n_tables == 11                      # streams 2 and 7 share one table
    assert c.info.n_with_sync_points == sum(s is not None for s in sps)
    assert c.info.total_symbols == sum(e.n_symbols for e, _ in streams)
    assert c.info.image_bytes == len(img)
    for i in range(12):
        got = c.sync_points(i)
        if sps[i] is None:
            assert got is None
        else:
            assert np.array_equal(got, sps[i])
    base, obase = 1 << 32, 1 << 40                    # device addresses (never dereferenced)
    for use in (True, False):
        ds, offs = c.descs(base, obase, c.info.total_symbols, use)
        for i, (es, _) in enumerate(streams):
            d = ds[i]
            uo = d.units - base
            assert uo % 256 == 0
            assert np.array_equal(img[uo:uo + 4 * es.n_units].view(np.uint32), es.units)
            to = d.table - base
            assert np.array_equal(img[to:to + 4 * es.L].view(np.uint32), es.table)
            assert (d.n_units, d.n_out, d.L) == (es.n_units, es.n_symbols, es.L)
            assert (d.first_state, d.first_bit) == (es.first_state, es.first_bit)
            assert (d.out or 0) == obase + int(offs[i])
            assert not d.index_out
            if use and sps[i] is not None:
                assert d.chunk_units == cus[i]
                so = d.index_in - base
                assert np.array_equal(img[so:so + 16 * len(sps[i])].view(sps[i].dtype), sps[i])
            else:
                assert not d.index_in and d.chunk_units == 0
        assert int(offs[-1]) == c.info.total_symbols


def test_decode_from_image(streams):
    """Every stream decoded (emulation build) from the image's units, table and sync
    points equals its symbols."""
    from multians_amd import container
    emu = C.CDLL(os.path.join(ROOT, "multians_amd", "lib", "libmans_emu.so"))
    emu.mans_emu_decode_index.argtypes = [_u32p, C.c_uint64, _u8p, C.c_uint64, _u32p, C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_void_p, C.c_void_p, _u64p]
    sps, cus = _sync(streams)
    img = container.pack([e for e, _ in streams], sps, cus)
    c = container.Container(img)
    base = 1 << 32                                    # offsets into the image = d.* - base
    ds, _ = c.descs(base, 1 << 40, 1 << 40, True)
    for i, (es, sym) in enumerate(streams):
        if es.n_symbols == 0:
            continue
        d = ds[i]
        uo, to = d.units - base, d.table - base
        units = np.ascontiguousarray(img[uo:uo + 4 * d.n_units].view(np.uint32))
        table = np.ascontiguousarray(img[to:to + 4 * d.L].view(np.uint32))
        S = d.chunk_units or 256
        rec = c.sync_points(i)
        out = np.zeros(d.n_out, np.uint8)
        stats = np.zeros(8, np.uint64)
        rin = np.ascontiguousarray(rec) if rec is not None else None
        rc = emu.mans_emu_decode_index(units, d.n_units, out, d.n_out, table, d.L,
                                       d.first_state, d.first_bit, S, 16,
                                       rin.ctypes.data if rin is not None else None, None, stats)
        assert rc == 0
        assert np.array_equal(out[::-1], sym), i
        if rec is not None:
            assert stats[2] == 0                      # every chunk started on the true path


def _u32(img, off, v):
    img[off:off + 4] = np.frombuffer(struct.pack("<I", v), np.uint8)


def _u64(img, off, v):
    img[off:off + 8] = np.frombuffer(struct.pack("<Q", v), np.uint8)


def _r64(img, off):
    return struct.unpack("<Q", img[off:off + 8].tobytes())[0]


def test_damaged_images_refused(streams):
    from multians_amd import _lib, container
    sps, cus = _sync(streams)
    good = container.pack([e for e, _ in streams], sps, cus)
    container.Container(good)
    toff, soff = _r64(good, 24), _r64(good, 32)

    def srec(i):
        return soff + 64 * i

    damage = {
        "magic": lambda m: m.__setitem__(0, ord("X")),
        "version": lambda m: _u32(m, 8, 2),
        "header_bytes": lambda m: _u32(m, 12, 64),
        "image_bytes_beyond": lambda m: _u64(m, 40, len(m) + 256),
        "directory_beyond": lambda m: _u32(m, 16, 1 << 30),
        "table_L": lambda m: _u32(m, toff, 1000),
        "table_beyond": lambda m: _u64(m, toff + 8, len(m) - 64),
        "stream_table": lambda m: _u32(m, srec(3) + 24, 99),
        "units_beyond": lambda m: _u64(m, srec(4) + 8, 1 << 28),
        "units_unaligned": lambda m: _u64(m, srec(4), _r64(m, srec(4)) + 4),
        "first_bit": lambda m: _u32(m, srec(5) + 32, 33),
        "first_state": lambda m: _u32(m, srec(5) + 28, 1),
        "symbols_without_units": lambda m: _u64(m, srec(0) + 16, 10),
        "chunk_units": lambda m: _u32(m, srec(3) + 36, 100),
        "chunk_units_without_sync": lambda m: _u32(m, srec(2) + 36, 256),
        "sync_beyond": lambda m: _u64(m, srec(3) + 40, len(m) - 16),
        "sync_not_ascending": lambda m: _u64(m, _r64(m, srec(3) + 40) + 16 * 2 + 8, 1 << 40),
        "sync_state": lambda m: _u32(m, _r64(m, srec(5) + 40) + 16, 1 << 20),
    }
    for name, f in damage.items():
        m = good.copy()
        f(m)
        with pytest.raises(_lib.MansError) as e:
            container.Container(m)
        assert e.value.code == ECORRUPT, name
    with pytest.raises(_lib.MansError):
        container.Container(good[:100])                       # shorter than the header
    with pytest.raises(_lib.MansError):
        container.Container(good[:len(good) - 256])           # truncated


def test_pack_rejects_bad_input(streams):
    from multians_amd import _lib, container
    sps, cus = _sync(streams)
    es = [e for e, _ in streams]
    bad = list(cus)
    bad[3] = 100                                              # not a power of two
    with pytest.raises(_lib.MansError):
        container.pack(es, sps, bad)
    sp2 = list(sps)
    r = sps[3].copy()
    r["symbols_before"][2:] = 1                               # decreasing at record 2
    sp2[3] = r
    with pytest.raises(_lib.MansError):
        container.pack(es, sp2, cus)


def test_container_edge_cases(tmp_path):
    """No streams; one empty stream; a 16,384-state table; the image written to a file and
    read back as bytes."""
    from multians_amd import _lib, container
    from multians_amd.decoder import EncodedStream
    img = container.pack([])
    c = container.Container(img)
    assert (c.info.n_streams, c.info.n_tables, c.info.total_symbols) == (0, 0, 0)
    assert c.descs(1 << 32, 0, 0)[1].tolist() == [0]
    units, fb, fs, dec, sym = _lib.make_stream(5, 0.7, 300_000, 16384, True)
    es = EncodedStream(units, fb, fs, dec, 16384, len(sym))
    L = 1024
    empty = EncodedStream(np.zeros(0, np.uint32), 0, L,
                          (np.arange(L, dtype=np.uint32) + L) | (1 << 24), L, 0)
    rec = _truth(units, fb, fs, dec, 16384, len(sym), 512)
    img = container.pack([es, empty], [rec, None], [512, 0])
    f = tmp_path / "batch.mans"
    f.write_bytes(img.tobytes())
    c = container.Container(f.read_bytes())
    assert c.info.n_streams == 2 and c.info.total_symbols == len(sym)
    assert np.array_equal(c.sync_points(0), rec)
    assert c.sync_points(1) is None
    ds, offs = c.descs(1 << 32, 1 << 40, len(sym), True)
    assert ds[0].L == 16384 and ds[0].chunk_units == 512 and ds[1].n_out == 0
    assert offs.tolist() == [0, len(sym), len(sym)]
    with pytest.raises(_lib.MansError):                      # output arena too small
        c.descs(1 << 32, 1 << 40, len(sym) - 1, True)
    with pytest.raises(_lib.MansError):                      # device image not 256-aligned
        c.descs((1 << 32) + 16, 1 << 40, len(sym), True)
