"""ctypes bindings for the oracle (oracle/libmans_oracle.so) and, when it was built,
the reference's own CPU code (oracle/_ref/libmans_ref.so).

Test infrastructure only: the checker, never the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libmans_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmans_ref.so")

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


class EncEntry(C.Structure):
    _fields_ = [("next_state", C.c_uint32), ("code", C.c_uint32), ("len", C.c_uint32)]


def _load_oracle():
    lib = C.CDLL(ORACLE_SO)
    lib.mo_gen_symbols.argtypes = [C.c_uint64, C.c_double, _u8p, C.c_size_t]
    lib.mo_gen_symbols.restype = None
    lib.mo_normalize.argtypes = [_u64p, C.c_int, C.c_uint32, _u32p]
    lib.mo_build_tables.argtypes = [_u32p, C.c_int, C.c_uint32, C.c_void_p, _u32p]
    lib.mo_max_units.argtypes = [C.c_size_t, C.c_uint32]
    lib.mo_max_units.restype = C.c_size_t
    lib.mo_encode.argtypes = [_u8p, C.c_size_t, C.c_void_p, C.c_uint32, _u32p,
                              C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.mo_encode.restype = C.c_size_t
    lib.mo_decode_seq.argtypes = [_u32p, C.c_size_t, _u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                  _u8p, C.c_size_t, C.c_uint32, C.c_void_p, C.c_void_p,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
    lib.mo_subseq_last.argtypes = [_u32p, C.c_size_t, _u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_size_t, C.c_uint32, _u32p, _u64p]
    lib.mo_encode_interleaved.argtypes = [_u8p, C.c_size_t, C.c_void_p, C.c_uint32, C.c_uint32,
                                          _u32p, C.POINTER(C.c_uint32), _u32p]
    lib.mo_encode_interleaved.restype = C.c_size_t
    lib.mo_decode_interleaved.argtypes = [_u32p, C.c_size_t, C.c_void_p, C.c_uint32, C.c_uint32,
                                          _u32p, C.c_uint32, _u8p, C.c_size_t, C.c_size_t,
                                          C.c_uint32, C.c_void_p, _u32p, C.POINTER(C.c_uint64)]
    lib.mo_decode_multicore.argtypes = [C.c_size_t, C.c_size_t, _u32p, C.c_size_t, _u32p,
                                        C.c_uint32, C.c_uint32, C.c_uint32, _u8p, C.c_size_t]
    return lib


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _oracle = _load_oracle()
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.ref_demo_tables.argtypes = [C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
                                        _u32p, _u32p, _u32p, _u32p]
        lib.ref_tables_from_counts.argtypes = [_u32p, C.c_uint32, C.c_uint32, _u32p, _u32p,
                                               _u32p, _u32p]
        lib.ref_gen_test_data.argtypes = [_u32p, C.c_uint32, C.c_size_t, C.c_uint32,
                                          C.c_uint32, _u8p]
        lib.ref_encode.argtypes = [_u32p, C.c_uint32, C.c_uint32, _u8p, C.c_size_t, _u32p,
                                   C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.ref_encode.restype = C.c_long
        lib.ref_multicore.argtypes = [C.c_size_t, C.c_size_t, _u32p, C.c_size_t, _u32p,
                                      C.c_uint32, C.c_uint32, C.c_uint32, _u8p, C.c_size_t]
        lib.ref_multicore.restype = C.c_double
        _ref = lib
    return _ref


# ---------------------------------------------------------------- oracle API

def gen_symbols(seed, lam, n):
    out = np.empty(n, np.uint8)
    oracle().mo_gen_symbols(seed, lam, out, n)
    return out


def normalize(hist, L):
    hist = np.ascontiguousarray(hist, np.uint64)
    counts = np.zeros(len(hist), np.uint32)
    rc = oracle().mo_normalize(hist, len(hist), L, counts)
    if rc:
        raise ValueError("normalize failed")
    return counts


def build_tables(counts, L):
    counts = np.ascontiguousarray(counts, np.uint32)
    nsym = len(counts)
    enc = (EncEntry * (nsym * L))()
    dec = np.zeros(L, np.uint32)
    rc = oracle().mo_build_tables(counts, nsym, L, C.cast(enc, C.c_void_p), dec)
    if rc:
        raise ValueError("build_tables failed: %d" % rc)
    encarr = np.frombuffer(enc, dtype=np.uint32).reshape(nsym, L, 3)
    return enc, encarr, dec


class Stream:
    """One encoded stream in the reference's decoder-order format."""

    def __init__(self, units, first_bit, first_state, table, L, n):
        self.units = units
        self.first_bit = first_bit
        self.first_state = first_state
        self.table = table
        self.L = L
        self.n = n


def encode(symbols, counts=None, L=4096):
    symbols = np.ascontiguousarray(symbols, np.uint8)
    if counts is None:
        hist = np.bincount(symbols, minlength=256).astype(np.uint64)
        counts = normalize(hist, L)
    enc, _, dec = build_tables(counts, L)
    n = len(symbols)
    out = np.zeros(oracle().mo_max_units(n, L), np.uint32)
    fb, fs = C.c_uint32(), C.c_uint32()
    nu = oracle().mo_encode(symbols, n, C.cast(enc, C.c_void_p), L, out, C.byref(fb), C.byref(fs))
    if nu == 0:
        raise ValueError("encode failed")
    return Stream(out[:nu].copy(), fb.value, fs.value, dec, L, n), counts


def decode_seq(st, n_out=None, S=0):
    n_out = st.n if n_out is None else n_out
    out = np.zeros(n_out, np.uint8)
    nsub = (len(st.units) + S - 1) // S if S else 0
    entries = np.zeros(max(nsub, 1), np.uint32)
    counts = np.zeros(max(nsub, 1), np.uint32)
    es, eb = C.c_uint32(), C.c_uint64()
    rc = oracle().mo_decode_seq(st.units, len(st.units), st.table, st.L, st.first_state,
                                st.first_bit, out, n_out, S,
                                entries.ctypes.data if S else None,
                                counts.ctypes.data if S else None, C.byref(es), C.byref(eb))
    if rc:
        raise ValueError("decode_seq failed")
    return out, entries[:nsub], counts[:nsub], es.value, eb.value


def subseq_last(st, S, n_out=None):
    """(state before, absolute start bit) of the last step of each subsequence of S units."""
    n_out = st.n if n_out is None else n_out
    nsub = (len(st.units) + S - 1) // S
    ls = np.zeros(max(nsub, 1), np.uint32)
    lb = np.zeros(max(nsub, 1), np.uint64)
    rc = oracle().mo_subseq_last(st.units, len(st.units), st.table, st.L, st.first_state,
                                 st.first_bit, n_out, S, ls, lb)
    if rc:
        raise ValueError("subseq_last failed")
    return ls[:nsub], lb[:nsub]


IL_RECORD_DTYPE = np.dtype([("state", "<u2", (2,)), ("bit", "<u4"), ("before", "<u8")])


class ILStream:
    """An interleaved-state stream: units (decoder order), first_bit, first_states[N],
    decoder tables[N] (u32[L] each), n symbols."""

    def __init__(self, units, first_bit, first_states, tables, L, n):
        self.units, self.first_bit = units, first_bit
        self.first_states = np.ascontiguousarray(first_states, np.uint32)
        self.tables = [np.ascontiguousarray(t, np.uint32) for t in tables]
        self.L, self.n, self.N = L, n, len(self.first_states)


def encode_interleaved(symbols, N, counts_list, L):
    """counts_list: one count vector per slot (the same vector N times = one shared
    table).  Returns (ILStream, encoder table arrays)."""
    symbols = np.ascontiguousarray(symbols, np.uint8)
    encs, decs = [], []
    for c in counts_list:
        enc, _, dec = build_tables(c, L)
        encs.append(enc)
        decs.append(dec)
    ptrs = (C.c_void_p * N)(*[C.cast(e, C.c_void_p) for e in encs])
    n = len(symbols)
    out = np.zeros(oracle().mo_max_units(n, L) + 2, np.uint32)
    fb = C.c_uint32()
    fs = np.zeros(N, np.uint32)
    nu = oracle().mo_encode_interleaved(symbols, n, C.cast(ptrs, C.c_void_p), L, N, out,
                                        C.byref(fb), fs)
    if nu == 0:
        raise ValueError("interleaved encode failed")
    return ILStream(out[:nu].copy(), fb.value, fs, decs, L, n), encs


def decode_interleaved(st, n_out=None, R=0):
    """Sequential decode (decode order); with R > 0 also the records every R units."""
    n_out = st.n if n_out is None else n_out
    N = st.N
    tabs = (C.c_void_p * N)(*[t.ctypes.data for t in st.tables])
    out = np.zeros(max(n_out, 1), np.uint8)
    nrec = (len(st.units) + R - 1) // R if R else 0
    rec = np.zeros(max(nrec, 1), IL_RECORD_DTYPE)
    es = np.zeros(N, np.uint32)
    eb = C.c_uint64()
    rc = oracle().mo_decode_interleaved(st.units, len(st.units), C.cast(tabs, C.c_void_p), st.L,
                                        N, st.first_states, st.first_bit, out, n_out, st.n, R,
                                        rec.ctypes.data if R else None, es, C.byref(eb))
    if rc:
        raise ValueError("interleaved decode failed")
    return out[:n_out], rec[:nrec], es, eb.value


def decode_multicore(st, S=4, T=4, n_out=None, pad=None):
    n_out = st.n if n_out is None else n_out
    pad = S + 4 if pad is None else pad
    units = np.concatenate([st.units, np.zeros(pad, np.uint32)])
    out = np.zeros(n_out, np.uint8)
    rc = oracle().mo_decode_multicore(S, T, units, len(st.units), st.table, st.L,
                                      st.first_state, st.first_bit, out, n_out)
    if rc:
        raise ValueError("multicore failed")
    return out
