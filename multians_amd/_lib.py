"""ctypes bindings of libmultians_amd.so (the C ABI declared in include/mans_c_api.h).

The library is the only decode path: importing a decode function with the library
missing raises, there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmultians_amd.so")

MANS_OK = 0
MANS_EINVAL = -1
MANS_EHIP = -2
MANS_ENOMEM = -3
MANS_ETABLE = -4
MANS_ECORRUPT = -5
MANS_F_SYNC_ERRORS = 1
MANS_F_TRUST_SYNC_POINTS = 2


class MansError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mans error %d: %s" % (code, msg))
        self.code = code


class DecodeArgs(C.Structure):
    _fields_ = [("units", C.c_void_p), ("n_units", C.c_uint64), ("out", C.c_void_p),
                ("n_out", C.c_uint64), ("table", C.c_void_p), ("L", C.c_uint32),
                ("first_state", C.c_uint32), ("first_bit", C.c_uint32),
                ("chunk_units", C.c_uint32), ("checkpoint_units", C.c_uint32),
                ("flags", C.c_uint32), ("workspace", C.c_void_p),
                ("workspace_bytes", C.c_uint64), ("stream", C.c_void_p),
                ("kernel_events", C.c_void_p), ("index_in", C.c_void_p),
                ("index_out", C.c_void_p), ("items_out", C.c_void_p),
                ("ref_sync", C.c_void_p)]


class RefSync(C.Structure):
    """mans_ref_sync: the reference decoder's sync state arrays (device pointers)."""
    _fields_ = [("sync_points", C.c_void_p), ("output_sizes", C.c_void_p),
                ("synced_flags", C.c_void_p), ("subsequence_units", C.c_uint32),
                ("threads", C.c_uint32)]


IL_RECORD_DTYPE = np.dtype([("state", "<u2", (2,)), ("bit", "<u4"), ("before", "<u8")])
IL_MIN_RECORD_UNITS = 16
IL_WORKSPACE_BYTES = 256


class IlArgs(C.Structure):
    """mans_il_args: a two-state interleaved stream on the device."""
    _fields_ = [("units", C.c_void_p), ("n_units", C.c_uint64), ("out", C.c_void_p),
                ("n_symbols", C.c_uint64), ("tables", C.c_void_p * 2), ("L", C.c_uint32),
                ("first_states", C.c_uint32 * 2), ("first_bit", C.c_uint32),
                ("record_units", C.c_uint32), ("records", C.c_void_p), ("flags", C.c_uint32),
                ("workspace", C.c_void_p), ("stream", C.c_void_p)]


N_KERNEL_EVENTS = 7
KERNEL_NAMES = ["init", "sync", "check", "fixup", "scan", "write"]


class StreamDesc(C.Structure):
    _fields_ = [("units", C.c_void_p), ("n_units", C.c_uint64), ("out", C.c_void_p),
                ("n_out", C.c_uint64), ("table", C.c_void_p), ("L", C.c_uint32),
                ("first_state", C.c_uint32), ("first_bit", C.c_uint32),
                ("chunk_units", C.c_uint32), ("index_in", C.c_void_p),
                ("index_out", C.c_void_p), ("items_out", C.c_void_p)]


# mans_sync_point: one record per sync chunk (state = x - L, bit offset, symbols before)
SYNC_POINT_DTYPE = np.dtype([("state", "<u4"), ("bit", "<u4"), ("symbols_before", "<u8")])


class SyncPoint(C.Structure):
    _fields_ = [("state", C.c_uint32), ("bit", C.c_uint32), ("symbols_before", C.c_uint64)]


class BatchOpts(C.Structure):
    _fields_ = [("workspace", C.c_void_p), ("workspace_bytes", C.c_uint64),
                ("stream", C.c_void_p), ("flags", C.c_uint32), ("reserved", C.c_uint32),
                ("kernel_events", C.c_void_p), ("stream_status", C.c_void_p)]


MANS_BATCH_ONE_PIPELINE = 1
MANS_BATCH_TRUST_SYNC_POINTS = 2


class TableJob(C.Structure):
    _fields_ = [("symbols", C.c_void_p), ("n", C.c_uint64), ("L", C.c_uint32),
                ("reserved", C.c_uint32), ("table", C.c_void_p), ("counts", C.c_void_p)]


class EncodeJob(C.Structure):
    _fields_ = [("symbols", C.c_void_p), ("n", C.c_uint64), ("table", C.c_void_p),
                ("L", C.c_uint32), ("reserved", C.c_uint32), ("units", C.c_void_p),
                ("units_capacity", C.c_uint64), ("items_out", C.c_void_p)]


class EncodeResult(C.Structure):
    _fields_ = [("n_units", C.c_uint64), ("first_bit", C.c_uint32), ("first_state", C.c_uint32),
                ("status", C.c_uint32), ("reserved", C.c_uint32)]


class GeometryRule(C.Structure):
    _fields_ = [("L_max", C.c_uint32), ("chunk_units", C.c_uint32), ("spu_max", C.c_double),
                ("units_max", C.c_uint64), ("warm_units", C.c_uint32), ("scope", C.c_uint32)]


WARM_AUTO = 0xFFFFFFFF
RULE_SINGLE = 1
RULE_BATCH = 2
RULE_LINEAGE = 4


class BatchStats(C.Structure):
    _fields_ = [("total_units", C.c_uint64), ("total_symbols", C.c_uint64),
                ("n_streams", C.c_uint32), ("n_chunks", C.c_uint32),
                ("n_items", C.c_uint32), ("status", C.c_uint32)]


# every symbol include/mans_c_api.h declares
EXPORTS = [
    "mans_error_string", "mans_last_error", "mans_set_device", "mans_get_device_count",
    "mans_malloc", "mans_free", "mans_memset", "mans_memcpy_h2d", "mans_memcpy_d2h",
    "mans_memcpy_d2d", "mans_host_alloc_pinned", "mans_host_free_pinned",
    "mans_stream_create", "mans_stream_destroy", "mans_sync", "mans_event_create",
    "mans_event_destroy", "mans_event_record", "mans_event_elapsed_ms",
    "mans_workspace_bytes", "mans_workspace_bytes_geometry", "mans_decode", "mans_decode_status",
    "mans_batch_workspace_bytes", "mans_decode_batch", "mans_gen_symbols",
    "mans_normalize_counts", "mans_build_tables", "mans_max_encoded_units", "mans_encode",
    "mans_make_stream", "mans_batch_graph_create", "mans_graph_launch", "mans_graph_destroy",
    "mans_set_geometry_rules", "mans_chunk_units", "mans_histogram", "mans_set_wait_polls",
    "mans_set_sync_round_limit",
    "mans_table_workspace_bytes", "mans_build_tables_device", "mans_table_status",
    "mans_sync_point_count", "mans_stream_range", "mans_sync_points_check",
    "mans_encode_workspace_bytes", "mans_encode_device", "mans_encode_results",
    "mans_gen_symbols_device", "mans_ref_sync_fill", "mans_decode_noalloc",
    "mans_il_record_count", "mans_encode_interleaved", "mans_decode_interleaved",
    "mans_set_lineage_cap", "mans_lineage_stats", "mans_lineage_spacing",
]

_lib = None


def lib():
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmultians_amd.so not built: run `python -m multians_amd.build`")
    try:
        # torch loads its own copy of the HIP runtime (same soname, by path); loaded
        # first, it also serves this library.  Loaded the other way round, the process
        # would hold two HIP runtimes, and the second finds no device.
        import torch  # noqa: F401
    except ImportError:
        pass
    l = C.CDLL(LIB_PATH)
    vp, u64, u32, i32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
    sig = {
        "mans_error_string": ([i32], C.c_char_p),
        "mans_last_error": ([], C.c_char_p),
        "mans_set_device": ([i32], i32),
        "mans_get_device_count": ([C.POINTER(i32)], i32),
        "mans_malloc": ([C.POINTER(vp), C.c_size_t], i32),
        "mans_free": ([vp], i32),
        "mans_memset": ([vp, i32, C.c_size_t, vp], i32),
        "mans_memcpy_h2d": ([vp, vp, C.c_size_t, vp], i32),
        "mans_memcpy_d2h": ([vp, vp, C.c_size_t, vp], i32),
        "mans_memcpy_d2d": ([vp, vp, C.c_size_t, vp], i32),
        "mans_host_alloc_pinned": ([C.POINTER(vp), C.c_size_t], i32),
        "mans_host_free_pinned": ([vp], i32),
        "mans_stream_create": ([C.POINTER(vp)], i32),
        "mans_stream_destroy": ([vp], i32),
        "mans_sync": ([vp], i32),
        "mans_event_create": ([C.POINTER(vp)], i32),
        "mans_event_destroy": ([vp], i32),
        "mans_event_record": ([vp, vp], i32),
        "mans_event_elapsed_ms": ([vp, vp, C.POINTER(C.c_float)], i32),
        "mans_workspace_bytes": ([u64, u32, C.POINTER(u64)], i32),
        "mans_workspace_bytes_geometry": ([u64, u32, u32, C.POINTER(u64)], i32),
        "mans_decode": ([C.POINTER(DecodeArgs)], i32),
        "mans_decode_status": ([vp, vp, C.POINTER(i32)], i32),
        "mans_batch_workspace_bytes": ([C.POINTER(StreamDesc), C.c_size_t, C.POINTER(u64)], i32),
        "mans_decode_batch": ([C.POINTER(StreamDesc), C.c_size_t, C.POINTER(BatchOpts),
                               C.POINTER(BatchStats)], i32),
        "mans_gen_symbols": ([u64, C.c_double, vp, u64, i32], i32),
        "mans_normalize_counts": ([vp, u32, u32, vp], i32),
        "mans_build_tables": ([vp, u32, u32, vp, vp], i32),
        "mans_max_encoded_units": ([u64], u64),
        "mans_encode": ([vp, u64, vp, u32, vp, C.POINTER(u64), C.POINTER(u32), C.POINTER(u32)], i32),
        "mans_make_stream": ([u64, C.c_double, u64, u32, vp, vp, u64, C.POINTER(u64),
                              C.POINTER(u32), C.POINTER(u32), vp], i32),
        "mans_batch_graph_create": ([C.POINTER(StreamDesc), C.c_size_t, C.POINTER(BatchOpts),
                                     C.POINTER(vp)], i32),
        "mans_graph_launch": ([vp, vp], i32),
        "mans_graph_destroy": ([vp], i32),
        "mans_set_geometry_rules": ([C.POINTER(GeometryRule), C.c_size_t], i32),
        "mans_chunk_units": ([u64, u64, u32, C.POINTER(u32)], i32),
        "mans_histogram": ([vp, u64, vp, i32], i32),
        "mans_set_wait_polls": ([u32], i32),
        "mans_gen_symbols_device": ([u64, C.c_double, C.c_void_p, u64, C.c_void_p], i32),
        "mans_ref_sync_fill": ([C.POINTER(DecodeArgs)], i32),
        "mans_decode_noalloc": ([C.POINTER(DecodeArgs)], i32),
        "mans_set_sync_round_limit": ([u32], i32),
        "mans_table_workspace_bytes": ([C.c_size_t, C.POINTER(u64)], i32),
        "mans_build_tables_device": ([C.POINTER(TableJob), C.c_size_t, vp, u64, vp], i32),
        "mans_table_status": ([vp, C.c_size_t, vp, C.POINTER(u32)], i32),
        "mans_sync_point_count": ([u64, u32, C.POINTER(u64)], i32),
        "mans_encode_workspace_bytes": ([C.POINTER(EncodeJob), C.c_size_t, C.POINTER(u64)], i32),
        "mans_encode_device": ([C.POINTER(EncodeJob), C.c_size_t, vp, u64, vp], i32),
        "mans_encode_results": ([vp, C.POINTER(EncodeJob), C.c_size_t, vp,
                                 C.POINTER(EncodeResult)], i32),
        "mans_sync_points_check": ([vp, u64, u32], i32),
        "mans_il_record_count": ([u64, u32, C.POINTER(u64)], i32),
        "mans_encode_interleaved": ([vp, u64, vp, vp, u32, u32, vp, C.POINTER(u64),
                                     C.POINTER(u32), vp, vp, u64, C.POINTER(u64)], i32),
        "mans_decode_interleaved": ([C.POINTER(IlArgs)], i32),
        "mans_set_lineage_cap": ([u32], i32),
        "mans_lineage_stats": ([vp, vp, C.POINTER(u32)], i32),
        "mans_lineage_spacing": ([u64, u64, u32, C.POINTER(u32)], i32),
        "mans_stream_range": ([C.POINTER(StreamDesc), vp, vp, u64, u64, C.POINTER(StreamDesc),
                               C.POINTER(u64)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(l, name)
        f.argtypes = args
        f.restype = res
    _lib = l
    return l


def check(rc):
    if rc != 0:
        raise MansError(rc, lib().mans_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ host codec

def gen_symbols(seed, lam, n, threads=8):
    out = np.empty(n, np.uint8)
    check(lib().mans_gen_symbols(seed, lam, _ptr(out), n, threads))
    return out


def histogram(symbols, threads=8):
    symbols = np.ascontiguousarray(symbols, np.uint8)
    h = np.zeros(256, np.uint64)
    check(lib().mans_histogram(_ptr(symbols), len(symbols), _ptr(h), threads))
    return h


def normalize_counts(hist, L):
    hist = np.ascontiguousarray(hist, np.uint64)
    counts = np.zeros(len(hist), np.uint32)
    check(lib().mans_normalize_counts(_ptr(hist), len(hist), L, _ptr(counts)))
    return counts


def build_tables(counts, L, want_enc=True):
    counts = np.ascontiguousarray(counts, np.uint32)
    nsym = len(counts)
    enc = np.zeros(3 * nsym * L, np.uint32) if want_enc else None
    dec = np.zeros(L, np.uint32)
    check(lib().mans_build_tables(_ptr(counts), nsym, L, _ptr(enc) if want_enc else None, _ptr(dec)))
    return enc, dec


def encode(symbols, enc, L):
    symbols = np.ascontiguousarray(symbols, np.uint8)
    n = len(symbols)
    units = np.empty(lib().mans_max_encoded_units(n), np.uint32)
    nu, fb, fs = C.c_uint64(), C.c_uint32(), C.c_uint32()
    check(lib().mans_encode(_ptr(symbols), n, _ptr(enc), L, _ptr(units), C.byref(nu), C.byref(fb),
                            C.byref(fs)))
    return units[:nu.value].copy(), fb.value, fs.value


def il_record_count(n_units, record_units):
    c = C.c_uint64()
    check(lib().mans_il_record_count(n_units, record_units, C.byref(c)))
    return c.value


def encode_interleaved(symbols, enc0, enc1, L, record_units):
    """mans_encode_interleaved: (units, first_bit, first_states[2], records)."""
    symbols = np.ascontiguousarray(symbols, np.uint8)
    n = len(symbols)
    cap_units = lib().mans_max_encoded_units(n)
    units = np.empty(cap_units, np.uint32)
    records = np.zeros(il_record_count(cap_units, record_units), IL_RECORD_DTYPE)
    nu, fb, nrec = C.c_uint64(), C.c_uint32(), C.c_uint64()
    fs = np.zeros(2, np.uint32)
    check(lib().mans_encode_interleaved(_ptr(symbols), n, _ptr(enc0), _ptr(enc1), L, record_units,
                                        _ptr(units), C.byref(nu), C.byref(fb), _ptr(fs),
                                        _ptr(records), len(records), C.byref(nrec)))
    return units[:nu.value].copy(), fb.value, fs, records[:nrec.value].copy()


def make_stream(seed, lam, n, L, keep_symbols=True):
    """Generate + encode one stream in one native call; returns (units, fb, fs, table, symbols)."""
    sym = np.empty(n, np.uint8) if keep_symbols else None
    units = np.empty(lib().mans_max_encoded_units(n), np.uint32)
    dec = np.empty(L, np.uint32)
    nu, fb, fs = C.c_uint64(), C.c_uint32(), C.c_uint32()
    check(lib().mans_make_stream(seed, lam, n, L, _ptr(sym) if keep_symbols else None, _ptr(units),
                                 len(units), C.byref(nu), C.byref(fb), C.byref(fs), _ptr(dec)))
    return units[:nu.value], fb.value, fs.value, dec, sym


def encode_device(jobs, stream=None):
    """Encode on the device.  jobs: list of (symbols u8 tensor, decoder table tensor, L,
    units tensor of int32 with capacity >= max_encoded_units(n)[, item records tensor of
    sync_point_count(capacity, 32) records]).  Returns a list of (n_units, first_bit,
    first_state, status); waits for the stream."""
    import torch
    n = len(jobs)
    js = (EncodeJob * n)()
    for k, job in enumerate(jobs):
        sym, table, L, units = job[:4]
        js[k].symbols = sym.data_ptr() if sym.numel() else units.data_ptr()
        js[k].n = sym.numel()
        js[k].table, js[k].L = table.data_ptr(), L
        js[k].units, js[k].units_capacity = units.data_ptr(), units.numel()
        if len(job) > 4 and job[4] is not None:
            js[k].items_out = job[4].data_ptr()
    b = C.c_uint64()
    check(lib().mans_encode_workspace_bytes(js, n, C.byref(b)))
    ws = torch.empty(max(b.value, 256), dtype=torch.uint8, device="cuda")
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().mans_encode_device(js, n, C.c_void_p(ws.data_ptr()), ws.numel(), C.c_void_p(st)))
    res = (EncodeResult * n)()
    check(lib().mans_encode_results(C.c_void_p(ws.data_ptr()), js, n, C.c_void_p(st), res))
    return [(r.n_units, r.first_bit, r.first_state, r.status) for r in res]


def sync_point_count(n_units, chunk_units):
    c = C.c_uint64()
    check(lib().mans_sync_point_count(n_units, chunk_units, C.byref(c)))
    return c.value


def sync_points_check(index_host, L):
    index_host = np.ascontiguousarray(index_host, SYNC_POINT_DTYPE)
    check(lib().mans_sync_points_check(_ptr(index_host), len(index_host), L))


def stream_range(full, index_host, index_dev_ptr, first, last):
    """mans_stream_range: (sub-stream StreamDesc, skip) for the output window [first, last)
    of the stream `full` (a StreamDesc with chunk_units), given its sync points on the host
    (SYNC_POINT_DTYPE array) and their device address."""
    index_host = np.ascontiguousarray(index_host, SYNC_POINT_DTYPE)
    sub, skip = StreamDesc(), C.c_uint64()
    check(lib().mans_stream_range(C.byref(full), _ptr(index_host), C.c_void_p(index_dev_ptr),
                                  first, last, C.byref(sub), C.byref(skip)))
    return sub, skip.value


class KernelTimer:
    """HIP events around the decode kernels (kernel_events argument)."""

    def __init__(self):
        L = lib()
        self.evs = [C.c_void_p() for _ in range(N_KERNEL_EVENTS)]
        for e in self.evs:
            check(L.mans_event_create(C.byref(e)))
        self.arr = (C.c_void_p * N_KERNEL_EVENTS)(*[e.value for e in self.evs])

    @property
    def ptr(self):
        return C.cast(self.arr, C.c_void_p)

    def times_ms(self):
        L = lib()
        out = []
        for k in range(N_KERNEL_EVENTS - 1):
            ms = C.c_float()
            check(L.mans_event_elapsed_ms(self.evs[k], self.evs[k + 1], C.byref(ms)))
            out.append(ms.value)
        return dict(zip(KERNEL_NAMES, out))

    def close(self):
        for e in self.evs:
            lib().mans_event_destroy(e)


def lineage_spacing(n_units, n_out, L):
    """Lineage spacing the rules give a single stream of this shape (0 = chunked)."""
    v = C.c_uint32()
    check(lib().mans_lineage_spacing(n_units, n_out, L, C.byref(v)))
    return v.value


def set_geometry_rules(rules):
    """rules: iterable of (L_max, spu_max, units_max, chunk_units[, warm_units[, scope]]);
    warm_units defaults to WARM_AUTO, scope to 0 (single and batch); [] = built-in."""
    rules = list(rules)
    arr = (GeometryRule * max(len(rules), 1))()
    for i, r in enumerate(rules):
        L_max, spu_max, units_max, chunk = r[:4]
        arr[i].L_max, arr[i].spu_max = int(L_max), float(spu_max)
        arr[i].units_max, arr[i].chunk_units = int(units_max), int(chunk)
        arr[i].warm_units = int(r[4]) if len(r) > 4 else WARM_AUTO
        arr[i].scope = int(r[5]) if len(r) > 5 else 0
    check(lib().mans_set_geometry_rules(arr, len(rules)))


def build_tables_device(jobs, stream=None):
    """Decoder tables on the device: jobs = [(symbols u8 tensor, L, table u32 tensor,
    counts u32 tensor or None)], all on one device.  Returns the per-job status list
    (0 = built); waits for the stream."""
    import torch
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    arr = (TableJob * len(jobs))()
    for i, (sym, L, table, counts) in enumerate(jobs):
        arr[i].symbols, arr[i].n, arr[i].L = sym.data_ptr(), sym.numel(), int(L)
        arr[i].table = table.data_ptr()
        arr[i].counts = counts.data_ptr() if counts is not None else None
    b = C.c_uint64()
    check(lib().mans_table_workspace_bytes(len(jobs), C.byref(b)))
    ws = torch.empty(b.value, dtype=torch.uint8, device=jobs[0][0].device)
    check(lib().mans_build_tables_device(arr, len(jobs), C.c_void_p(ws.data_ptr()), b.value,
                                         C.c_void_p(st)))
    status = (C.c_uint32 * len(jobs))()
    check(lib().mans_table_status(C.c_void_p(ws.data_ptr()), len(jobs), C.c_void_p(st), status))
    return list(status)


WAIT_POLLS_DEFAULT = 1 << 20


def set_wait_polls(polls=WAIT_POLLS_DEFAULT):
    """Polls of the sync kernel's wait for the previous range's exit (see mans_c_api.h)."""
    check(lib().mans_set_wait_polls(int(polls)))


def gen_symbols_device(seed, lam, out, stream=None):
    """mans_gen_symbols_device into the u8 device tensor `out` (its whole length)."""
    import torch
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().mans_gen_symbols_device(int(seed), float(lam), C.c_void_p(out.data_ptr()),
                                        out.numel(), C.c_void_p(st)))


def set_sync_round_limit(rounds=0):
    """Fixup rounds per stream before non-convergence is reported (0 = the proven
    bound; see mans_c_api.h)."""
    check(lib().mans_set_sync_round_limit(int(rounds)))


def chunk_units(n_units, n_out, L):
    v = C.c_uint32()
    check(lib().mans_chunk_units(int(n_units), int(n_out), int(L), C.byref(v)))
    return v.value
