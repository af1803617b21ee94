"""Python host layer over the C ABI, used by tests and bench.py.

Mirrors the reference's host objects (src/multians_input_buffer.cc, src/multians_codetable.cc,
src/multians_output_buffer.cc) and decode entry point (CUHDGPUDecoder::decode,
include/cuhd_gpu_decoder.h:24-35).  Device memory and streams come from torch;
all decoding runs in libmultians_amd.so.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class EncodedStream:
    """A compressed stream in the reference's CUHDInputBuffer format."""
    units: np.ndarray          # u32, decoder order
    first_bit: int
    first_state: int
    table: np.ndarray          # u32[L], CUHDCodetableItem layout
    L: int
    n_symbols: int
    counts: np.ndarray = None

    @property
    def n_units(self):
        return len(self.units)


def encode_symbols(symbols, L, counts=None):
    """Histogram -> counts -> tables -> encoded stream (host, product codec)."""
    symbols = np.ascontiguousarray(symbols, np.uint8)
    if counts is None:
        hist = np.bincount(symbols, minlength=256).astype(np.uint64)
        counts = _lib.normalize_counts(hist, L)
    enc, dec = _lib.build_tables(counts, L)
    units, fb, fs = _lib.encode(symbols, enc, L)
    return EncodedStream(units, fb, fs, dec, L, len(symbols), counts)


def _torch():
    import torch
    return torch


def workspace_bytes(n_units, L, chunk_units=0, checkpoint_units=0):
    b = C.c_uint64()
    if chunk_units and checkpoint_units:
        _lib.check(_lib.lib().mans_workspace_bytes_geometry(n_units, chunk_units,
                                                            checkpoint_units, C.byref(b)))
    else:
        _lib.check(_lib.lib().mans_workspace_bytes(n_units, L, C.byref(b)))
    return b.value


class DeviceStream:
    """A stream resident in device memory (torch tensors), ready to decode."""

    def __init__(self, es: EncodedStream, device="cuda", chunk_units=0, checkpoint_units=0):
        torch = _torch()
        self.es = es
        self.units = torch.from_numpy(es.units.view(np.int32)).to(device)
        self.table = torch.from_numpy(es.table.view(np.int32)).to(device)
        self.out = torch.zeros(max(es.n_symbols, 1), dtype=torch.uint8, device=device)
        self.ws = torch.zeros(workspace_bytes(es.n_units, es.L, chunk_units, checkpoint_units),
                              dtype=torch.uint8, device=device)

    def args(self, stream=None, chunk_units=0, checkpoint_units=0, flags=0, n_out=None,
             kernel_events=None, index_in=None, index_out=None, items_out=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        a = _lib.DecodeArgs()
        a.units = self.units.data_ptr()
        a.n_units = self.es.n_units
        a.out = self.out.data_ptr()
        a.n_out = self.es.n_symbols if n_out is None else n_out
        a.table = self.table.data_ptr()
        a.L = self.es.L
        a.first_state = self.es.first_state
        a.first_bit = self.es.first_bit
        a.chunk_units = chunk_units
        a.checkpoint_units = checkpoint_units
        a.flags = flags
        a.workspace = self.ws.data_ptr()
        a.workspace_bytes = self.ws.numel()
        a.stream = st
        a.kernel_events = kernel_events
        a.index_in = index_in.data_ptr() if index_in is not None else None
        a.index_out = index_out.data_ptr() if index_out is not None else None
        a.items_out = items_out.data_ptr() if items_out is not None else None
        return a

    def decode(self, **kw):
        """index_in / index_out: device tensors of sync_point_count(n_units, chunk_units)
        records (new_index), with chunk_units given."""
        a = self.args(**kw)
        _lib.check(_lib.lib().mans_decode(C.byref(a)))
        return self.out

    def new_index(self, chunk_units):
        """Zeroed device buffer for this stream's sync points at this chunk length."""
        torch = _torch()
        n = _lib.sync_point_count(self.es.n_units, chunk_units)
        return torch.zeros(max(n, 1) * 16, dtype=torch.uint8, device=self.units.device)

    def desc(self, chunk_units=0):
        d = _lib.StreamDesc()
        d.units, d.n_units = self.units.data_ptr(), self.es.n_units
        d.out, d.n_out = self.out.data_ptr(), self.es.n_symbols
        d.table, d.L = self.table.data_ptr(), self.es.L
        d.first_state, d.first_bit = self.es.first_state, self.es.first_bit
        d.chunk_units = chunk_units
        return d

    def decode_range(self, first, last, index, chunk_units, index_host=None, stream=None):
        """Output symbols [first, last) (decode order) decoded alone from the sync points
        `index` (device tensor from a decode with index_out at this chunk length): the
        chunks that hold the window, starting at the recorded sync point of the first.
        Returns a new device tensor of last - first symbols."""
        torch = _torch()
        if index_host is None:
            index_host = index.cpu().numpy().view(_lib.SYNC_POINT_DTYPE)
        sub, skip = _lib.stream_range(self.desc(chunk_units), index_host, index.data_ptr(),
                                      first, last)
        out = torch.empty(max(sub.n_out, 1), dtype=torch.uint8, device=self.units.device)
        ws = torch.empty(workspace_bytes(sub.n_units, sub.L, chunk_units, 16), dtype=torch.uint8,
                         device=self.units.device)
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        a = _lib.DecodeArgs()
        a.units, a.n_units, a.out, a.n_out = sub.units, sub.n_units, out.data_ptr(), sub.n_out
        a.table, a.L, a.first_state, a.first_bit = sub.table, sub.L, sub.first_state, sub.first_bit
        a.chunk_units, a.checkpoint_units = chunk_units, 0
        a.workspace, a.workspace_bytes, a.stream = ws.data_ptr(), ws.numel(), st
        a.index_in = sub.index_in
        _lib.check(_lib.lib().mans_decode(C.byref(a)))
        self._range_ws = ws                 # kept alive until the next call
        return out[skip:skip + (last - first)]

    def status(self, stream=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        s = C.c_int()
        _lib.check(_lib.lib().mans_decode_status(C.c_void_p(self.ws.data_ptr()), C.c_void_p(st),
                                                 C.byref(s)))
        return s.value

    def lineage_stats(self, stream=None):
        """Control words of the last lineage decode on this stream's workspace: (fell back
        to the chunked pipeline, reason bits, lineages run, checkpoints they decoded)."""
        return lineage_stats(self.ws, stream)


def lineage_stats(ws, stream=None):
    torch = _torch()
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    out = (C.c_uint32 * 4)()
    _lib.check(_lib.lib().mans_lineage_stats(C.c_void_p(ws.data_ptr()), C.c_void_p(st), out))
    return tuple(int(v) for v in out)


@dataclass
class InterleavedEncoded:
    """A two-state interleaved stream (mans_encode_interleaved) on the host."""
    units: np.ndarray          # u32, decoder order
    first_bit: int
    first_states: np.ndarray   # u32[2], slot 0 and slot 1
    tables: tuple              # two u32[L] decoder tables (the same array when shared)
    L: int
    n_symbols: int
    record_units: int
    records: np.ndarray        # _lib.IL_RECORD_DTYPE

    @property
    def n_units(self):
        return len(self.units)


def encode_interleaved(symbols, L, record_units=256, per_slot_tables=False):
    """Tables (one shared from the whole histogram, or one per slot from the histogram of
    that slot's symbols: even input positions for slot 0, odd for slot 1) -> host encode
    with records."""
    symbols = np.ascontiguousarray(symbols, np.uint8)
    parts = [symbols[0::2], symbols[1::2]] if per_slot_tables else [symbols]
    encs, decs = [], []
    for p in parts:
        src = p if len(p) else symbols[:1]
        hist = np.bincount(src, minlength=256).astype(np.uint64)
        if not hist.any():
            hist[0] = 1
        enc, dec = _lib.build_tables(_lib.normalize_counts(hist, L), L)
        encs.append(enc)
        decs.append(dec)
    if len(encs) == 1:
        encs.append(encs[0])
        decs.append(decs[0])
    units, fb, fs, rec = _lib.encode_interleaved(symbols, encs[0], encs[1], L, record_units)
    return InterleavedEncoded(units, fb, fs, (decs[0], decs[1]), L, len(symbols), record_units, rec)


class DeviceInterleaved:
    """An interleaved stream resident in device memory; decode() runs
    mans_decode_interleaved into `out` (decode order)."""

    def __init__(self, ie: InterleavedEncoded, device="cuda"):
        torch = _torch()
        self.ie = ie
        self.units = torch.from_numpy(ie.units.view(np.int32)).to(device)
        self.tables = [torch.from_numpy(ie.tables[0].view(np.int32)).to(device)]
        if ie.tables[1] is ie.tables[0]:
            self.tables.append(self.tables[0])
        else:
            self.tables.append(torch.from_numpy(ie.tables[1].view(np.int32)).to(device))
        rec = ie.records if len(ie.records) else np.zeros(1, _lib.IL_RECORD_DTYPE)
        self.records = torch.from_numpy(rec.view(np.uint8).copy()).to(device)
        self.out = torch.zeros(max(ie.n_symbols, 1), dtype=torch.uint8, device=device)
        self.ws = torch.zeros(_lib.IL_WORKSPACE_BYTES, dtype=torch.uint8, device=device)

    def args(self, stream=None, flags=0, records=None):
        torch = _torch()
        a = _lib.IlArgs()
        a.units, a.n_units = self.units.data_ptr(), self.ie.n_units
        a.out, a.n_symbols = self.out.data_ptr(), self.ie.n_symbols
        a.tables[0], a.tables[1] = self.tables[0].data_ptr(), self.tables[1].data_ptr()
        a.L = self.ie.L
        a.first_states[0], a.first_states[1] = int(self.ie.first_states[0]), int(self.ie.first_states[1])
        a.first_bit = self.ie.first_bit
        a.record_units = self.ie.record_units
        a.records = (records if records is not None else self.records).data_ptr()
        a.flags = flags
        a.workspace = self.ws.data_ptr()
        a.stream = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        return a

    def decode(self, **kw):
        a = self.args(**kw)
        _lib.check(_lib.lib().mans_decode_interleaved(C.byref(a)))
        return self.out

    def status(self, stream=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        s = C.c_int()
        _lib.check(_lib.lib().mans_decode_status(C.c_void_p(self.ws.data_ptr()), C.c_void_p(st),
                                                 C.byref(s)))
        return s.value


def plan_windows(full_descs, records, windows):
    """Sub-stream descriptors for windows [(stream, first, last), ...] (decode order):
    full_descs[stream] has chunk_units and its sync points on the device in index_in,
    records(stream) returns the same records on the host.  Returns (descs, skip per
    window, output offset per window) for one mans_decode_batch over the sub-streams."""
    n = len(windows)
    subs = (_lib.StreamDesc * max(n, 1))()
    skips = np.zeros(n, np.int64)
    ooff = np.zeros(n + 1, np.int64)
    for k, (i, first, last) in enumerate(windows):
        rec = records(i)
        if rec is None or not full_descs[i].index_in:
            raise ValueError("stream %d has no sync points" % i)
        sub, skip = _lib.stream_range(full_descs[i], rec, full_descs[i].index_in, first, last)
        subs[k] = sub
        skips[k] = skip
        ooff[k + 1] = ooff[k] + sub.n_out
    return subs, skips, ooff


class WindowDecoder:
    """Decodes planned windows (plan_windows) with one mans_decode_batch call into an
    arena it keeps (and grows) on the device."""

    def __init__(self, device="cuda"):
        self.device = device
        self.arena = None
        self.ws = None

    def decode(self, planned, windows, stream=None, one_pipeline=False, trust=False):
        """Returns (arena, [(offset, length)]): window k at arena[offset:offset+length].
        trust: decode from the records without proving them (MANS_BATCH_TRUST_SYNC_POINTS)."""
        torch = _torch()
        subs, skips, ooff = planned
        n = len(windows)
        need = max(int(ooff[-1]), 1)
        if self.arena is None or self.arena.numel() < need:
            self.arena = torch.empty(need, dtype=torch.uint8, device=self.device)
        base = self.arena.data_ptr()
        for k in range(n):
            subs[k].out = base + int(ooff[k])
        b = C.c_uint64()
        _lib.check(_lib.lib().mans_batch_workspace_bytes(subs, n, C.byref(b)))
        if self.ws is None or self.ws.numel() < b.value:
            self.ws = torch.zeros(max(b.value, 256), dtype=torch.uint8, device=self.device)
        o = _lib.BatchOpts()
        o.workspace, o.workspace_bytes = self.ws.data_ptr(), self.ws.numel()
        o.stream = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        o.flags = (_lib.MANS_BATCH_ONE_PIPELINE if one_pipeline else 0) | \
            (_lib.MANS_BATCH_TRUST_SYNC_POINTS if trust else 0)
        stats = _lib.BatchStats()
        _lib.check(_lib.lib().mans_decode_batch(subs, n, C.byref(o), C.byref(stats)))
        spans = [(int(ooff[k] + skips[k]), int(w[2] - w[1])) for k, w in enumerate(windows)]
        return self.arena, spans

    def status(self, stream=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        s = C.c_int()
        _lib.check(_lib.lib().mans_decode_status(C.c_void_p(self.ws.data_ptr()), C.c_void_p(st),
                                                 C.byref(s)))
        return s.value


class DeviceBatch:
    """Many independent streams (each with its own table) resident on the device,
    decoded together by mans_decode_batch.  Units of every stream start 16-byte
    aligned and are padded to a multiple of four units; outputs are packed."""

    def __init__(self, streams, device="cuda", sync_points=False):
        """sync_points: give every stream a sync-point buffer (self.index, packed, stream i
        at record ioff[i]) at its tuned chunk length, for decode(index=...)."""
        torch = _torch()
        self.streams = list(streams)
        n = len(self.streams)
        uoff = np.zeros(n + 1, np.int64)
        toff = np.zeros(n + 1, np.int64)
        ooff = np.zeros(n + 1, np.int64)
        for i, es in enumerate(self.streams):
            uoff[i + 1] = uoff[i] + ((es.n_units + 3) // 4 + 1) * 4
            toff[i + 1] = toff[i] + es.L
            ooff[i + 1] = ooff[i] + es.n_symbols
        units = np.zeros(uoff[-1], np.uint32)
        tables = np.zeros(toff[-1], np.uint32)
        for i, es in enumerate(self.streams):
            units[uoff[i]:uoff[i] + es.n_units] = es.units
            tables[toff[i]:toff[i] + es.L] = es.table
        self.uoff, self.toff, self.ooff = uoff, toff, ooff
        self.units = torch.from_numpy(units.view(np.int32)).to(device)
        self.tables = torch.from_numpy(tables.view(np.int32)).to(device)
        self.out = torch.zeros(max(int(ooff[-1]), 1), dtype=torch.uint8, device=device)
        self.descs = (_lib.StreamDesc * n)()
        ub, tb, ob = self.units.data_ptr(), self.tables.data_ptr(), self.out.data_ptr()
        for i, es in enumerate(self.streams):
            d = self.descs[i]
            d.units = ub + 4 * int(uoff[i])
            d.n_units = es.n_units
            d.out = ob + int(ooff[i])
            d.n_out = es.n_symbols
            d.table = tb + 4 * int(toff[i])
            d.L = es.L
            d.first_state = es.first_state
            d.first_bit = es.first_bit
        self.device = device
        self.index = None
        self._idescs = {}
        self.ws = None
        if sync_points:
            self.use_index()
        else:
            self._fit_workspace([self.descs])

    def _fit_workspace(self, desc_sets):
        torch = _torch()
        b = C.c_uint64()
        need = 0
        for ds in desc_sets:
            _lib.check(_lib.lib().mans_batch_workspace_bytes(ds, len(self.streams), C.byref(b)))
            need = max(need, b.value)
        if self.ws is None or self.ws.numel() < need:
            self.ws = None
            self.ws = torch.zeros(need, dtype=torch.uint8, device=self.device)

    def use_index(self, chunk_units=None):
        """(Re)build the sync-point buffer and descriptors: every stream at its tuned
        chunk length, or at `chunk_units` (a power of two >= 16) for all."""
        torch = _torch()
        n = len(self.streams)
        cu = np.zeros(n, np.int64)
        ioff = np.zeros(n + 1, np.int64)
        for i, es in enumerate(self.streams):
            if chunk_units:
                cu[i] = chunk_units
            else:
                c = C.c_uint32()
                _lib.check(_lib.lib().mans_chunk_units(es.n_units, es.n_symbols, es.L,
                                                       C.byref(c)))
                cu[i] = max(c.value, 16)
            ioff[i + 1] = ioff[i] + _lib.sync_point_count(es.n_units, int(cu[i]))
        self.chunk_units, self.ioff = cu, ioff
        self.index = None
        self.index = torch.zeros(max(int(ioff[-1]), 1) * 16, dtype=torch.uint8, device=self.device)
        ib = self.index.data_ptr()
        self._idescs = {}
        for mode in ("out", "in", "both"):
            ds = (_lib.StreamDesc * n)()
            C.memmove(ds, self.descs, C.sizeof(ds))
            for i in range(n):
                ds[i].chunk_units = int(cu[i])
                p = ib + 16 * int(ioff[i])
                if mode in ("in", "both"):
                    ds[i].index_in = p
                if mode in ("out", "both"):
                    ds[i].index_out = p
            self._idescs[mode] = ds
        self._fit_workspace([self.descs] + list(self._idescs.values()))

    def use_items(self):
        """Record buffer at 32-unit granularity (self.items, stream i at record
        item_off[i]) and the descriptors "items_out" (a decode at the tuned chunk lengths
        that writes them) and "trusted" (the write pass alone from them, decode(index=
        "trusted", trust=True))."""
        torch = _torch()
        n = len(self.streams)
        ioff = np.zeros(n + 1, np.int64)
        for i, es in enumerate(self.streams):
            ioff[i + 1] = ioff[i] + _lib.sync_point_count(es.n_units, 32)
        self.item_off = ioff
        self.items = None
        self.items = torch.zeros(max(int(ioff[-1]), 1) * 16, dtype=torch.uint8, device=self.device)
        ib = self.items.data_ptr()
        for mode in ("items_out", "trusted"):
            ds = (_lib.StreamDesc * n)()
            C.memmove(ds, self.descs, C.sizeof(ds))
            for i in range(n):
                p = ib + 16 * int(ioff[i])
                if mode == "items_out":
                    ds[i].items_out = p
                else:
                    ds[i].chunk_units = 32
                    ds[i].index_in = p
            self._idescs[mode] = ds
        self._fit_workspace([self.descs] + list(self._idescs.values()))

    def sync_points(self, i):
        """Host copy of stream i's sync points (SYNC_POINT_DTYPE records)."""
        a, b = int(self.ioff[i]), int(self.ioff[i + 1])
        return self.index[16 * a:16 * b].cpu().numpy().view(_lib.SYNC_POINT_DTYPE)

    def plan_windows(self, windows, index_host=None, items=False):
        """plan_windows over this batch's streams (after a decode with index="out", or
        with items: index="items_out"); index_host: host copy of the records (copied here
        if not given)."""
        src = self.items if items else self.index
        if index_host is None:
            index_host = src.cpu().numpy().view(_lib.SYNC_POINT_DTYPE)
        ioff = self.item_off if items else self.ioff
        descs = self._idescs["trusted" if items else "in"]
        return plan_windows(descs, lambda i: index_host[ioff[i]:ioff[i + 1]], windows)

    def decode(self, stream=None, kernel_events=None, one_pipeline=False, index=None,
               trust=False):
        """one_pipeline: no split into concurrent parts (MANS_BATCH_ONE_PIPELINE).
        index (batches built with sync_points): "out" writes every stream's sync points,
        "in" starts the decode from them, "both" does both.  trust (with index="in"):
        one pass from the records, unproven (MANS_BATCH_TRUST_SYNC_POINTS)."""
        torch = _torch()
        descs = self.descs if index is None else self._idescs[index]
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        o = _lib.BatchOpts()
        o.kernel_events = kernel_events
        o.flags = (_lib.MANS_BATCH_ONE_PIPELINE if one_pipeline else 0) | \
            (_lib.MANS_BATCH_TRUST_SYNC_POINTS if trust else 0)
        o.workspace = self.ws.data_ptr()
        o.workspace_bytes = self.ws.numel()
        o.stream = st
        stats = _lib.BatchStats()
        _lib.check(_lib.lib().mans_decode_batch(descs, len(self.streams), C.byref(o),
                                                C.byref(stats)))
        return stats

    def graph(self, index=None):
        """Capture this batch's decode into a hipGraph (see mans_batch_graph_create);
        returns a BatchGraph whose launch() replays it.  index as in decode()."""
        descs = self.descs if index is None else self._idescs[index]
        o = _lib.BatchOpts()
        o.workspace = self.ws.data_ptr()
        o.workspace_bytes = self.ws.numel()
        h = C.c_void_p()
        _lib.check(_lib.lib().mans_batch_graph_create(descs, len(self.streams), C.byref(o),
                                                      C.byref(h)))
        return BatchGraph(h, self)

    def plain_args(self, i, ws, stream=None):
        """mans_decode arguments for stream i alone (per-stream plain launch), using the
        batch's device buffers and the caller's workspace tensor."""
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        d = self.descs[i]
        a = _lib.DecodeArgs()
        a.units, a.n_units, a.out, a.n_out = d.units, d.n_units, d.out, d.n_out
        a.table, a.L, a.first_state, a.first_bit = d.table, d.L, d.first_state, d.first_bit
        a.workspace, a.workspace_bytes, a.stream = ws.data_ptr(), ws.numel(), st
        return a

    def output(self, i):
        return self.out[int(self.ooff[i]):int(self.ooff[i + 1])]

    def status(self, stream=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        s = C.c_int()
        _lib.check(_lib.lib().mans_decode_status(C.c_void_p(self.ws.data_ptr()), C.c_void_p(st),
                                                 C.byref(s)))
        return s.value


class BatchGraph:
    """An instantiated hipGraph of one batch's decode; keeps the batch's buffers alive."""

    def __init__(self, handle, batch):
        self.h = handle
        self.batch = batch

    def launch(self, stream=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().mans_graph_launch(self.h, C.c_void_p(st)))

    def close(self):
        if self.h:
            _lib.lib().mans_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GeneratedBatch:
    """Streams made on the device alone, for workloads too large to generate and encode on
    the host quickly: symbols (mans_gen_symbols_device, the same bytes as the host
    generator), decoder tables from their histograms (mans_build_tables_device, the host
    codec's bytes) and units (mans_encode_device, unit-exact against the host codec).
    specs: [(seed, lambda, n_symbols, L)].  Buffers are packed, each stream 16-byte
    aligned; `symbols(i)` is stream i in encoder order (the decoder emits it reversed)."""

    def __init__(self, specs, device="cuda"):
        torch = _torch()
        self.specs = list(specs)
        n = len(self.specs)
        self.n = n
        L = np.array([sp[3] for sp in self.specs], np.int64)
        nsym = np.array([sp[2] for sp in self.specs], np.int64)
        cap = np.array([(_lib.lib().mans_max_encoded_units(int(k)) + 3) // 4 * 4 + 4
                        for k in nsym], np.int64)
        self.soff = np.concatenate([[0], np.cumsum((nsym + 15) // 16 * 16)])
        self.toff = np.concatenate([[0], np.cumsum(L)])
        self.uoff = np.concatenate([[0], np.cumsum(cap)])
        self.L, self.n_symbols = L, nsym
        self.sym = torch.empty(max(int(self.soff[-1]), 16), dtype=torch.uint8, device=device)
        self.tables = torch.zeros(max(int(self.toff[-1]), 1), dtype=torch.int32, device=device)
        self.units = torch.zeros(max(int(self.uoff[-1]), 4), dtype=torch.int32, device=device)
        for i, (seed, lam, k, _) in enumerate(self.specs):
            if k:
                _lib.gen_symbols_device(seed, lam, self.symbols(i))
        jobs = [(self.symbols(i), int(L[i]), self.table(i), None) for i in range(n)]
        st = _lib.build_tables_device(jobs)
        if any(st):
            raise RuntimeError("device tables: status %s" % sorted(set(st)))
        ej = [(self.symbols(i), self.table(i), int(L[i]),
               self.units[int(self.uoff[i]):int(self.uoff[i + 1])]) for i in range(n)]
        res = _lib.encode_device(ej)
        if any(r[3] for r in res):
            raise RuntimeError("device encode: status %s" % sorted(set(r[3] for r in res)))
        self.n_units = np.array([r[0] for r in res], np.int64)
        self.first_bit = np.array([r[1] for r in res], np.int64)
        self.first_state = np.array([r[2] for r in res], np.int64)
        self.device = device

    def symbols(self, i):
        return self.sym[int(self.soff[i]):int(self.soff[i]) + int(self.n_symbols[i])]

    def table(self, i):
        return self.tables[int(self.toff[i]):int(self.toff[i + 1])]

    def stream_units(self, i):
        return self.units[int(self.uoff[i]):int(self.uoff[i]) + int(self.n_units[i])]

    def encoded(self, i):
        """Host copy of stream i as an EncodedStream (for the oracle)."""
        return EncodedStream(self.stream_units(i).cpu().numpy().view(np.uint32).copy(),
                             int(self.first_bit[i]), int(self.first_state[i]),
                             self.table(i).cpu().numpy().view(np.uint32).copy(),
                             int(self.L[i]), int(self.n_symbols[i]))

    def descs(self, out, ooff, index=None):
        """StreamDesc array decoding stream i into out[ooff[i]:ooff[i + 1]]."""
        d = (_lib.StreamDesc * self.n)()
        ob = out.data_ptr()
        for i in range(self.n):
            x = d[i]
            x.units = self.units.data_ptr() + 4 * int(self.uoff[i])
            x.n_units = int(self.n_units[i])
            x.out = ob + int(ooff[i])
            x.n_out = int(self.n_symbols[i])
            x.table = self.tables.data_ptr() + 4 * int(self.toff[i])
            x.L = int(self.L[i])
            x.first_state = int(self.first_state[i])
            x.first_bit = int(self.first_bit[i])
        return d

    def decode_one(self, i, out, ws, stream=None, **kw):
        """mans_decode of stream i into the u8 tensor out with workspace tensor ws."""
        torch = _torch()
        a = _lib.DecodeArgs()
        a.units = self.units.data_ptr() + 4 * int(self.uoff[i])
        a.n_units = int(self.n_units[i])
        a.out = out.data_ptr()
        a.n_out = int(self.n_symbols[i])
        a.table = self.tables.data_ptr() + 4 * int(self.toff[i])
        a.L = int(self.L[i])
        a.first_state = int(self.first_state[i])
        a.first_bit = int(self.first_bit[i])
        a.workspace = ws.data_ptr()
        a.workspace_bytes = ws.numel()
        a.stream = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        a.kernel_events = kw.get("kernel_events")
        a.chunk_units = kw.get("chunk_units", 0)
        idx = kw.get("index_out")
        a.index_out = idx.data_ptr() if idx is not None else None
        _lib.check(_lib.lib().mans_decode(C.byref(a)))
        return out

    def status(self, ws, stream=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        s = C.c_int()
        _lib.check(_lib.lib().mans_decode_status(C.c_void_p(ws.data_ptr()), C.c_void_p(st),
                                                 C.byref(s)))
        return s.value
