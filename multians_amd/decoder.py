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
             kernel_events=None):
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
        return a

    def decode(self, **kw):
        a = self.args(**kw)
        _lib.check(_lib.lib().mans_decode(C.byref(a)))
        return self.out

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

    def __init__(self, streams, device="cuda"):
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
        b = C.c_uint64()
        _lib.check(_lib.lib().mans_batch_workspace_bytes(self.descs, n, C.byref(b)))
        self.ws = torch.zeros(b.value, dtype=torch.uint8, device=device)

    def decode(self, stream=None, kernel_events=None, one_pipeline=False):
        """one_pipeline: no split into concurrent parts (MANS_BATCH_ONE_PIPELINE)."""
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        o = _lib.BatchOpts()
        o.kernel_events = kernel_events
        o.flags = _lib.MANS_BATCH_ONE_PIPELINE if one_pipeline else 0
        o.workspace = self.ws.data_ptr()
        o.workspace_bytes = self.ws.numel()
        o.stream = st
        stats = _lib.BatchStats()
        _lib.check(_lib.lib().mans_decode_batch(self.descs, len(self.streams), C.byref(o),
                                                C.byref(stats)))
        return stats

    def graph(self):
        """Capture this batch's decode into a hipGraph (see mans_batch_graph_create);
        returns a BatchGraph whose launch() replays it."""
        o = _lib.BatchOpts()
        o.workspace = self.ws.data_ptr()
        o.workspace_bytes = self.ws.numel()
        h = C.c_void_p()
        _lib.check(_lib.lib().mans_batch_graph_create(self.descs, len(self.streams), C.byref(o),
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
