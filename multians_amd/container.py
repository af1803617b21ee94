"""Batch container (wire format, include/mans_container.h) over the C ABI.

`pack` writes many encoded streams, their tables and optional sync points into one
image; `Container` validates an image on the host; `DeviceContainer` copies it to the
device once and decodes every stream with one mans_decode_batch call, or windows of its
streams alone from the stored sync points (mans_stream_range), many windows per call.
"""
import ctypes as C

import numpy as np

from . import _lib
from .decoder import WindowDecoder, plan_windows

EXPORTS = ["mans_container_bytes", "mans_container_pack", "mans_container_open",
           "mans_container_descs", "mans_container_sync_points"]


class ContainerTable(C.Structure):
    _fields_ = [("words", C.c_void_p), ("L", C.c_uint32), ("reserved", C.c_uint32)]


class ContainerStream(C.Structure):
    _fields_ = [("units", C.c_void_p), ("n_units", C.c_uint64), ("n_symbols", C.c_uint64),
                ("table", C.c_uint32), ("first_state", C.c_uint32), ("first_bit", C.c_uint32),
                ("chunk_units", C.c_uint32), ("sync_points", C.c_void_p)]


class ContainerInfo(C.Structure):
    _fields_ = [("n_streams", C.c_uint32), ("n_tables", C.c_uint32),
                ("n_with_sync_points", C.c_uint32), ("reserved", C.c_uint32),
                ("total_units", C.c_uint64), ("total_symbols", C.c_uint64),
                ("image_bytes", C.c_uint64)]


_done = False


def lib():
    global _done
    l = _lib.lib()
    if not _done:
        vp, u64, u32, i32, sz = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_size_t
        sig = {
            "mans_container_bytes": ([C.POINTER(ContainerTable), sz, C.POINTER(ContainerStream),
                                      sz, C.POINTER(u64)], i32),
            "mans_container_pack": ([C.POINTER(ContainerTable), sz, C.POINTER(ContainerStream),
                                     sz, vp, u64], i32),
            "mans_container_open": ([vp, u64, C.POINTER(ContainerInfo)], i32),
            "mans_container_descs": ([vp, u64, vp, vp, u64, i32, C.POINTER(_lib.StreamDesc), sz,
                                      vp], i32),
            "mans_container_sync_points": ([vp, u64, u32, C.POINTER(u64)], vp),
        }
        for name, (args, res) in sig.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = res
        _done = True
    return l


def pack(streams, sync_points=None, chunk_units=None):
    """Image (np.uint8) holding the EncodedStreams `streams`, one table each (identical
    tables are stored once), with sync_points[i] (SYNC_POINT_DTYPE records at chunk length
    chunk_units[i]) where given."""
    n = len(streams)
    keys, tabs, tid = {}, [], []
    for es in streams:
        t = np.ascontiguousarray(es.table, np.uint32)
        k = (es.L, t.tobytes())
        if k not in keys:
            keys[k] = len(tabs)
            tabs.append(t)
        tid.append(keys[k])
    ct = (ContainerTable * max(len(tabs), 1))()
    for i, t in enumerate(tabs):
        ct[i].words = t.ctypes.data
        ct[i].L = len(t)
    keep = []
    cs = (ContainerStream * max(n, 1))()
    for i, es in enumerate(streams):
        u = np.ascontiguousarray(es.units, np.uint32)
        keep.append(u)
        s = cs[i]
        s.units, s.n_units, s.n_symbols = u.ctypes.data, len(u), es.n_symbols
        s.table, s.first_state, s.first_bit = tid[i], es.first_state, es.first_bit
        if sync_points is not None and sync_points[i] is not None:
            r = np.ascontiguousarray(sync_points[i], _lib.SYNC_POINT_DTYPE)
            keep.append(r)
            s.sync_points = r.ctypes.data
            s.chunk_units = int(chunk_units[i])
    b = C.c_uint64()
    _lib.check(lib().mans_container_bytes(ct, len(tabs), cs, n, C.byref(b)))
    img = np.zeros(b.value, np.uint8)
    _lib.check(lib().mans_container_pack(ct, len(tabs), cs, n, img.ctypes.data, b.value))
    return img


class Container:
    """A validated container image (host copy)."""

    def __init__(self, image):
        self.image = np.ascontiguousarray(np.frombuffer(image, np.uint8) if isinstance(
            image, (bytes, bytearray)) else image, np.uint8)
        info = ContainerInfo()
        _lib.check(lib().mans_container_open(self.image.ctypes.data, len(self.image),
                                             C.byref(info)))
        self.info = info

    def sync_points(self, i):
        """Stream i's stored sync points (a view into the image), or None."""
        cnt = C.c_uint64()
        p = lib().mans_container_sync_points(self.image.ctypes.data, len(self.image), i,
                                             C.byref(cnt))
        if not p:
            return None
        off = p - self.image.ctypes.data
        return self.image[off:off + 16 * cnt.value].view(_lib.SYNC_POINT_DTYPE)

    def descs(self, device_image_ptr, device_out_ptr, out_bytes, use_sync_points=True):
        n = self.info.n_streams
        ds = (_lib.StreamDesc * max(n, 1))()
        offs = np.zeros(n + 1, np.uint64)
        _lib.check(lib().mans_container_descs(self.image.ctypes.data, len(self.image),
                                              device_image_ptr, device_out_ptr, out_bytes,
                                              1 if use_sync_points else 0, ds, n,
                                              offs.ctypes.data))
        return ds, offs


class DeviceContainer:
    """A container resident on the device (one copy of the image) with an output arena
    for all of its streams."""

    def __init__(self, container: Container, device="cuda"):
        import torch
        self.c = container
        self.image = torch.from_numpy(container.image).to(device)
        self.out = torch.zeros(max(container.info.total_symbols, 1), dtype=torch.uint8,
                               device=device)
        self.device = device
        self._win = None
        self._ws = {}
        self.descs = {}
        for use in (True, False):
            ds, offs = container.descs(self.image.data_ptr(), self.out.data_ptr(),
                                       self.out.numel(), use)
            self.descs[use] = ds
            self.offs = offs

    def _workspace(self, key, descs, n):
        import torch
        b = C.c_uint64()
        _lib.check(_lib.lib().mans_batch_workspace_bytes(descs, n, C.byref(b)))
        ws = self._ws.get(key)
        if ws is None or ws.numel() < b.value:
            ws = torch.zeros(max(b.value, 256), dtype=torch.uint8, device=self.device)
            self._ws[key] = ws
        return ws

    def decode(self, use_sync_points=True, stream=None, one_pipeline=False, trust=False):
        """Decode every stream (mans_decode_batch); returns the batch stats.  trust: the
        stored records are 32-unit item records to decode from without proof
        (MANS_BATCH_TRUST_SYNC_POINTS; every stream must hold them)."""
        import torch
        n = self.c.info.n_streams
        ds = self.descs[use_sync_points]
        ws = self._workspace(("all", use_sync_points), ds, n)
        o = _lib.BatchOpts()
        o.workspace, o.workspace_bytes = ws.data_ptr(), ws.numel()
        o.stream = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        o.flags = (_lib.MANS_BATCH_ONE_PIPELINE if one_pipeline else 0) | \
            (_lib.MANS_BATCH_TRUST_SYNC_POINTS if trust else 0)
        stats = _lib.BatchStats()
        _lib.check(_lib.lib().mans_decode_batch(ds, n, C.byref(o), C.byref(stats)))
        return stats

    def output(self, i):
        return self.out[int(self.offs[i]):int(self.offs[i + 1])]

    def status(self, key=("all", True), stream=None):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        s = C.c_int()
        _lib.check(_lib.lib().mans_decode_status(C.c_void_p(self._ws[key].data_ptr()),
                                                 C.c_void_p(st), C.byref(s)))
        return s.value

    def plan_windows(self, windows):
        """decoder.plan_windows over this container's streams (stored sync points)."""
        return plan_windows(self.descs[True], self.c.sync_points, windows)

    def decode_windows(self, windows, stream=None, one_pipeline=False, planned=None,
                       trust=False):
        """Decode windows [(stream, first, last), ...] (decode order) in one batch call,
        each from the stored sync point of the chunk holding its first symbol.  Returns
        (arena, [(offset, length)]) with window k at arena[offset:offset + length]."""
        if self._win is None:
            self._win = WindowDecoder(self.device)
        planned = planned if planned is not None else self.plan_windows(windows)
        return self._win.decode(planned, windows, stream, one_pipeline, trust)
