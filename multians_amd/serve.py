"""Continuous-batching decode server over mans_decode_batch.

Requests are decodes of tANS streams (each with its own table).  The loop admits arrived
requests under an HBM budget, coalesces everything admitted since the last launch into the
next batched launch, and keeps up to `max_inflight` batches on the device at once, each on
its own HIP stream, so a batch that arrives while others run overlaps them.  A request's
latency runs from its arrival to the moment its batch's completion event is observed on
the host.

Two input modes:
  * host inputs (StreamPool(host_inputs=True), the served configuration): a request brings
    its compressed stream and table from host memory (the pool plays the clients: pinned
    host buffers, one per stream).  On admission they are copied to the device on a copy
    stream, into the request's block of the device arena, and the batch's decode waits
    for the copy; nothing of a stream stays on the device after its request.  The budget
    covers inputs, outputs, the batch workspace and status words, and latency includes
    the copy.
  * resident inputs (the default): compressed streams and tables stay on the device; the
    budget covers what a request holds while admitted (output, workspace share).
"""
import bisect
import ctypes as C
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib

_ALIGN = 256


def _up(v, a=_ALIGN):
    return (v + a - 1) // a * a


class DeviceArena:
    """First-fit allocator over one device byte tensor (offsets, 256-byte aligned)."""

    def __init__(self, nbytes, device="cuda"):
        import torch
        self.size = _up(int(nbytes))
        self.buf = torch.empty(self.size, dtype=torch.uint8, device=device)
        self.base = self.buf.data_ptr()
        self.free_off = [0]           # sorted offsets of free blocks
        self.free_len = {0: self.size}
        self.used = 0
        self.peak = 0

    def alloc(self, n):
        n = _up(max(int(n), 1))
        for i, off in enumerate(self.free_off):
            ln = self.free_len[off]
            if ln >= n:
                del self.free_off[i]
                del self.free_len[off]
                if ln > n:
                    self.free_off.insert(i, off + n)
                    self.free_len[off + n] = ln - n
                self.used += n
                self.peak = max(self.peak, self.used)
                return off, n
        return None

    def free(self, off, n):
        self.used -= n
        i = bisect.bisect_left(self.free_off, off)
        # merge with the following and the preceding free block
        if i < len(self.free_off) and off + n == self.free_off[i]:
            nxt = self.free_off.pop(i)
            n += self.free_len.pop(nxt)
        if i > 0 and self.free_off[i - 1] + self.free_len[self.free_off[i - 1]] == off:
            prev = self.free_off[i - 1]
            self.free_len[prev] += n
            return
        self.free_off.insert(i, off)
        self.free_len[off] = n

    def largest_free(self):
        return max(self.free_len.values()) if self.free_len else 0


class StreamPool:
    """Encoded streams resident in device memory: packed units (each stream 16-byte
    aligned, padded to whole 16-byte blocks), packed tables, and optionally the
    original symbols (decode order) for checking results."""

    def __init__(self, streams, symbols=None, device="cuda", records=False, host_inputs=False):
        """records: keep a buffer of 32-unit item records per stream (written by the
        stream's first decode, read by later ones; Server(use_records=True)).
        host_inputs: keep the streams in pinned host memory instead, each as one image
        [table | units] (256-byte aligned), copied to the device per request."""
        import torch
        self.streams = list(streams)
        n = len(self.streams)
        self.n = n
        self.host_inputs = bool(host_inputs)
        if self.host_inputs and records:
            raise ValueError("item records need resident inputs")
        self.n_units = np.array([es.n_units for es in self.streams], np.int64)
        self.L = np.array([es.L for es in self.streams], np.int64)
        if self.host_inputs:
            # per stream: table (4L bytes), then units padded to whole 16-byte blocks
            self.in_bytes = np.array([_up(4 * es.L + 16 * ((es.n_units + 3) // 4 + 1))
                                      for es in self.streams], np.int64)
            ioff = np.concatenate([[0], np.cumsum(self.in_bytes)])
            self.inputs = torch.zeros(max(int(ioff[-1]), 1), dtype=torch.uint8, pin_memory=True)
            host = self.inputs.numpy()
            for i, es in enumerate(self.streams):
                o = int(ioff[i])
                host[o:o + 4 * es.L] = es.table.view(np.uint8)
                host[o + 4 * es.L:o + 4 * es.L + 4 * es.n_units] = es.units.view(np.uint8)
            self.in_ptr = self.inputs.data_ptr() + ioff[:-1]
            self.uptr = self.tptr = None
        else:
            uoff = np.zeros(n + 1, np.int64)
            toff = np.zeros(n + 1, np.int64)
            for i, es in enumerate(self.streams):
                uoff[i + 1] = uoff[i] + ((es.n_units + 3) // 4 + 1) * 4
                toff[i + 1] = toff[i] + es.L
            units = np.zeros(uoff[-1], np.uint32)
            tables = np.zeros(toff[-1], np.uint32)
            for i, es in enumerate(self.streams):
                units[uoff[i]:uoff[i] + es.n_units] = es.units
                tables[toff[i]:toff[i] + es.L] = es.table
            self.units = torch.from_numpy(units.view(np.int32)).to(device)
            self.tables = torch.from_numpy(tables.view(np.int32)).to(device)
            del units, tables
            self.uptr = self.units.data_ptr() + 4 * uoff[:-1]
            self.tptr = self.tables.data_ptr() + 4 * toff[:-1]
        self.n_sym = np.array([es.n_symbols for es in self.streams], np.int64)
        self.first_state = np.array([es.first_state for es in self.streams], np.int64)
        self.first_bit = np.array([es.first_bit for es in self.streams], np.int64)
        self.have_rec = np.zeros(n, bool)
        self.rec = None
        self.device = device
        if records:
            self.attach_records()
        self.sym = None
        if symbols is not None:
            soff = np.zeros(n + 1, np.int64)
            soff[1:] = np.cumsum(self.n_sym)
            host = np.empty(int(soff[-1]), np.uint8)
            for i, s in enumerate(symbols):
                host[soff[i]:soff[i + 1]] = s[::-1]          # decode order
            self.sym = torch.from_numpy(host).to(device)
            self.soff = soff

    def attach_records(self):
        """Allocate the item-record buffer (empty: no stream has records yet)."""
        import torch
        roff = np.zeros(self.n + 1, np.int64)
        for i, es in enumerate(self.streams):
            roff[i + 1] = roff[i] + _lib.sync_point_count(es.n_units, 32)
        self.rec = torch.zeros(max(int(roff[-1]), 1) * 16, dtype=torch.uint8, device=self.device)
        self.rptr = self.rec.data_ptr() + 16 * roff[:-1]
        self.have_rec[:] = False

    def expected(self, i):
        return self.sym[int(self.soff[i]):int(self.soff[i + 1])]


@dataclass
class Request:
    rid: int
    stream: int           # index into the pool
    t_arrival: float      # seconds from the start of the run
    t_admit: float = 0.0
    t_done: float = 0.0
    out_off: int = 0      # arena offset of the output
    in_off: int = 0       # arena offset of the input image (host inputs)
    status: int = -1      # the decoder's status word for this request (0 = ok)


@dataclass
class _Batch:
    reqs: list
    block: tuple
    event: object
    hip_stream: object
    descs: object
    t_launch: float
    status: object = None     # pinned host copy of every request's status word
    check: object = None      # pinned host copy of the checked requests' mismatch flags
    check_idx: list = None    # positions in reqs of the checked requests
    fresh: list = None        # (position in reqs, pool stream) whose item records it writes


@dataclass
class ServeStats:
    latencies_ms: np.ndarray
    symbols: int
    wall_s: float
    n_batches: int
    batch_sizes: list = field(default_factory=list)
    peak_bytes: int = 0
    checked: int = 0
    mismatches: int = -1
    failed: int = 0           # requests whose decode reported a nonzero status

    def summary(self):
        lat = self.latencies_ms
        return {
            "requests": int(len(lat)),
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
            "max_ms": round(float(lat.max()), 3) if len(lat) else None,
            "decoded_GBps": round(self.symbols / self.wall_s / 1e9, 3),
            "batches": self.n_batches,
            "mean_batch": round(float(np.mean(self.batch_sizes)), 1) if self.batch_sizes else 0,
            "peak_hbm_GB": round(self.peak_bytes / 1e9, 3),
            "checked": self.checked,
            "mismatches": self.mismatches,
            "failed": self.failed,
        }


class Server:
    """Continuous-batching decode loop over a StreamPool."""

    def __init__(self, pool, budget_bytes, max_inflight=3, max_batch=4096, device="cuda",
                 use_records=False):
        """use_records (pool built with records=True): a stream's first decode also
        writes its item records; later requests for it run the write pass alone from
        them (MANS_BATCH_TRUST_SYNC_POINTS), as a second launch of their batch."""
        import torch
        self.torch = torch
        self.pool = pool
        self.use_records = bool(use_records and pool.rec is not None)
        self.arena = DeviceArena(budget_bytes, device)
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        self.streams = [torch.cuda.Stream(device=device) for _ in range(max_inflight)]
        self.host = pool.host_inputs
        # host inputs: copies on their own stream (one link; batches queue there in order)
        self.copy_stream = torch.cuda.Stream(device=device) if self.host else None
        self.lib = _lib.lib()

    def _groups(self, reqs):
        """The batch's launches: [(requests, trusted)], trusted = decode from records."""
        if not self.use_records:
            return [(reqs, False)]
        have = self.pool.have_rec
        cached = [r for r in reqs if have[r.stream]]
        fresh = [r for r in reqs if not have[r.stream]]
        return [(g, t) for g, t in ((fresh, False), (cached, True)) if g]

    def _batch_block(self, reqs):
        """Bytes of one batch's block: outputs (256-aligned each) + workspace (the
        largest of its launches', which run one after the other on one stream)."""
        p = self.pool
        out_bytes = 0
        for r in reqs:
            r.out_off = out_bytes
            out_bytes += _up(int(p.n_sym[r.stream]))
        in_bytes = 0
        if self.host:
            for r in reqs:
                r.in_off = out_bytes + in_bytes
                in_bytes += int(p.in_bytes[r.stream])
        launches = []
        ws = 0
        for group, trusted in self._groups(reqs):
            n = len(group)
            descs = (_lib.StreamDesc * n)()
            for j, r in enumerate(group):
                i = r.stream
                d = descs[j]
                # workspace sizing only reads sizes and pointers: out (and, with host
                # inputs, units and table) are set at launch
                d.units = 1 << 12 if self.host else int(p.uptr[i])
                d.n_units = int(p.n_units[i])
                d.out = 1 << 12
                d.n_out = int(p.n_sym[i])
                d.table = 1 << 12 if self.host else int(p.tptr[i])
                d.L = int(p.L[i])
                d.first_state = int(p.first_state[i])
                d.first_bit = int(p.first_bit[i])
                if self.use_records:
                    if trusted:
                        d.chunk_units = 32
                        d.index_in = int(p.rptr[i])
                    else:
                        d.items_out = int(p.rptr[i])
            b = C.c_uint64()
            _lib.check(self.lib.mans_batch_workspace_bytes(descs, n, C.byref(b)))
            ws = max(ws, int(b.value))
            launches.append((group, trusted, descs))
        return launches, out_bytes + in_bytes, ws

    def _launch(self, reqs, t_now, hs, check_every):
        torch = self.torch
        launches, out_bytes, ws_bytes = self._batch_block(reqs)
        # the block also holds one status word per request (the decoder writes them)
        st_bytes = _up(4 * len(reqs))
        blk = self.arena.alloc(out_bytes + ws_bytes + st_bytes)
        if blk is None:
            return None
        off, nb = blk
        base = self.arena.base + off
        st_off = off + out_bytes + ws_bytes
        for r in reqs:
            r.out_off += off
        fresh = []
        g0 = 0                                  # first status word of the group
        if self.host:
            # the requests' inputs: pinned host -> their block, on the copy stream; the
            # decode stream waits for them
            cs = self.copy_stream
            for r in reqs:
                r.in_off += off
                _lib.check(self.lib.mans_memcpy_h2d(
                    C.c_void_p(self.arena.base + r.in_off),
                    C.c_void_p(int(self.pool.in_ptr[r.stream])),
                    int(self.pool.in_bytes[r.stream]), C.c_void_p(cs.cuda_stream)))
            copied = self.torch.cuda.Event()
            copied.record(cs)
            hs.wait_event(copied)
        for group, trusted, descs in launches:
            for j, r in enumerate(group):
                descs[j].out = self.arena.base + r.out_off
                if self.host:
                    descs[j].table = self.arena.base + r.in_off
                    descs[j].units = self.arena.base + r.in_off + 4 * int(self.pool.L[r.stream])
            o = _lib.BatchOpts()
            o.workspace = base + out_bytes
            o.workspace_bytes = ws_bytes
            o.stream = hs.cuda_stream
            o.stream_status = self.arena.base + st_off + 4 * g0
            # one pipeline per batch: the server already keeps max_inflight batches in
            # flight on its own streams
            o.flags = _lib.MANS_BATCH_ONE_PIPELINE | \
                (_lib.MANS_BATCH_TRUST_SYNC_POINTS if trusted else 0)
            _lib.check(self.lib.mans_decode_batch(descs, len(group), C.byref(o), None))
            if self.use_records and not trusted:
                fresh.extend((g0 + j, r.stream) for j, r in enumerate(group))
            g0 += len(group)
        descs = [d for _, _, d in launches]
        order = [r for g, _, _ in launches for r in g]       # status word order
        status = torch.empty(len(reqs), dtype=torch.int32, pin_memory=True)
        check, check_idx = None, None
        with torch.cuda.stream(hs):
            words = self.arena.buf[st_off:st_off + 4 * len(reqs)].view(torch.int32)
            status.copy_(words, non_blocking=True)
            if check_every and self.pool.sym is not None:
                flags, check_idx = [], []
                for k, r in enumerate(order):
                    if r.rid % check_every:
                        continue
                    n = int(self.pool.n_sym[r.stream])
                    got = self.arena.buf[r.out_off:r.out_off + n]
                    flags.append((got != self.pool.expected(r.stream)).any())
                    check_idx.append(k)
                if flags:
                    check = torch.empty(len(flags), dtype=torch.bool, pin_memory=True)
                    check.copy_(torch.stack(flags), non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(hs)
        return _Batch(order, blk, ev, hs, descs, t_now, status, check, check_idx, fresh)

    def run(self, arrivals, stream_ids, check_every=0, poll_sleep=20e-6):
        """arrivals: sorted arrival times (s); stream_ids: pool index per request.
        Runs until every request is done.  Returns ServeStats."""
        torch = self.torch
        reqs = [Request(k, int(s), float(t)) for k, (t, s) in enumerate(zip(arrivals, stream_ids))]
        pending = deque()
        inflight = []
        free_streams = list(self.streams)
        nxt = 0
        done = []
        sizes = []
        n_batches = 0
        checked = 0
        failed = 0
        bad_total = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while nxt < len(reqs) or pending or inflight:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].t_arrival <= now:
                pending.append(reqs[nxt])
                nxt += 1
            # completions
            still = []
            for b in inflight:
                if b.event.query():
                    t = time.perf_counter() - t0
                    status = b.status.numpy()
                    for r, stw in zip(b.reqs, status):
                        r.t_done = t
                        r.status = int(stw)
                    done.extend(b.reqs)
                    failed += int(np.count_nonzero(status))
                    if b.check is not None:
                        flags = b.check.numpy()
                        for k, bad in zip(b.check_idx, flags):
                            if status[k] == 0:      # a failed request's output is undefined
                                checked += 1
                                bad_total += int(bad)
                    self.arena.free(*b.block)
                    free_streams.append(b.hip_stream)
                    for k, sid in (b.fresh or ()):
                        if status[k] == 0:      # its records are complete and proven now
                            self.pool.have_rec[sid] = True
                else:
                    still.append(b)
            inflight = still
            # launch everything admitted so far as one batch on a free stream
            if pending and free_streams:
                take = []
                need = 0
                room = self.arena.largest_free()
                while pending and len(take) < self.max_batch:
                    r = pending[0]
                    n = int(self.pool.n_sym[r.stream])
                    # outputs plus a workspace bound (about 1.5 B per unit at K = 32,
                    # about 2.5 B per unit for a decode from 32-unit item records) and
                    # the request's status word
                    est = _up(n) + (3 if self.use_records else 2) * \
                        int(self.pool.n_units[r.stream]) + 4096 + 4
                    if self.host:
                        est += int(self.pool.in_bytes[r.stream])
                    if take and need + est > room:
                        break
                    take.append(pending.popleft())
                    need += est
                if take:
                    b = self._launch(take, now, free_streams[-1], check_every)
                    if b is None and not inflight:
                        raise RuntimeError("request of %d symbols does not fit the HBM budget"
                                           % int(self.pool.n_sym[take[0].stream]))
                    if b is None:
                        for r in reversed(take):
                            pending.appendleft(r)
                    else:
                        free_streams.pop()
                        for r in take:
                            r.t_admit = now
                        inflight.append(b)
                        sizes.append(len(take))
                        n_batches += 1
                        continue
            if nxt < len(reqs) and not pending:
                wait = reqs[nxt].t_arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, poll_sleep) if inflight else wait)
            else:
                time.sleep(poll_sleep)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array([(r.t_done - r.t_arrival) * 1e3 for r in done])
        syms = int(sum(int(self.pool.n_sym[r.stream]) for r in done if r.status == 0))
        st = ServeStats(lat, syms, wall, n_batches, sizes, self.arena.peak, checked,
                        bad_total if check_every else -1, failed)
        return st


def poisson_arrivals(rate_per_s, duration_s, seed=7):
    """Open-loop arrival times of a Poisson process on [0, duration_s)."""
    rng = np.random.default_rng(seed)
    n = int(rate_per_s * duration_s * 1.2) + 16
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, n))
    return t[t < duration_s]


def request_mix(n, seed=20261204, min_bytes=1 << 10, max_bytes=64 << 20, lam_lo=0.005,
                lam_hi=2.0, states=(1024, 4096)):
    """Served-configuration request streams: (seed, lambda, n_symbols, L) per stream,
    size log-uniform in [min_bytes, max_bytes], lambda log-uniform, L drawn 50/50."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(min_bytes), np.log(max_bytes), n))).astype(np.int64)
    lams = np.exp(rng.uniform(np.log(lam_lo), np.log(lam_hi), n))
    Ls = rng.choice(np.array(states), n)
    return [(seed * 7919 + i, float(lams[i]), int(sizes[i]), int(Ls[i])) for i in range(n)]
