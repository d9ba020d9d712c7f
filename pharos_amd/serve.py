"""Continuous-batching request loop over the batched engine (one GPU).

Requests are runs (scenarios) arriving over time. The loop works in iterations:
1. It admits every arrived request whose buffers fit the device memory left (free HBM less a reserve when the loop
   starts, carved by an arena), oldest first.
2. It submits one persistent batch holding the next slice of every admitted request. A slice is a whole number
   of 100-epoch AGC windows, sized so it carries about SLICE_WORK channel-samples.
3. While the GPU runs that batch, worker threads make each request's following slice (its segments and masks) and
   copy it into the request's second buffer set, through pinned staging on copy streams of their own, so the
   copies run beside the batch instead of waiting for it.
4. It retires the requests that completed.

A request's AGC state (pharos_synth_agc) carries from slice to slice, so its bytes are those of the same run alone,
whatever else shares its batches. Latency is measured from a request's arrival to the end of the batch that
writes its last epoch. Its output is in device memory when `on_complete` is called with it, and is freed after
that unless the loop is asked to keep outputs.
"""
import bisect
import collections
import ctypes as C
import os
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib
from ._lib import check
from .engine import Engine, DeviceBuffer, Stream, require_gpu

BLOCK = 100             # epochs per segment (one AGC window)
HOST_CHUNK = 256 << 20  # bytes per device-to-host copy of a finished output (serve(to_host=True))
SLICE_MIN = 1000        # epochs of a request per iteration, at least (ten AGC windows) ...
SLICE_WORK = 16e9       # ... or about this many channel-samples (a few ms of the GPU), in whole AGC windows
RESERVE = 8 << 30       # HBM left free for the library's own allocations: engines and the batch-state blocks it
                        # caches between batches (at most 4 GiB of them); each request's share of a batch's state
                        # is held against the arena while it is admitted (Request.state_bytes)
PREP_BYTES = 272        # the batched form's row pre-pass per row (RowPrep) of a job with several tiles per epoch
ITEM_BYTES = 4          # the batched form's segment of each work item (a 4096-sample tile of an epoch)
MIN_COPY = 64 << 10     # bytes a copy carries at least: the runtime moves a copy this large with a DMA engine, beside
                        # a running batch; a small one (4 KiB measured) goes to a copy kernel, which waits for the
                        # batch to leave the compute units
STATE_SLACK = 2         # a batch's state block rounds up to a power of two below 256 MiB


class Request:
    def __init__(self, scenario, arrival_s=0.0, rid=0):
        self.scenario = scenario
        self.arrival = float(arrival_s)
        self.id = rid
        self.engine = None
        self.out = None
        self.done_epochs = 0
        self.agc = _lib.SynthAgc(1.0, 0, 0, 0, 0)
        self.adjust = (_lib.SynthAdjustment * 4096)()
        self.clipped_total = 0
        self.slice_clip = C.c_uint64(0)
        self.staged = None       # (first_epoch, epochs, DeviceBuffer rows, DeviceBuffer masks, mask_len)
        self.admitted = None
        self.completed = None

    @property
    def epochs(self):
        return self.scenario.epochs

    @property
    def out_bytes(self):
        return self.scenario.epochs * self.scenario.bytes_per_epoch

    @property
    def slice_epochs(self):
        per_epoch = max(1, len(self.scenario.channels)) * self.scenario.samples_per_epoch
        n = max(SLICE_MIN, int(SLICE_WORK // per_epoch) // 100 * 100)
        return min(n, self.epochs)

    def slice_bytes(self):
        """Device bytes of one slice's rows, masks and segments (the scenario stage's bounds)."""
        s, chans = self.scenario.spec()
        n = max(1, len(self.scenario.channels))
        blocks = -(-self.slice_epochs // BLOCK)
        masks = (int(_lib.lib().pharos_scenario_mask_bound(C.byref(s), chans, BLOCK)) + 16 * n) * blocks + 16
        return self.slice_epochs * n * 64, masks, blocks * n * _lib.SEGMENT_DTYPE.itemsize

    def state_bytes(self):
        """Device bytes of this request's share of a batch's state: the row pre-pass of one slice (when an epoch
        spans several tiles), its work items' segments and its job and segment records, with the rounding of the
        block they live in."""
        rows = self.slice_epochs * max(1, len(self.scenario.channels))
        prep = rows * PREP_BYTES if self.scenario.samples_per_epoch > 4096 else 0
        items = self.slice_epochs * -(-self.scenario.samples_per_epoch // 4096) * ITEM_BYTES
        segments = self.slice_epochs // 100 + 2
        return STATE_SLACK * (prep + items + 256 + segments * 64)

    def spec(self):
        """The scenario's ctypes records (built once: slices reuse them)."""
        if getattr(self, "_spec", None) is None:
            self._spec = self.scenario.spec()
        return self._spec

    def mask_cap(self, epochs):
        """Bytes the masks of a slice of `epochs` epochs can take (the host stage's bound)."""
        s, chans = self.spec()
        n = max(1, len(self.scenario.channels))
        blocks = -(-epochs // BLOCK)
        return (int(_lib.lib().pharos_scenario_mask_bound(C.byref(s), chans, BLOCK)) + 16 * n) * blocks + 16

    def io_split(self):
        """Offset of the masks in a slice's segment-and-mask buffer (the segments come first)."""
        return (self.slice_bytes()[2] + 255) // 256 * 256

    def buffer_sizes(self):
        """Bytes of the request's device buffers: its output, then two (rows, segments + masks) sets.  A set's
        segments and masks share one buffer, so one copy of at least MIN_COPY bytes brings both."""
        rows, masks, _ = self.slice_bytes()
        io = max(self.io_split() + masks, MIN_COPY)
        return [self.out_bytes, rows, io, rows, io]

    def gain_mode(self):
        return 1 if self.scenario.gain_db is not None else 0

    def constant_gain(self):
        g = self.scenario.gain_db
        return 0.0 if g is None else float(np.float32(10.0 ** (g / 20.0)))

    def adjustments(self):
        n = min(self.agc.adjustments, len(self.adjust))
        return [(int(a.epoch), float(a.gain), float(a.clip_rate)) for a in self.adjust[:n]]


def _prepare(req, first):
    """The host's part of the slice starting at `first` (runs in a worker thread; the C stage releases the GIL):
    one segment per channel and 100-epoch block and the symbol masks; the rows are derived on the device.  A
    request's first slice is SLICE_MIN epochs, so it starts soon; the later ones carry about SLICE_WORK."""
    n = min(SLICE_MIN if first == 0 else req.slice_epochs, req.epochs - first)
    segs, masks = req.scenario.segments(first, n, BLOCK)
    return first, n, segs, masks


def _arrive(req, device):
    """Work done as soon as a request arrives, before it is admitted: its engine (code tables on the device, a few
    MB) and its first slice."""
    req.engine = Engine(req.scenario, device)
    return _prepare(req, 0)


class View:
    """A device range inside a larger buffer: .ptr and .nbytes, as the loop's buffers have."""
    __slots__ = ("ptr", "nbytes")

    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes = int(ptr), int(nbytes)


class ArenaBuffer:
    """A block of an arena: .ptr, .nbytes, .at() and .to_numpy() as a DeviceBuffer has; free() returns it."""
    __slots__ = ("arena", "block", "offset", "nbytes", "ptr")

    def __init__(self, arena, block, offset, nbytes):
        self.arena, self.block, self.offset, self.nbytes = arena, block, offset, nbytes
        self.ptr = arena.blocks[block].at(offset)

    def at(self, offset):
        return self.ptr + int(offset)

    def to_numpy(self, nbytes=None, offset=0):
        n = self.nbytes - offset if nbytes is None else int(nbytes)
        out = np.empty(n, dtype=np.uint8)
        if n:
            check(_lib.lib().pharos_synth_memcpy(out.ctypes.data, self.at(offset), n, 1))
        return out

    def free(self):
        if self.ptr is not None:
            self.arena.release(self)
            self.ptr = None


class DeviceArena:
    """Device memory for the loop's buffers, carved from a few large blocks: blocks of `block_bytes` (or one
    buffer's size, if larger) are allocated from the driver as needed, up to `limit_bytes` in all, and are only
    returned at close().  Buffers are placed first fit in a block's free ranges, which merge with their neighbours
    when freed, so admissions and retirements make no driver calls in steady state (a large device allocation or
    free takes milliseconds)."""
    ALIGN = 4096

    def __init__(self, device, limit_bytes, block_bytes=16 << 30):
        self.device = device
        self.limit = int(limit_bytes)
        self.block_bytes = int(block_bytes)
        self.blocks = []        # DeviceBuffer per block
        self.ranges = []        # per block: sorted [offset, size] free ranges
        self.used = 0
        self.held = 0           # bytes held against the limit for memory the library allocates outside the arena

    def hold(self, nbytes):
        """Holds nbytes of the limit for an allocation outside the arena; False if the limit has no room."""
        if self.used + self.held + int(nbytes) > self.limit:
            return False
        self.held += int(nbytes)
        return True

    def unhold(self, nbytes):
        self.held -= int(nbytes)

    def _round(self, n):
        return (max(int(n), 1) + self.ALIGN - 1) // self.ALIGN * self.ALIGN

    def reserved(self):
        return sum(b.nbytes for b in self.blocks)

    def alloc(self, nbytes):
        """An ArenaBuffer of at least nbytes, or None if it does not fit within the limit."""
        n = self._round(nbytes)
        if self.used + self.held + n > self.limit:
            return None
        for bi, free in enumerate(self.ranges):
            for k, (off, size) in enumerate(free):
                if size >= n:
                    if size == n:
                        del free[k]
                    else:
                        free[k] = [off + n, size - n]
                    self.used += n
                    return ArenaBuffer(self, bi, off, n)
        size = max(self.block_bytes, n)
        if self.reserved() + self.held + size > self.limit:
            size = n
            if self.reserved() + self.held + size > self.limit:
                return None
        self.blocks.append(DeviceBuffer(size, self.device))
        self.ranges.append([[n, size - n]] if size > n else [])
        self.used += n
        return ArenaBuffer(self, len(self.blocks) - 1, 0, n)

    def alloc_all(self, sizes):
        """Buffers for every size, or None (and nothing held) if they do not all fit."""
        got = []
        for n in sizes:
            b = self.alloc(n)
            if b is None:
                for g in got:
                    g.free()
                return None
            got.append(b)
        return got

    def release(self, buf):
        free = self.ranges[buf.block]
        off, n = buf.offset, buf.nbytes
        k = bisect.bisect_left(free, [off, 0])
        if k < len(free) and free[k][0] == off + n:       # merge with the following range
            n += free[k][1]
            del free[k]
        if k > 0 and free[k - 1][0] + free[k - 1][1] == off:  # ... and the preceding one
            free[k - 1][1] += n
        else:
            free.insert(k, [off, n])
        self.used -= buf.nbytes

    def close(self):
        for b in self.blocks:
            b.free()
        self.blocks, self.ranges, self.used = [], [], 0


class Server:
    def __init__(self, device=0, threads=None, reserve=RESERVE, workgroups_per_cu=0, channels_per_pass=0):
        require_gpu()
        self.device = device
        self.reserve = reserve
        self.tuning = _lib.SynthTuning(workgroups_per_cu, channels_per_pass, 0, 0)
        threads = threads or min(16, os.cpu_count() or 4)
        # Two pools: the next slices of admitted requests, which the next batch waits for, never queue behind the
        # work of requests that just arrived (their engines and first slices).
        self.pool = ThreadPoolExecutor(max_workers=threads)
        self.arrivals = ThreadPoolExecutor(max_workers=max(1, threads // 4))
        self.stream = Stream(device)
        # Copy streams, made right after the loop's stream so that each gets a hardware queue of its own (streams
        # share a process's few hardware queues in turn, and a copy queued behind the running batch on a shared
        # queue would wait for it).  Every worker's slice copies go on one, finished outputs on the other.
        self.copy_stream = Stream(device)
        self.host_stream = Stream(device)
        # Copies of finished outputs to host memory (serve(to_host=True)): two copier threads, each with a pinned
        # staging chunk; requests whose copy ended come back through _copied for their device memory to be freed.
        self.copier = ThreadPoolExecutor(max_workers=2)
        # Staging records for copies (a pinned buffer and a copy stream each), one per thread that copies.
        self._pinned = []
        self._pinned_lock = threading.Lock()
        self._stagings = queue.SimpleQueue()       # slices: the workers and the loop
        self._host_stagings = queue.SimpleQueue()  # finished outputs: the copier threads
        for _ in range(threads + 1):
            self._stagings.put({"ptr": 0, "nbytes": 0, "stream": self.copy_stream})
        for _ in range(2):
            self._host_stagings.put({"ptr": 0, "nbytes": 0, "stream": self.host_stream})
        self._copied = collections.deque()
        free, total = C.c_uint64(), C.c_uint64()
        check(_lib.lib().pharos_synth_mem_info(device, C.byref(free), C.byref(total)))
        self.arena = DeviceArena(device, max(0, int(free.value) - reserve))
        self.iterations = 0
        self.batch_sizes = []
        self.phase_s = collections.Counter()  # host seconds per loop phase

    def _borrow(self, nbytes, pool=None):
        """A staging record (pinned host buffer of at least nbytes, and the copy stream) for one task; _give_back
        returns it.  Copies go through them as asynchronous copies on a copy stream: a synchronous copy waits for
        the batch running on the device, an asynchronous one from pinned memory runs beside it.  The streams are
        made when the loop starts, since creating one waits for the device to go idle."""
        st = (pool or self._stagings).get()
        st["pool"] = pool or self._stagings
        if st["nbytes"] < nbytes:
            p = C.c_void_p()
            size = max(int(nbytes), st["nbytes"] * 2)
            check(_lib.lib().pharos_synth_host_alloc(size, C.byref(p)))
            with self._pinned_lock:
                self._pinned.append(p.value)
            st["ptr"], st["nbytes"] = p.value, size
        return st

    def _give_back(self, st):
        st["pool"].put(st)

    def _copy_into(self, req, buffers, prepared):
        """Copies a prepared slice's segments and masks into one of the request's buffer sets (worker thread) in one
        copy of at least MIN_COPY bytes; its rows are derived into the set's row buffer on the loop's stream just
        before the batch that reads them."""
        first, n, segs, masks = prepared
        d_rows, d_io = buffers
        split = req.io_split()
        if segs.nbytes > split or split + masks.nbytes > d_io.nbytes:
            raise RuntimeError("slice larger than its buffers")
        count = min(max(split + masks.nbytes, MIN_COPY), d_io.nbytes)
        st = self._borrow(count)
        try:
            staging = np.ctypeslib.as_array((C.c_uint8 * count).from_address(st["ptr"]))
            staging[:segs.nbytes] = segs.view(np.uint8).reshape(-1)
            staging[split:split + masks.nbytes] = masks
            check(_lib.lib().pharos_synth_memcpy_async(d_io.ptr, st["ptr"], count, 0, st["stream"].handle))
            st["stream"].sync()
        finally:
            self._give_back(st)
        return first, n, d_rows, View(d_io.ptr + split, masks.nbytes), len(masks), View(d_io.ptr, segs.nbytes)

    def _next_slice(self, req, first, buffers):
        """The slice at `first` (one worker task): the host stage writes its segments and masks straight into a
        pinned staging buffer, and one copy of at least MIN_COPY bytes takes them to the buffer set.  (Writing into
        fresh arrays instead made the workers contend for the page faults of the new memory.)"""
        n = min(SLICE_MIN if first == 0 else req.slice_epochs, req.epochs - first)
        d_rows, d_io = buffers
        split = req.io_split()
        nch = len(req.scenario.channels)
        blocks = -(-n // BLOCK)
        cap = req.mask_cap(n)
        seg_bytes = blocks * nch * _lib.SEGMENT_DTYPE.itemsize
        if seg_bytes > split or split + cap > d_io.nbytes:
            raise RuntimeError("slice larger than its buffers")
        t0 = time.perf_counter()
        st = self._borrow(max(split + cap, MIN_COPY))
        try:
            s, chans = req.spec()
            used = _lib.lib().pharos_scenario_segments(C.byref(s), chans, first, n, BLOCK, st["ptr"],
                                                       st["ptr"] + split, cap)
            if used < 0:
                raise _lib.PharosError(int(-used), "scenario segments failed")
            if used == 0:
                C.memset(st["ptr"] + split, 0, 1)
            count = min(max(split + int(used), MIN_COPY), d_io.nbytes)
            t1 = time.perf_counter()
            check(_lib.lib().pharos_synth_memcpy_async(d_io.ptr, st["ptr"], count, 0, st["stream"].handle))
            st["stream"].sync()
            t2 = time.perf_counter()
        finally:
            self._give_back(st)
        with self._pinned_lock:  # per-thread seconds, summed over the workers
            self.phase_s["slice_host_thread"] += t1 - t0
            self.phase_s["slice_copy_thread"] += t2 - t1
        mask_len = max(int(used), 1)
        return first, n, d_rows, View(d_io.ptr + split, mask_len), mask_len, View(d_io.ptr, seg_bytes)

    def _admit(self, req):
        """Places the request's buffers in the arena and holds its share of the batch state; False (and nothing
        held) if they do not fit now."""
        if not self.arena.hold(req.state_bytes()):
            return False
        got = self.arena.alloc_all(req.buffer_sizes())
        if got is None:
            self.arena.unhold(req.state_bytes())
            return False
        req.out = got[0]
        # Two row/mask/segment sets per request: the next slice is written into one while the batch reads the other.
        req.buffers = [(got[1], got[2]), (got[3], got[4])]
        req.flip = 0
        req.admitted = time.perf_counter()
        return True

    def _spare(self, req):
        """The row/mask buffers the next slice goes into (not the ones the running batch reads)."""
        b = req.buffers[req.flip]
        req.flip ^= 1
        return b

    def _retire(self, req, keep_output):
        self.arena.unhold(req.state_bytes())
        for pair in req.buffers:
            for b in pair:
                b.free()
        req.buffers = None
        req.staged = None
        if not keep_output:
            req.out.free()
            req.out = None
            req.engine.close()

    def _to_host(self, req):
        """Copies a finished request's output to host memory in chunks through a pinned staging buffer on a copy
        stream of its own (copier thread), so the copies run beside the next batches; stamps when its last byte
        arrived, and hands the request back for its device output to be freed.  A last chunk below MIN_COPY is
        copied as the output's last MIN_COPY bytes."""
        st = self._borrow(min(HOST_CHUNK, max(req.out_bytes, 1)), self._host_stagings)
        try:
            L = _lib.lib()
            for off in range(0, req.out_bytes, HOST_CHUNK):
                n = min(HOST_CHUNK, req.out_bytes - off)
                if n < MIN_COPY <= req.out_bytes:
                    off, n = req.out_bytes - MIN_COPY, MIN_COPY
                check(L.pharos_synth_memcpy_async(st["ptr"], req.out.at(off), n, 1, st["stream"].handle))
                st["stream"].sync()
        finally:
            self._give_back(st)
        req.host_done = time.perf_counter()
        self._copied.append(req)

    def _release_copied(self):
        while self._copied:
            req = self._copied.popleft()
            req.out.free()
            req.out = None
            req.engine.close()

    def close(self):
        """Returns the arena's device memory (outputs kept by serve(keep_outputs=True) included) and the pinned
        staging chunks."""
        self.copier.shutdown(wait=True)
        self.pool.shutdown(wait=True)
        self.arrivals.shutdown(wait=True)
        self._release_copied()
        self.arena.close()
        for p in self._pinned:
            _lib.lib().pharos_synth_host_free(p)
        self._pinned = []
        for st in (self.stream, self.copy_stream, self.host_stream):
            st.close()

    def _derive(self, req):
        """Queues the device derive of the request's staged slice on the loop's stream (ahead of its batch)."""
        first, n, d_rows, d_masks, mask_len, d_segs = req.staged
        sc = req.scenario
        check(_lib.lib().pharos_synth_derive(d_segs.ptr, BLOCK, first, n, len(sc.channels), sc.start_gps_ms,
                                             sc.start_gps_sub_ms, sc.samples_per_epoch, d_rows.ptr,
                                             self.stream.handle))

    def _job(self, req):
        first, n, d_rows, d_masks, mask_len, _ = req.staged
        req.current = (first, n)
        prior = min(req.agc.adjustments, len(req.adjust))
        adj = C.cast(C.c_void_p(C.addressof(req.adjust) + prior * C.sizeof(_lib.SynthAdjustment)),
                     C.POINTER(_lib.SynthAdjustment))
        cap = len(req.adjust) - prior
        req.slice_clip.value = 0
        return _lib.SynthJob(req.engine.handle.value, first, n, len(req.scenario.channels), d_rows.ptr, d_masks.ptr,
                             mask_len, req.out.at(first * req.scenario.bytes_per_epoch), req.gain_mode(),
                             req.constant_gain(), C.pointer(req.agc), adj, cap, 0, C.pointer(req.slice_clip))

    def serve(self, requests, on_complete=None, keep_outputs=False, to_host=False):
        """Runs `requests` (arrival times relative to the call) to completion; returns them with .completed set (the
        end of the batch that wrote their last epoch to device memory).  With to_host, each finished output is also
        copied to host memory by copier threads while the loop goes on, and .host_done is when its last byte
        arrived there."""
        L = _lib.lib()
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting, active, finished = collections.deque(), [], []
        copies = []
        t0 = time.perf_counter()
        for r in requests:
            r.arrival_abs = t0 + r.arrival
        while pending or waiting or active:
            self._release_copied()
            now = time.perf_counter()
            while pending and pending[0].arrival_abs <= now:
                req = pending.popleft()
                req.arrived = self.arrivals.submit(_arrive, req, self.device)
                waiting.append(req)
            # Admission by device memory, oldest first, of requests whose first slice is ready.
            admitted = []
            ta = time.perf_counter()
            while waiting and (waiting[0].arrived.done() or not active):
                if not self._admit(waiting[0]):
                    if not active and not admitted:
                        raise RuntimeError(f"request {waiting[0].id} needs more device memory than the loop has")
                    break
                admitted.append(waiting.popleft())
            tb = time.perf_counter()
            for req in admitted:
                prepared = req.arrived.result()
                req.arrived = None
                req.staged = self._copy_into(req, self._spare(req), prepared)
            active.extend(admitted)
            tc = time.perf_counter()
            self.phase_s["admit"] += tb - ta
            self.phase_s["first_slice"] += tc - tb
            if not active:
                if pending and not waiting:
                    time.sleep(max(0.0, min(0.002, pending[0].arrival_abs - time.perf_counter())))
                continue
            # One persistent batch over every active request's staged slice, after the derives of their rows.
            for r in active:
                self._derive(r)
            jobs = (_lib.SynthJob * len(active))(*[self._job(r) for r in active])
            h = C.c_void_p()
            check(L.pharos_synth_batch_submit(jobs, len(active), C.byref(self.tuning), self.stream.handle, C.byref(h)))
            td = time.perf_counter()
            self.phase_s["submit"] += td - tc
            self.iterations += 1
            self.batch_sizes.append(len(active))
            # While it runs: the next slice of every request that has one, derived on the host and copied into the
            # request's other row/mask buffer.
            nxt = [(r, sum(r.current)) for r in active if sum(r.current) < r.epochs]
            futures = [(r, self.pool.submit(self._next_slice, r, f, self._spare(r))) for r, f in nxt]
            for r, f in futures:
                r.staged = f.result()
            te = time.perf_counter()
            check(L.pharos_synth_batch_wait(h))
            end = time.perf_counter()
            self.phase_s["next_slices"] += te - td
            self.phase_s["wait"] += end - te
            still = []
            for r in active:
                r.clipped_total += int(r.slice_clip.value)
                r.done_epochs = sum(r.current)
                if r.done_epochs >= r.epochs:
                    r.completed = end
                    finished.append(r)
                    if on_complete:
                        on_complete(r)
                    self._retire(r, keep_outputs or to_host)
                    if to_host:
                        copies.append(self.copier.submit(self._to_host, r))
                else:
                    still.append(r)
            active = still
        for f in copies:
            f.result()
        self._release_copied()
        return finished


def make_requests(rate_hz, duration_s, seed=1, max_out_bytes=2 << 30, epochs_cap=None):
    """Poisson arrivals at `rate_hz` over `duration_s`, each a run drawn as the uneven batch draws them (4-48
    channels, 4.092-61.44 Msps, 1-30 s, IQ2-IQ16), its length cut so its output stays within max_out_bytes."""
    from . import scenario as S
    rng = np.random.default_rng([seed, 4])
    t, times = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= duration_s:
            break
        times.append(t)
    scs = S.c3(seed=seed + 1000, count=len(times), epochs_cap=epochs_cap)
    reqs = []
    for i, (at, sc) in enumerate(zip(times, scs)):
        fit = max(100, (max_out_bytes // sc.bytes_per_epoch) // 100 * 100)
        sc.epochs = int(min(sc.epochs, fit))
        reqs.append(Request(sc, at, i))
    return reqs


def measure(duration_s=10.0, rate_hz=20.0, seed=1, epochs_cap=None, to_host=False):
    """p50/p99 latency and aggregate channel-samples/s of a Poisson request stream: latency from arrival to the
    last byte in device memory and, with to_host, to the last byte in host memory."""
    reqs = make_requests(rate_hz, duration_s, seed, epochs_cap=epochs_cap)
    server = Server()
    t0 = time.perf_counter()
    done = server.serve(reqs, to_host=to_host)
    makespan = time.perf_counter() - t0
    server.close()
    lat = np.array([r.completed - r.arrival_abs for r in done]) * 1e3
    cs = sum(r.scenario.channel_samples for r in done)
    host = {}
    if to_host:
        hl = np.array([r.host_done - r.arrival_abs for r in done]) * 1e3
        host = {"host_p50_ms": float(np.percentile(hl, 50)), "host_p99_ms": float(np.percentile(hl, 99)),
                "host_bytes": int(sum(r.out_bytes for r in done))}
    return {"requests": len(done), "rate_hz": rate_hz, "arrival_window_s": duration_s, "makespan_s": makespan,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)), **host,
            "latency_to": "last byte in device memory (p50/p99); last byte in host memory (host_p50/p99)",
            "rows": "derived on the device from per-block segments (pharos_synth_derive)",
            "mean_ms": float(lat.mean()), "channel_samples": int(cs), "channel_samples_per_s": cs / makespan,
            "iterations": server.iterations, "mean_batch_jobs": float(np.mean(server.batch_sizes)),
            "host_phase_s": {k: round(v, 3) for k, v in server.phase_s.items()},
            "slice": {"min_epochs": SLICE_MIN, "channel_samples": SLICE_WORK}, "epochs_cap": epochs_cap}
