"""The HIP engine from Python, over the pharos_synth C ABI.

Device memory, streams and events come from the library itself (pharos_synth_malloc and friends), so one HIP
runtime owns the device; every sample is computed by libpharos_amd.so's gfx950 kernels.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check


def device_count():
    return int(_lib.lib().pharos_synth_device_count())


def require_gpu():
    if device_count() < 1:
        raise RuntimeError("no ROCm GPU visible: the pharos engine runs on the GPU only")


class DeviceBuffer:
    """A block of device memory owned by the library's HIP runtime."""

    def __init__(self, nbytes, device=0):
        self.nbytes = int(nbytes)
        self.device = device
        p = C.c_void_p()
        check(_lib.lib().pharos_synth_malloc(device, max(self.nbytes, 16), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, array, device=0):
        a = np.ascontiguousarray(array)
        buf = cls(a.nbytes, device)
        if a.nbytes:
            check(_lib.lib().pharos_synth_memcpy(buf.ptr, a.ctypes.data, a.nbytes, 0))
        return buf

    def at(self, offset):
        """Device address `offset` bytes in."""
        return self.ptr + int(offset)

    def to_numpy(self, nbytes=None, offset=0):
        n = self.nbytes - offset if nbytes is None else int(nbytes)
        out = np.empty(n, dtype=np.uint8)
        if n:
            check(_lib.lib().pharos_synth_memcpy(out.ctypes.data, self.at(offset), n, 1))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            _lib.lib().pharos_synth_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def digest(ptr, nbytes, stream=None):
    """pharos_synth_digest of nbytes of device memory at ptr."""
    d = C.c_uint64()
    check(_lib.lib().pharos_synth_digest(ptr, int(nbytes), C.byref(d), stream.handle if stream else None))
    return int(d.value)


def host_digest(data):
    """The same checksum of host bytes (numpy), for comparing device outputs with host ones."""
    b = np.frombuffer(bytes(data) if not isinstance(data, np.ndarray) else data.tobytes(), dtype=np.uint8)
    n = len(b)
    pad = np.zeros((n + 7) // 8 * 8, dtype=np.uint8)
    pad[:n] = b
    w = pad.view("<u8")
    m = np.uint64(0xFFFFFFFFFFFFFFFF)

    def fmix(x):
        x = x ^ (x >> np.uint64(33))
        x = x * np.uint64(0xff51afd7ed558ccd)
        x = x ^ (x >> np.uint64(33))
        x = x * np.uint64(0xc4ceb9fe1a85ec53)
        return x ^ (x >> np.uint64(33))

    with np.errstate(over="ignore"):
        i = np.arange(len(w), dtype=np.uint64)
        s = int(fmix(w ^ (i * np.uint64(0x9E3779B97F4A7C15))).sum(dtype=np.uint64)) & int(m)
        return int(fmix(np.array([s ^ n], dtype=np.uint64))[0])


class Stream:
    def __init__(self, device=0):
        s = C.c_void_p()
        check(_lib.lib().pharos_synth_stream_create(device, C.byref(s)))
        self.handle = s.value

    def sync(self):
        check(_lib.lib().pharos_synth_stream_sync(self.handle))

    def close(self):
        """Destroys the stream now (a process's streams share a few hardware queues: fewer live streams, fewer
        shared)."""
        if getattr(self, "handle", None):
            _lib.lib().pharos_synth_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(_lib.lib().pharos_synth_event_create(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None):
        check(_lib.lib().pharos_synth_event_record(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(_lib.lib().pharos_synth_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            _lib.lib().pharos_synth_event_destroy(self.handle)
        except Exception:
            pass


class Engine:
    """pharos_synth_engine for one scenario configuration (its tables, rate, format, noise and seed)."""

    def __init__(self, scenario, device=0, format_override=None):
        require_gpu()
        L = _lib.lib()
        self.scenario = scenario
        self.device = device
        self.format = scenario.format_id if format_override is None else format_override
        tabs = scenario.tables()
        cfg = _lib.SynthConfig(scenario.samples_per_epoch, self.format, 0 if scenario.noise else 1, 0,
                               scenario.seed & (2**64 - 1))
        h = C.c_void_p()
        check(L.pharos_synth_create(C.byref(cfg), tabs.array, tabs.count, device, C.byref(h)))
        self.handle = h
        self.bytes_per_epoch = int(L.pharos_synth_bytes_per_epoch(h))

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().pharos_synth_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return int(_lib.lib().pharos_synth_device_bytes(self.handle))

    def window(self, first_epoch, epochs, rows_ptr, masks_ptr, mask_len, gain, out_ptr, clips=True, stream=None):
        """Plain launch of `epochs` epochs (device pointers); returns per-epoch clip counts or None."""
        L = _lib.lib()
        nch = len(self.scenario.channels)
        clipped = (C.c_uint64 * epochs)() if clips else None
        check(L.pharos_synth_window(self.handle, first_epoch, epochs, nch, rows_ptr if nch else None, masks_ptr,
                                    mask_len, float(gain), out_ptr, clipped, stream.handle if stream else None))
        return np.frombuffer(clipped, dtype=np.uint64).copy() if clips else None


class HostAgc:
    """The adaptive gain carried between plain windows on the host, by the library's AGC step
    (pharos_synth_agc_account: Agc.account, gain.zig:19-35, applied per window as engine/run.zig:238-247 does)."""

    def __init__(self, cap=4096):
        self.state = _lib.SynthAgc(1.0, 0, 0, 0, 0)
        self.log = (_lib.SynthAdjustment * cap)()

    @property
    def gain(self):
        return float(self.state.gain)

    def account(self, clipped, components, epoch_after):
        check(_lib.lib().pharos_synth_agc_account(C.byref(self.state), int(clipped), int(components),
                                                  int(epoch_after), self.log, len(self.log)))

    def adjustments(self):
        n = min(self.state.adjustments, len(self.log))
        return [(int(a.epoch), float(a.gain), float(a.clip_rate)) for a in self.log[:n]]


def run_plain(scenario, first_epoch=0, epochs=None, device=0, format_override=None, window=100, gain=None,
              derive=False):
    """Per-window plain launches with the host AGC between windows (gain.zig:19-35, engine/run.zig:238-247); rows
    from the host stage, or with derive=True from the scenario's segments on the device.

    Returns (bytes, per-epoch clips, the AGC's gain after the run, its adjustments as (epoch, gain, rate))."""
    epochs = scenario.epochs if epochs is None else epochs
    eng = Engine(scenario, device, format_override)
    if derive:
        d_rows, d_masks, mask_len = derive_rows(scenario, first_epoch, epochs, device=device)
        masks = np.zeros(mask_len, dtype=np.uint8)
    else:
        rows, masks = scenario.window(first_epoch, epochs)
        d_rows, d_masks = DeviceBuffer.from_array(rows, device), DeviceBuffer.from_array(masks, device)
    out = DeviceBuffer(epochs * eng.bytes_per_epoch, device)
    nch = len(scenario.channels)
    constant = scenario.gain_db is not None or gain is not None
    g = gain if gain is not None else (float(np.float32(10.0 ** (scenario.gain_db / 20.0)))
                                       if scenario.gain_db is not None else 1.0)
    agc = HostAgc()
    clips = np.zeros(epochs, dtype=np.uint64)
    e = 0
    while e < epochs:
        absolute = first_epoch + e
        n = min(window - absolute % window, epochs - e)
        gw = g if constant else float(np.float32(agc.gain))
        c = eng.window(absolute, n, d_rows.at(e * nch * 64), d_masks.ptr, len(masks), gw, out.at(e * eng.bytes_per_epoch))
        clips[e:e + n] = c
        if not constant:
            agc.account(int(c.sum()), n * scenario.samples_per_epoch * 2, absolute + n)
        e += n
    result = out.to_numpy()
    eng.close()
    return result, clips, agc.gain, agc.adjustments()


def derive_rows(scenario, first_epoch, epochs, block_epochs=100, device=0, stream=None, segments=None, levels=None):
    """Rows on the device from the scenario's per-block segments (pharos_synth_derive): the host computes one
    segment per channel and block plus the masks; the device evaluates every epoch's row.  levels (LEVEL_DTYPE,
    sorted by channel then epoch) are power-level steps the device applies to the rows' amplitudes at their epochs
    (pharos_synth_derive_levels).  Returns (rows, masks, mask_len) as device buffers."""
    if segments is None:
        segments = scenario.segments(first_epoch, epochs, block_epochs)
    segs, masks = segments
    nch = len(scenario.channels)
    d_segs = DeviceBuffer.from_array(segs, device)
    d_masks = DeviceBuffer.from_array(masks, device)
    d_rows = DeviceBuffer(max(1, epochs * nch) * 64, device)
    nlev = 0 if levels is None else len(levels)
    d_levels = DeviceBuffer.from_array(np.ascontiguousarray(levels, dtype=_lib.LEVEL_DTYPE), device) if nlev else None
    check(_lib.lib().pharos_synth_derive_levels(d_segs.ptr if nch else None, block_epochs, first_epoch, epochs, nch,
                                                scenario.start_gps_ms, scenario.start_gps_sub_ms,
                                                scenario.samples_per_epoch, d_levels.ptr if nlev else None, nlev,
                                                d_rows.ptr, stream.handle if stream else None))
    if stream:
        stream.sync()
    else:
        check(_lib.lib().pharos_synth_device_sync(device))
    return d_rows, d_masks, len(masks)


class Job:
    """One run in a batch: its engine, device rows/masks/output and its gain state.  Rows come from host rows and
    masks (rows=, masks=, or the scenario's host stage), or with derive=True from the scenario's segments, evaluated
    on the device with the power-level steps `levels`, if any."""

    def __init__(self, scenario, first_epoch=0, epochs=None, device=0, rows=None, masks=None, adjustments_cap=4096,
                 derive=False, block_epochs=100, levels=None):
        self.scenario = scenario
        self.first_epoch = first_epoch
        self.epochs = scenario.epochs if epochs is None else epochs
        self.engine = Engine(scenario, device)
        if derive:
            self.rows, self.masks, self.mask_len = derive_rows(scenario, first_epoch, self.epochs, block_epochs, device,
                                                               levels=levels)
        else:
            if rows is None:
                rows, masks = scenario.window(first_epoch, self.epochs)
            self.rows = DeviceBuffer.from_array(rows, device)
            self.masks = DeviceBuffer.from_array(masks, device)
            self.mask_len = len(masks)
        self.out = DeviceBuffer(self.epochs * self.engine.bytes_per_epoch, device)
        self.agc = _lib.SynthAgc(1.0, 0, 0, 0, 0)
        self.adjust = (_lib.SynthAdjustment * adjustments_cap)()
        self.clipped = C.c_uint64(0)

    @property
    def constant_gain(self):
        g = self.scenario.gain_db
        return None if g is None else float(np.float32(10.0 ** (g / 20.0)))

    def record(self):
        cg = self.constant_gain
        return _lib.SynthJob(self.engine.handle.value, self.first_epoch, self.epochs, len(self.scenario.channels),
                             self.rows.ptr, self.masks.ptr, self.mask_len, self.out.ptr,
                             1 if cg is not None else 0, cg if cg is not None else 0.0, C.pointer(self.agc),
                             C.cast(self.adjust, C.POINTER(_lib.SynthAdjustment)), len(self.adjust), 0,
                             C.pointer(self.clipped))

    def adjustments(self):
        n = min(self.agc.adjustments, len(self.adjust))
        return [(int(a.epoch), float(a.gain), float(a.clip_rate)) for a in self.adjust[:n]]

    def result(self):
        return self.out.to_numpy()

    def to_file(self, path):
        """The run written to `path` by the library's file sink (pharos_synth_to_file: blocks synthesised on the
        device while the previous block is copied and written); returns the bytes on file.  The AGC state,
        adjustments and clip total come back as after a batched run."""
        n = C.c_uint64(0)
        rec = self.record()
        check(_lib.lib().pharos_synth_to_file(C.byref(rec), os.fsencode(str(path)), C.byref(n)))
        return n.value


def plain_job(job, stream=None, window=100):
    """The plain form over a Job's device buffers: per-window launches with the host AGC between windows (the
    reference's window loop, engine/run.zig:225-256 with gain.zig:19-35).  Returns (clip total, final gain,
    adjustments as (epoch, gain, rate))."""
    sc, eng = job.scenario, job.engine
    nch = len(sc.channels)
    cg = job.constant_gain
    agc = HostAgc()
    total = 0
    e = 0
    while e < job.epochs:
        absolute = job.first_epoch + e
        n = min(window - absolute % window, job.epochs - e)
        gw = cg if cg is not None else float(np.float32(agc.gain))
        c = eng.window(absolute, n, job.rows.at(e * nch * 64), job.masks.ptr, job.mask_len, gw,
                       job.out.at(e * eng.bytes_per_epoch), clips=True, stream=stream)
        total += int(c.sum())
        if cg is None:
            agc.account(int(c.sum()), n * sc.samples_per_epoch * 2, absolute + n)
        e += n
    return total, agc.gain, agc.adjustments()


def run_batch(jobs, workgroups_per_cu=0, stream=None, channels_per_pass=0, wait_limit=0, flags=0):
    """Runs jobs (a list of Job) in one persistent launch and waits for it."""
    arr = (_lib.SynthJob * len(jobs))(*[j.record() for j in jobs])
    tuning = _lib.SynthTuning(workgroups_per_cu, channels_per_pass, wait_limit, flags)
    check(_lib.lib().pharos_synth_batch(arr, len(jobs), C.byref(tuning), stream.handle if stream else None))
    return jobs


class Batch:
    """A submitted batch (pharos_synth_batch_submit): query, abort and wait."""

    def __init__(self, jobs, workgroups_per_cu=0, channels_per_pass=0, wait_limit=0, flags=0, stream=None):
        self.jobs = jobs
        self._arr = (_lib.SynthJob * len(jobs))(*[j.record() for j in jobs])
        tuning = _lib.SynthTuning(workgroups_per_cu, channels_per_pass, wait_limit, flags)
        h = C.c_void_p()
        check(_lib.lib().pharos_synth_batch_submit(self._arr, len(jobs), C.byref(tuning),
                                                   stream.handle if stream else None, C.byref(h)))
        self.handle = h.value

    def done(self):
        st = _lib.lib().pharos_synth_batch_query(self.handle)
        if st < 0:
            raise _lib.PharosError(-st, _lib.last_error())
        return st == 1

    def abort(self):
        check(_lib.lib().pharos_synth_batch_abort(self.handle))

    def wait(self):
        h, self.handle = self.handle, None
        check(_lib.lib().pharos_synth_batch_wait(h))
        return self.jobs


class Graph:
    """A job's window loop captured as a HIP graph (pharos_synth_graph_*): AGC steps run on the device between
    windows, and a launch replays the whole run."""

    def __init__(self, job):
        self.job = job
        self._record = job.record()  # kept alive: the graph holds its host pointers
        h = C.c_void_p()
        check(_lib.lib().pharos_synth_graph_create(C.byref(self._record), C.byref(h)))
        self.handle = h.value

    @property
    def windows(self):
        return int(_lib.lib().pharos_synth_graph_windows(self.handle))

    def launch(self, stream=None):
        check(_lib.lib().pharos_synth_graph_launch(self.handle, stream.handle if stream else None))

    def finish(self, stream=None):
        check(_lib.lib().pharos_synth_graph_finish(self.handle, stream.handle if stream else None))

    def run(self, stream=None):
        self.launch(stream)
        self.finish(stream)
        return self.job

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().pharos_synth_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
