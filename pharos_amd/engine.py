"""The HIP engine from Python: device buffers (torch on ROCm) around the pharos_synth C ABI.

torch provides device memory and streams only; every sample is computed by libpharos_amd.so's gfx950 kernels.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, ROW_DTYPE


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU visible: the pharos engine runs on the GPU only")
    return torch


class Engine:
    """pharos_synth_engine for one scenario configuration (its tables, rate, format, noise and seed)."""

    def __init__(self, scenario, device=0, format_override=None):
        torch = _torch()
        L = _lib.lib()
        self.scenario = scenario
        self.device = device
        self.format = scenario.format_id if format_override is None else format_override
        tabs = scenario.tables()
        cfg = _lib.SynthConfig(scenario.samples_per_epoch, self.format, 0 if scenario.noise else 1, 0,
                               scenario.seed & (2**64 - 1))
        h = C.c_void_p()
        torch.cuda.set_device(device)
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

    def window(self, first_epoch, epochs, rows_dev, masks_dev, mask_len, gain, out_dev, clips=True, stream=None):
        """Plain launch of `epochs` epochs; returns per-epoch clip counts (numpy) or None."""
        torch = _torch()
        L = _lib.lib()
        nch = len(self.scenario.channels)
        clipped = (C.c_uint64 * epochs)() if clips else None
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream(self.device).cuda_stream
        check(L.pharos_synth_window(self.handle, first_epoch, epochs, nch, rows_dev.data_ptr() if nch else None,
                                    masks_dev.data_ptr(), mask_len, float(gain), out_dev.data_ptr(), clipped, s))
        return np.frombuffer(clipped, dtype=np.uint64).copy() if clips else None


def upload(array, device=0):
    """A numpy array's bytes as a device uint8 tensor."""
    torch = _torch()
    a = np.ascontiguousarray(array)
    t = torch.from_numpy(a.view(np.uint8).reshape(-1).copy())
    return t.to(f"cuda:{device}")


def run_plain(scenario, first_epoch=0, epochs=None, device=0, format_override=None, window=100, gain=None):
    """Synthesises epochs with per-window plain launches and the host AGC between windows (the plain per-run form).

    Returns (bytes as numpy uint8, clip counts per epoch, final AGC gain)."""
    torch = _torch()
    epochs = scenario.epochs if epochs is None else epochs
    eng = Engine(scenario, device, format_override)
    rows, masks = scenario.window(first_epoch, epochs)
    d_rows, d_masks = upload(rows, device), upload(masks, device)
    out = torch.empty(epochs * eng.bytes_per_epoch, dtype=torch.uint8, device=f"cuda:{device}")
    nch = len(scenario.channels)
    agc_gain = 1.0
    constant = scenario.gain_db is not None or gain is not None
    g = gain if gain is not None else (float(np.float32(10.0 ** (scenario.gain_db / 20.0))) if scenario.gain_db is not None else 1.0)
    clips = np.zeros(epochs, dtype=np.uint64)
    acc_clip, acc_comp = 0, 0
    e = 0
    while e < epochs:
        absolute = first_epoch + e
        n = min(window - absolute % window, epochs - e)
        gw = g if constant else float(np.float32(agc_gain))
        c = eng.window(absolute, n, d_rows[e * nch * 64:], d_masks, len(masks), gw, out[e * eng.bytes_per_epoch:])
        clips[e:e + n] = c
        if not constant:  # gain.zig:19-35
            acc_clip += int(c.sum())
            acc_comp += n * scenario.samples_per_epoch * 2
            if (absolute + n) % 100 == 0:
                rate = acc_clip / acc_comp
                acc_clip = acc_comp = 0
                if rate > 0.01:
                    agc_gain *= 0.95
                elif rate < 0.001 and agc_gain < 1.0:
                    agc_gain = min(agc_gain * 1.02, 1.0)
        e += n
    torch.cuda.synchronize(device)
    result = out.cpu().numpy()
    eng.close()
    return result, clips, agc_gain
