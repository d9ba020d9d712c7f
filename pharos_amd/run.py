"""pharos_run.h from Python: a run created from a scenario bundle, pulled, sought and commanded (ctypes over
libpharos_amd.so; every epoch is synthesised on the GPU)."""
import ctypes as C

import numpy as np

from . import _lib

STATUS = _lib.RUN_STATUS


class RunError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"status {status}: {message}")
        self.status = status


class Run:
    def __init__(self, bundle, seed=0, noise=0, gain=0, gain_db=0.0, workers=0):
        L = _lib.lib()
        self.bundle = bytes(bundle)
        opts = _lib.RunOptions(workers, seed & (2**64 - 1), noise, gain, gain_db)
        h = C.c_void_p()
        st = L.pharos_run_create(self.bundle, len(self.bundle), C.byref(opts), C.byref(h))
        if st:
            raise RunError(st, L.pharos_run_last_error().decode(errors="replace"))
        self.handle = h

    def _check(self, st, ok=(0,)):
        if st not in ok:
            raise RunError(st, self.error())
        return st

    def error(self):
        return _lib.lib().pharos_run_error(self.handle).decode(errors="replace")

    def info(self):
        i = _lib.RunInfo()
        self._check(_lib.lib().pharos_run_info_get(self.handle, C.byref(i)))
        return i

    def pull(self, max_epochs, cap=None):
        """(status, block fields, bytes, applied commands) of one pull."""
        info = self.info()
        cap = max_epochs * info.bytes_per_epoch if cap is None else cap
        buf = np.zeros(max(cap, 1), dtype=np.uint8)
        b = _lib.RunBlock()
        st = _lib.lib().pharos_run_pull(self.handle, buf.ctypes.data, cap, max_epochs, C.byref(b))
        applied = [(b.applied[k].epoch, b.applied[k].command.kind, b.applied[k].command.value)
                   for k in range(b.applied_count)]
        return st, (b.epoch_first, b.epochs, b.bytes, b.clipped), buf[:b.bytes].copy(), applied

    def pull_all(self, chunk):
        out = []
        while True:
            st, blk, data, _ = self.pull(chunk)
            if st == STATUS["end"]:
                return np.concatenate(out) if out else np.zeros(0, dtype=np.uint8)
            self._check(st)
            out.append(data)

    def seek(self, epoch):
        return _lib.lib().pharos_run_seek(self.handle, epoch)

    def submit(self, kind, value=0.0, apply_at_epoch=0, system=0, svid=0, unit=0, use_default=0):
        c = _lib.RunCommand(apply_at_epoch, value, kind, system, svid, unit, use_default)
        return _lib.lib().pharos_run_submit(self.handle, C.byref(c))

    def events(self):
        out, ev = [], _lib.RunEvent()
        while _lib.lib().pharos_run_next_event(self.handle, C.byref(ev)):
            out.append((ev.kind, ev.epoch, ev.a, ev.b, ev.c, ev.code))
        return out

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().pharos_run_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
