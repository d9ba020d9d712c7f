"""The reference's trajectory cache files (.pht, src/pht/read.zig and write.zig) through the library's reader and
writer (csrc/trajectory.cpp over the cache header and codec of csrc/cache_codec.h)."""
import ctypes as C

import numpy as np

from . import _lib


class CacheRefused(ValueError):
    """A file the reference would refuse: .code is its diagnostic (600-604), .offset the byte it names."""

    def __init__(self, code, offset, message):
        super().__init__(f"{code} at byte {offset}: {message}")
        self.code, self.offset = code, offset


def read(data):
    """(name, knots) of a .pht file's bytes; knots is a KNOT_DTYPE array (t, pos, vel, acc)."""
    L = _lib.lib()
    data = bytes(data)
    n, fault = C.c_size_t(), _lib.CacheFault()
    name = C.create_string_buffer(4096)
    st = L.pharos_trajectory_read(data, len(data), None, 0, C.byref(n), name, len(name), C.byref(fault))
    if st == 5:
        raise CacheRefused(fault.code, fault.offset, fault.message.decode(errors="replace"))
    knots = np.zeros(n.value, dtype=_lib.KNOT_DTYPE)
    if n.value:
        st = L.pharos_trajectory_read(data, len(data), knots.ctypes.data, n.value, C.byref(n), name, len(name),
                                      C.byref(fault))
    if st:
        raise RuntimeError(f"trajectory read: status {st}")
    return name.value.decode(), knots


def write(name, knots):
    """The .pht bytes of a trajectory."""
    L = _lib.lib()
    knots = np.ascontiguousarray(knots, dtype=_lib.KNOT_DTYPE)
    n = C.c_size_t()
    L.pharos_trajectory_write(name.encode(), knots.ctypes.data, len(knots), None, 0, C.byref(n))
    out = np.zeros(n.value, dtype=np.uint8)
    st = L.pharos_trajectory_write(name.encode(), knots.ctypes.data, len(knots), out.ctypes.data, n.value,
                                   C.byref(n))
    if st:
        raise RuntimeError(f"trajectory write: status {st}")
    return out.tobytes()
