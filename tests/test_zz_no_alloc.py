"""The window entry point that allocates nothing (pharos_synth_window_static) against the plain window
(pharos_synth_window): the same bytes and per-epoch clip counts, on two calls in a row with the second smaller,
and a call beyond its static scratch refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from pharos_amd import _lib, scenario as S
from pharos_amd.engine import DeviceBuffer, Engine

pytestmark = pytest.mark.gpu


def both_forms(sc, first, epochs, gain):
    L = _lib.lib()
    rows, masks = sc.window(first, epochs)
    eng = Engine(sc)
    d_rows, d_masks = DeviceBuffer.from_array(rows), DeviceBuffer.from_array(masks)
    nch = len(sc.channels)
    res = []
    for fn in (L.pharos_synth_window, L.pharos_synth_window_static):
        out = DeviceBuffer(epochs * eng.bytes_per_epoch)
        clipped = (C.c_uint64 * epochs)()
        st = fn(eng.handle, first, epochs, nch, d_rows.ptr, d_masks.ptr, len(masks), gain, out.ptr, clipped, None)
        assert st == 0, _lib.lib().pharos_synth_last_error()
        res.append((out.to_numpy(), np.frombuffer(clipped, dtype=np.uint64).copy()))
    eng.close()
    return res


def test_static_window_equals_the_plain_window():
    print("no-alloc: c1 subset 120 epochs, then c0 37 epochs", flush=True)
    big = S.c1(seed=3, epochs=1000)
    (b0, c0), (b1, c1) = both_forms(big, 400, 120, 0.8)
    assert np.array_equal(b0, b1) and np.array_equal(c0, c1)
    assert c0.sum() > 0  # IQ4 with noise clips: the counts are compared, not just zeros
    small = S.c0(seed=4, epochs=100)
    small.noise = True
    (b0, c0), (b1, c1) = both_forms(small, 55, 37, 1.7)
    assert np.array_equal(b0, b1) and np.array_equal(c0, c1)


def test_static_window_refuses_calls_beyond_its_scratch():
    L = _lib.lib()
    sc = S.c1(seed=3, epochs=10)
    eng = Engine(sc)
    tiles = -(-sc.samples_per_epoch // 4096)
    epochs = (1 << 18) // tiles + 1
    out = DeviceBuffer(16)
    st = L.pharos_synth_window_static(eng.handle, 0, epochs, len(sc.channels), out.ptr, out.ptr, 16, 1.0, out.ptr,
                                      None, None)
    assert st == 2 and b"static scratch" in L.pharos_synth_last_error()
    eng.close()
