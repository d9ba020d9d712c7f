"""The per-block polynomial hand-off (pharos_synth_segment): the host stage's cubics against its own channel model,
the oracle's segment derive pinned in closed form, and rows from segments against rows from the model.

The segments replace one row per channel and epoch by one record per channel and 100-epoch block; the device
evaluates the rows (tests/test_gpu_derive.py holds the device against the oracle bit for bit)."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib, scenario as S


def horner(c, tau):
    v = c[..., 3]
    v = v * tau + c[..., 2]
    v = v * tau + c[..., 1]
    return v * tau + c[..., 0]


@pytest.mark.parametrize("make,sub", [(S.c1, 0.0), (S.c2, 0.37), (S.c0, 0.0)])
def test_cubics_follow_the_channel_model(make, sub):
    sc = make(seed=2, epochs=1000)
    sc.start_gps_sub_ms = sub
    L = _lib.lib()
    spec, arr = sc.spec()
    first, epochs = 700, 300
    segs, _ = sc.segments(first, epochs, 100)
    worst_phase = worst_delay = 0.0
    for c in range(len(sc.channels)):
        for e in range(0, epochs + 1, 7):
            b = min(e // 100, segs.shape[0] - 1)
            tau = e - 100 * b
            t = sc.start_gps_ms + first + e
            ph, ms, sb = C.c_double(), C.c_int64(), C.c_double()
            L.pharos_scenario_endpoint(C.byref(spec), C.byref(arr[c]), t, C.byref(ph), C.byref(ms), C.byref(sb))
            g = segs[b, c]
            worst_phase = max(worst_phase, abs(horner(g["phase"], float(tau)) - ph.value))
            # the model's delay less the sub-ms start, from its transmit-time split
            worst_delay = max(worst_delay, abs((horner(g["delay"], float(tau)) - sub) - ((t - ms.value) - sb.value)))
    assert worst_phase < 1e-6, worst_phase      # cycles
    assert worst_delay < 1e-9, worst_delay      # ms


def test_segment_derive_in_closed_form():
    # Delay 70.25 ms falling 0.5 us per ms, phase falling 1.5 cycles per ms, receive time 123456.5 ms (sub 0.5):
    # transmit time = 123456.5 - 70.25 = 123386.25, so tx_ms 123386 and 0.25 of a ms.
    g = np.zeros(1, dtype=_lib.SEGMENT_DTYPE)
    g["phase"] = [[10.0, -1.5, 0.0, 0.0]]
    g["delay"] = [[70.25, -0.0005, 0.0, 0.0]]
    g["amp"] = [[0.125, 0.0]]
    g["if_hz"] = 1_250_000
    g["code_ms"] = 1
    g["entries_per_ms"] = 1023
    g["table"] = 2
    g["mask_first"] = 123_380
    g["mask_count"] = 20
    g["mask_offset"] = 500
    row = np.zeros(1, dtype=O.ROW_DTYPE)
    spe = 4092
    assert O.lib().orc_derive_segment(g.ctypes.data, 123_456, 0.5, 0, spe, row.ctypes.data) == 1
    r = row[0]
    assert r["period0"] == 123_386 and r["table"] == 2
    assert abs(r["entry_phase0"] - 0.25 * 1023) < 1e-9
    assert abs(r["entries_per_sample"] - (1 + 0.0005) * 1023 / spe) < 1e-12   # the code runs fast as the delay falls
    assert abs(r["carrier_step"] - (1_250_000 / (spe * 1000.0) + 1.5 / spe)) < 1e-15
    # ifPhase at 123456.5 ms: 1.25e6 * 123.4565 s = 154320625 cycles exactly, so phase0 = frac(0 - 10) = 0
    assert abs(((r["carrier_phase0"] + 0.5) % 1.0) - 0.5) < 1e-9
    assert r["amp"] == np.float32(0.125)
    assert r["mask_offset"] == 506 and r["mask_count"] == 14
    # A period outside the segment's mask range leaves the row active with no mask bytes.
    g["mask_first"] = 123_390
    assert O.lib().orc_derive_segment(g.ctypes.data, 123_456, 0.5, 0, spe, row.ctypes.data) == 1
    assert row[0]["mask_count"] == 0


def test_rows_from_segments_follow_rows_from_the_model():
    sc = S.c1(seed=4, epochs=600)
    first, epochs = 200, 250
    segs, smasks = sc.segments(first, epochs, 100)
    rows_s = O.derive_segments(segs, sc.start_gps_ms, sc.start_gps_sub_ms, first, epochs, 100, sc.samples_per_epoch)
    rows_m, mmasks = sc.window(first, epochs)
    assert (rows_s["table"] == rows_m["table"]).all()
    assert (rows_s["period0"] == rows_m["period0"]).all()
    d = np.abs(rows_s["carrier_phase0"] - rows_m["carrier_phase0"])
    assert np.minimum(d, 1 - d).max() < 1e-6
    assert np.abs(rows_s["entry_phase0"] - rows_m["entry_phase0"]).max() < 2e-5
    assert np.abs(rows_s["entries_per_sample"] / rows_m["entries_per_sample"] - 1).max() < 1e-9
    assert (rows_s["amp"] == rows_m["amp"]).all()
    # The same mask bytes for every period a row can read.
    for e in range(0, epochs, 17):
        for c in range(len(sc.channels)):
            a, b = rows_s[e, c], rows_m[e, c]
            n = min(int(a["mask_count"]), int(b["mask_count"]))
            assert n >= 3
            sa = smasks[int(a["mask_offset"]):int(a["mask_offset"]) + n]
            sb = mmasks[int(b["mask_offset"]):int(b["mask_offset"]) + n]
            assert np.array_equal(sa, sb), (e, c)


def test_segments_cost_less_host_time_than_rows():
    import time
    sc = S.c1(seed=1, epochs=60_000)
    sc.segments(0, 100)
    t0 = time.perf_counter()
    sc.segments(0, 6000)
    t_seg = time.perf_counter() - t0
    t0 = time.perf_counter()
    sc.window(0, 6000)
    t_rows = time.perf_counter() - t0
    assert t_seg * 5 < t_rows, (t_seg, t_rows)
