"""parameters.derive (src/products/run/synthesis/parameters.zig:51-81): the oracle pinned by exact arithmetic, and
the product's host stage (pharos_derive, pharos_amd/csrc/scenario.cpp) against the oracle row for row.

ifPhase is the IF carrier's phase at the receive time t = ms + sub_ms, in cycles: frac(if_hz * t / 1000).  The
reference splits it into an exact integer part on whole milliseconds and an f64 part on the sub-millisecond
fraction; the pin evaluates frac(if_hz * t / 1000) with exact rationals, so a dropped or mis-scaled term fails.
"""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib
from pharos_amd.scenario import Channel, Scenario, LEAP_SECONDS

BDS_OFFSET_MS = 14_000            # domain/time/gps.zig:20
MOSCOW_OFFSET_MS = 3 * 3_600_000  # UTC + 3 h


def if_phase_of(if_hz, t_ms, sub_ms, spe=4092):
    """ifPhase through orc_derive: endpoint a at phase 0 makes carrier_phase0 = ifPhase(if_hz, t_rx)."""
    L = O.lib()
    ch = O.ChannelConstants(if_hz, 1, 1023, 0, 0)
    a = O.Endpoint(0.0, 1_000_000, 0.0)
    b = O.Endpoint(0.0, 1_000_001, 0.0)
    row = np.zeros(1, dtype=O.ROW_DTYPE)
    assert L.orc_derive(C.byref(ch), 1_000_000, C.byref(a), C.byref(b), t_ms, sub_ms, spe, 45.0, row.ctypes.data)
    return float(row["carrier_phase0"][0])


def test_if_phase_is_the_exact_carrier_phase_at_the_receive_time():
    rng = np.random.default_rng(5)
    for _ in range(400):
        if_hz = int(rng.integers(-30_720_000, 30_720_000))
        t_ms = int(rng.integers(0, 2**41))
        sub = float(rng.choice([0.0, 0.5, 0.25, 0.125, 0.999755859375, rng.random()]))
        exact = (Fraction(if_hz) * (Fraction(t_ms) + Fraction(sub)) / 1000) % 1
        got = if_phase_of(if_hz, t_ms, sub)
        err = abs(Fraction(got) - exact)
        err = min(err, 1 - err)  # the phase is taken mod 1
        assert err < Fraction(1, 10**9), (if_hz, t_ms, sub, got, float(exact))


def test_derive_closed_form_for_a_static_channel():
    # Constant range: both endpoints carry the same phase and sub-ms transmit time, so the code advances exactly
    # one ms per epoch and the carrier step is the IF alone.
    L = O.lib()
    for sig, if_hz, spe in (("L1CA", 0, 4092), ("E1", -18_000_000, 61_440), ("B3I", 24_000_000, 61_440)):
        info = O.signal_info(sig)
        ch = O.ChannelConstants(if_hz, info.code_period_ms, info.entries_per_ms, info.system, 3)
        base = 1_234_567_891
        a = O.Endpoint(0.375, base, 0.8125)
        b = O.Endpoint(0.375, base + 1, 0.8125)
        row = np.zeros(1, dtype=O.ROW_DTYPE)
        assert L.orc_derive(C.byref(ch), base, C.byref(a), C.byref(b), base + 70, 0.0, spe, 40.0, row.ctypes.data)
        r = row[0]
        assert r["period0"] == base // info.code_period_ms
        assert r["entry_phase0"] == (base % info.code_period_ms + 0.8125) * info.entries_per_ms
        assert r["entries_per_sample"] == info.entries_per_ms / spe
        assert r["carrier_step"] == if_hz / (spe * 1000.0)
        assert r["table"] == 3
        exact_if = (Fraction(if_hz) * (base + 70) / 1000) % 1
        assert abs(float(r["carrier_phase0"]) - float((exact_if - Fraction(0.375)) % 1)) < 1e-12


def test_derive_rejects_spans_outside_three_percent_and_negative_times():
    L = O.lib()
    ch = O.ChannelConstants(0, 1, 1023, 0, 0)
    row = np.zeros(1, dtype=O.ROW_DTYPE)
    for span_sub, ok in ((0.0, 1), (0.0299, 1), (0.0301, 0), (-0.0299, 1), (-0.0301, 0)):
        a = O.Endpoint(0.0, 100, 0.5)
        b = O.Endpoint(0.0, 101, 0.5 + span_sub)
        assert L.orc_derive(C.byref(ch), 100, C.byref(a), C.byref(b), 170, 0.0, 4092, 45.0, row.ctypes.data) == ok
    a, b = O.Endpoint(0.0, -1, 0.5), O.Endpoint(0.0, 0, 0.5)
    assert L.orc_derive(C.byref(ch), -1, C.byref(a), C.byref(b), 70, 0.0, 4092, 45.0, row.ctypes.data) == 0


def base_ms(system, tx_ms):
    """GpsTime.baseMs (domain/time/gps.zig:109-115) with a fixed leap count (civil.zig:225-229)."""
    if system == 1:
        return tx_ms - BDS_OFFSET_MS
    if system == 3:
        return tx_ms - LEAP_SECONDS * 1000 + MOSCOW_OFFSET_MS
    return tx_ms


CASES = [("L1CA", 7, 0, 0), ("E1", 12, -18_000_000, 0), ("B1I", 8, -6_000_000, 0), ("G1", 3, 6_000_000 - 3 * 562_500, -3),
         ("L5", 24, 18_000_000, 0), ("B3I", 33, 24_000_000, 0), ("E5b", 19, 15_345_000, 0)]


@pytest.mark.parametrize("sub_ms", [0.0, 0.37, 0.9990234375])
def test_product_rows_equal_the_oracle(sub_ms):
    L, Lo = _lib.lib(), O.lib()
    chans = [Channel(s, p, freq_k=k, if_hz=f, range0_m=20_500_000.0 + 611_111.0 * i, doppler_hz=-3100.0 + 900.0 * i,
                     doppler_rate_hz_s=0.7, wobble_accel_mps2=2.5, wobble_period_s=31.4, cn0_dbhz=36.0 + 2.0 * i)
             for i, (s, p, f, k) in enumerate(CASES)]
    sc = Scenario("derive", 61_440, "IQ4", 0, chans, start_gps_sub_ms=sub_ms)
    spec, arr = sc.spec()
    for c, ch in enumerate(chans):
        info = O.signal_info(ch.signal)
        cc = O.ChannelConstants(ch.if_hz, info.code_period_ms, info.entries_per_ms, info.system, c)
        for epoch in (0, 1, 999, 12_345, 59_999):
            got = np.zeros(1, dtype=_lib.ROW_DTYPE)
            active = L.pharos_derive(C.byref(spec), C.byref(arr[c]), c, epoch, got.ctypes.data)
            t_rx = sc.start_gps_ms + epoch
            ends = []
            for t in (t_rx, t_rx + 1):
                ph, ms, sub = C.c_double(), C.c_int64(), C.c_double()
                L.pharos_scenario_endpoint(C.byref(spec), C.byref(arr[c]), t, C.byref(ph), C.byref(ms), C.byref(sub))
                assert 0.0 <= sub.value < 1.0
                ends.append(O.Endpoint(ph.value, ms.value, sub.value))
            want = np.zeros(1, dtype=O.ROW_DTYPE)
            ok = Lo.orc_derive(C.byref(cc), base_ms(info.system, ends[0].tx_ms), C.byref(ends[0]), C.byref(ends[1]),
                               t_rx, sub_ms, sc.samples_per_epoch, ch.cn0_dbhz, want.ctypes.data)
            assert active == ok == 1
            for field in ("carrier_phase0", "carrier_step", "entry_phase0", "entries_per_sample", "period0", "table",
                          "amp"):
                assert got[field][0] == want[field][0], (ch.signal, epoch, field, got[field][0], want[field][0])


def test_fractional_start_shifts_the_transmit_time_split():
    # The same channel model started 0.37 ms later in the ms: the range at a given epoch is unchanged (the model runs
    # on time since the start), so the transmit time (whole ms + sub) moves by exactly the 0.37 ms.
    L = _lib.lib()
    ch = Channel("L1CA", 1, range0_m=21_000_000.0, doppler_hz=500.0)
    res = []
    for sub in (0.0, 0.37):
        spec, arr = Scenario("t", 4092, "IQ8", 0, [ch], start_gps_sub_ms=sub).spec()
        ph, ms, s = C.c_double(), C.c_int64(), C.c_double()
        L.pharos_scenario_endpoint(C.byref(spec), C.byref(arr[0]), spec.start_gps_ms + 100, C.byref(ph), C.byref(ms),
                                   C.byref(s))
        res.append((ms.value, s.value))
    assert abs((res[1][0] - res[0][0]) + (res[1][1] - res[0][1]) - 0.37) < 1e-9
