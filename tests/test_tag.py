"""The .tag sidecar (pharos_synth_tag_text; output/tag.zig:12-34, docs/guide/samples.md:139-158): the seven lines
with their exact spacing, the GPS calendar time of the first sample (no leap seconds) against Python's calendar,
six-decimal MHz values, no form for IQ2/IQ4, and pharos_run_tag_text's buffer rules (c_api/run.zig:299-305)."""
import ctypes as C
import datetime as dt

import numpy as np

from pharos_amd import _lib

L = _lib.lib()
IQ2, IQ4, IQ8, IQ16 = 0, 1, 2, 3
GPS_EPOCH = dt.datetime(1980, 1, 6)


def tag(fmt, fs_hz, center_hz, first_ms, cap=512):
    buf = C.create_string_buffer(cap)
    n = C.c_size_t(0)
    st = L.pharos_synth_tag_text(fmt, fs_hz, center_hz, first_ms, buf, cap, C.byref(n))
    return st, buf.raw[: n.value].decode() if st == 0 else None, n.value


def test_layout_at_the_gps_epoch():
    st, text, n = tag(IQ8, 4_092_000, 1_575_420_000, 0)
    assert st == 0 and n == len(text)
    assert text == ("PROG = pharos-run\n"
                    "TIME = 1980/01/06 00:00:00.000\n"
                    "FMT  = CS8\n"
                    "F_S  = 4.092000\n"
                    "F_LO = 1575.420000\n"
                    "IQ   = 2\n"
                    "BITS = 8\n")


def test_iq16_and_negative_centre():
    st, text, _ = tag(IQ16, 122_880_000, -1_500_000, 1)
    assert st == 0
    lines = text.splitlines()
    assert lines[1] == "TIME = 1980/01/06 00:00:00.001"
    assert lines[2:] == ["FMT  = CS16", "F_S  = 122.880000", "F_LO = -1.500000", "IQ   = 2", "BITS = 16"]


def test_time_is_the_gps_calendar():
    rng = np.random.default_rng(8)
    cases = [int(x) for x in rng.integers(0, 60 * 365 * 86_400_000, 500)]
    cases += [86_400_000 - 1, 86_400_000, 59 * 86_400_000, 1_400_000_000_000, 1_700_000_000_123]
    for ms in cases:
        st, text, _ = tag(IQ8, 61_440_000, 0, ms)
        t = GPS_EPOCH + dt.timedelta(milliseconds=ms)
        want = f"TIME = {t:%Y/%m/%d %H:%M:%S}.{t.microsecond // 1000:03d}"
        assert st == 0 and text.splitlines()[1] == want, ms


def test_no_form_for_two_and_four_bit_pairs():
    for fmt in (IQ2, IQ4):
        st, _, n = tag(fmt, 16_368_000, 0, 0)
        assert st == 2 and n == 0


def test_buffer_rules():
    n = C.c_size_t(0)
    assert L.pharos_synth_tag_text(IQ8, 4_092_000, 0, 0, None, 0, C.byref(n)) == 2
    full = n.value
    assert full == len(tag(IQ8, 4_092_000, 0, 0)[1])
    buf = C.create_string_buffer(full)
    assert L.pharos_synth_tag_text(IQ8, 4_092_000, 0, 0, buf, full - 1, C.byref(n)) == 2 and n.value == full
    assert L.pharos_synth_tag_text(IQ8, 4_092_000, 0, 0, buf, full, C.byref(n)) == 0
