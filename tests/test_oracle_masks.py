"""Symbol masks (SymbolSource.fill, src/products/run/symbols/source.zig:53-91): the oracle's restatement pinned in
closed form, then the product's host stage (pharos_amd/csrc/scenario.cpp fill_masks) against the oracle.

A mask byte of code period k is (nav bit ^ NH chip) | secondary[k % len] << 1.  The closed-form cases state
each signal's layout from its ICD (bit length, NH code as published, frame length and offset) and check
the oracle's masks against it; a mistake in the frame/bit/NH index arithmetic shifts some byte.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib
from pharos_amd.scenario import Channel, Scenario

NH20 = [int(c) for c in "00000100110101001110"]   # BDS D1 NH code; GPS L5Q NH20 (IS-GPS-705)
NH10 = [int(c) for c in "0000110101"]             # GPS L5I NH10 (IS-GPS-705)
CS25 = [int(c) for c in format(0x380AD90, "028b")[:25]]  # Galileo E1-C secondary code, hex 380AD90
E5A_CS20 = [int(c) for c in format(0x842E9, "020b")]     # Galileo E5a-I secondary CS20, hex 842E9


def random_bits(frames, bits, seed):
    return np.random.default_rng(seed).integers(0, 2, size=(frames, bits), dtype=np.uint8)


def test_b1i_meo_masks_are_the_nh_code_with_no_nav_bits():
    # D1 (MEO/IGSO): 1 ms code periods, 20-period bits carrying the NH20 code, 6000 ms frames.
    k0 = 3 * 6000 + 7 * 20
    m = O.symbol_masks("B1I", 10, k0, 60)
    assert list(m) == NH20 * 3


def test_b1i_geo_masks_use_the_geo_frame():
    # D2 (GEO, PRN <= 5 or >= 59): 2 ms bits, no NH code, 600 ms frames.
    for prn in (3, 60):
        bits = random_bits(4, 300, prn)
        k0 = 600 * 5 + 17
        m = O.symbol_masks("B1I", prn, k0, 1500, bits, frame0=5)
        k = np.arange(k0, k0 + 1500)
        f, b = k // 600 - 5, (k % 600) // 2
        expect = np.where(f < 4, bits[np.minimum(f, 3), b], 0)
        assert np.array_equal(m, expect), prn
    # A MEO satellite of the same signal keeps the 20 ms NH bits.
    assert O.symbol_locate("B1I", 30, 6000 + 45) == (1, 2, NH20[5])


def test_l5_masks_combine_nav_nh10_and_the_nh20_pilot_code():
    bits = random_bits(3, 600, 1)        # 6 s frames of 10 ms bits
    k0 = 6000 + 123
    m = O.symbol_masks("L5", 4, k0, 9000, bits, frame0=1)
    k = np.arange(k0, k0 + 9000)
    f, in_frame = k // 6000 - 1, k % 6000
    nav = bits[f, in_frame // 10]
    expect = (nav ^ np.array(NH10)[in_frame % 10]) | (np.array(NH20)[k % 20] << 1)
    assert np.array_equal(m, expect)


def test_e1_masks_follow_the_1000_ms_frame_offset_and_cs25():
    # E1: 4 ms code periods, one symbol per period, 2 s pages that start 1 s into the code-period count.
    bits = random_bits(6, 500, 2)
    k0 = 1000
    m = O.symbol_masks("E1", 11, k0, 2000, bits, frame0=2)
    k = np.arange(k0, k0 + 2000)
    ms = 4 * k + 1000
    f, b = ms // 2000 - 2, (ms % 2000) // 4
    expect = bits[f, b] | (np.array(CS25)[k % 25] << 1)
    assert np.array_equal(m, expect)
    assert O.symbol_locate("E1", 11, 0) == (0, 250, 0)   # period 0 sits half way through page 0


def test_e5a_masks_use_the_cs20_nh_code():
    k0 = 10_000 * 2
    m = O.symbol_masks("E5a", 7, k0, 40)
    sec = O.code("secondary", "E5a", 7)
    expect = np.array(E5A_CS20 * 2) | (sec[np.arange(k0, k0 + 40) % 100] << 1)
    assert np.array_equal(m, expect)


def test_frames_without_bits_read_as_zero():
    bits = np.ones((1, 300), dtype=np.uint8)         # only frame 0 of L1 C/A (20 ms bits, 6 s frames) has bits
    m = O.symbol_masks("L1CA", 5, 5990, 20, bits, frame0=0)
    assert list(m) == [1] * 10 + [0] * 10


def nav_frame_bits(ch, signal, prn, k_first, k_last):
    """The stand-in's nav bits over the frames periods k_first..k_last span, as a [frames, bits] array."""
    L = _lib.lib()
    spec = ch.spec()
    f0 = O.symbol_locate(signal, prn, k_first)[0]
    f1 = O.symbol_locate(signal, prn, k_last)[0]
    info = O.signal_info(signal)
    geo = info.system == 1 and info.geo_frame_length_ms and (prn <= 5 or prn >= 59)
    nh_len = info.geo_nh_length if geo else info.nh_length
    frame_ms = info.geo_frame_length_ms if geo else info.frame_length_ms
    bpf = -(-frame_ms // (info.code_period_ms * nh_len))
    bits = np.array([[L.pharos_scenario_nav_bit(C.byref(spec), f, b) for b in range(bpf)]
                     for f in range(f0, f1 + 1)], dtype=np.uint8)
    return bits, f0


CHANNELS = [("L5", 4), ("E1", 11), ("E5a", 7), ("B1I", 3), ("B3I", 20), ("B2a", 61), ("L1CA", 9), ("G1", 2),
            ("E5b", 30), ("L1C", 5)]


@pytest.mark.parametrize("signal,prn", CHANNELS)
def test_product_masks_equal_the_oracle(signal, prn):
    ch = Channel(signal, prn, range0_m=21_345_678.0, doppler_hz=1234.5, cn0_dbhz=44.0, nav_seed=0xC0FFEE + prn)
    sc = Scenario("masks", 30_690, "IQ8", 0, [ch], noise=False)
    first = 5_900 if signal != "L1C" else 17_950  # windows that cross a frame edge
    rows, masks = sc.window(first, 250)
    active = rows[:, 0][rows[:, 0]["table"] != _lib.INACTIVE]
    assert len(active) == 250
    k_first = int(active["period0"].min())
    k_last = int((active["period0"] + active["mask_count"]).max())
    bits, f0 = nav_frame_bits(ch, signal, prn, k_first, k_last)
    expect = O.symbol_masks(signal, prn, k_first, k_last - k_first, bits, frame0=f0)
    for r in active:
        got = masks[int(r["mask_offset"]):int(r["mask_offset"]) + int(r["mask_count"])]
        k = int(r["period0"]) - k_first
        assert np.array_equal(got, expect[k:k + int(r["mask_count"])]), (signal, int(r["period0"]))
