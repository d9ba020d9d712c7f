"""Pins of the oracle's per-sample kernel: code sources, rotor, noise, quantisers, AGC and derive.

Each check compares the oracle with something it does not share code with: a direct floor formula for the
Q32.32 code walk, FP64/long-double trigonometry for the rotor, published SplitMix64 outputs and moments for the
noise, the documented quantiser tables (docs/guide/samples.md:120-136), the documented AGC rule and exact rational
arithmetic for derive.
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import binding as O

L = O.lib()


def _row(table=0, phase0=0.0, step=0.0, entry_phase0=0.0, eps=0.25, period0=1000, amp=0.5, mask_offset=0,
         mask_count=64, lead_bits=0, lead_entries=0.0):
    r = np.zeros(1, dtype=O.ROW_DTYPE)
    r["carrier_phase0"], r["carrier_step"], r["entry_phase0"], r["entries_per_sample"] = phase0, step, entry_phase0, eps
    r["period0"], r["table"], r["amp"], r["mask_offset"], r["mask_count"] = period0, table, amp, mask_offset, mask_count
    r["lead_bits"], r["lead_entries"] = lead_bits, lead_entries
    return r


def _direct_values(tables, row, masks, n0, length):
    idx = np.zeros(length, dtype=np.uint32)
    L.orc_entry_direct(tables.handles[0], row.ctypes.data, masks.ctypes.data, n0, length, idx.ctypes.data)
    info = O.signal_info(SIG_OF[id(tables)])
    amp = float(row["amp"][0])
    vre = np.array([np.float32(amp * info.values_re[k]) for k in range(8)], dtype=np.float32)
    vim = np.array([np.float32(amp * info.values_im[k]) for k in range(8)], dtype=np.float32)
    return vre[idx], vim[idx]


SIG_OF = {}

CASES = [
    # signal, prn, eps, lead bits (0 = none), lead entries
    ("L1CA", 5, 0.25, 0, 0.0),           # expand (c0)
    ("L1CA", 7, 1023 / 61440 * 1.00002, 0, 0.0),  # runs
    ("L1CA", 9, 0.9, 0, 0.0),            # gather
    ("L5", 3, 10230 / 61440, 0, 0.0),    # expand, two axes
    ("L5", 4, 10230 / 16368, 0, 0.0),    # expand near the limit
    ("L5", 6, 10230 / 4092, 0, 0.0),     # gather
    ("E1", 11, 12276 / 61440, 0, 0.0),   # 8 values, expand
    ("E1", 12, 12276 / 4092, 0, 0.0),    # 8 values, gather
    ("B1C", 2, 12276 / 30690, 6, 0.37),  # lead (pilot leads), expand
    ("B1C", 3, 12276 / 4092, 6, 1.81),   # lead, gather
    ("L5", 8, 10230 / 61440, 2, 2.5),    # lead of the pilot component, expand
    ("L5", 9, 10230 / 8000, 1, 0.2),     # lead of the data component, gather
    ("L2C", 1, 1023 / 4092, 0, 0.0),     # TDM expand
    ("L2C", 2, 1023 / 1000, 0, 0.0),     # TDM gather
    ("G1", 4, 511 / 61440, 0, 0.0),      # runs
]


@pytest.mark.parametrize("sig,prn,eps,lead,lead_entries", CASES)
def test_sources_follow_the_direct_floor_formula(sig, prn, eps, lead, lead_entries):
    rng = np.random.default_rng(prn * 131 + int(eps * 1e6))
    tabs = O.Tables([(sig, prn)])
    SIG_OF[id(tabs)] = sig
    per = L.orc_table_entries_per_period(tabs.handles[0])
    masks = rng.integers(0, 4, size=64).astype(np.uint8)
    for trial in range(6):
        e0 = float(rng.uniform(0, per)) if trial else float(per) - 2.0 ** -33  # frac -> 2^32 - 1 clamp
        row = _row(entry_phase0=e0, eps=eps * float(rng.uniform(0.97, 1.03)), lead_bits=lead, lead_entries=lead_entries,
                   period0=int(rng.integers(0, 10**6)), amp=float(rng.uniform(0.01, 1.0)))
        n0, length = (0, 2560) if trial % 2 == 0 else (2560, 1532)
        dre, dim = _direct_values(tabs, row, masks, n0, length)
        axis = L.orc_table_axis(tabs.handles[0])
        if axis == 1:
            dim[:] = 0
        if axis == 2:
            dre[:] = 0
        for path in (0, 1, 2):
            if path == 0 and (lead or sig == "L2C"):
                continue
            xr = np.zeros(length, dtype=np.float32)
            xi = np.zeros(length, dtype=np.float32)
            got = L.orc_source_values(tabs.handles[0], row.ctypes.data, masks.ctypes.data, n0, length, path,
                                      xr.ctypes.data, xi.ctypes.data)
            assert got == path
            assert np.array_equal(xr, dre) and np.array_equal(xi, dim), (sig, trial, path)


def test_rotor_phase_error_against_fp64_is_below_1e6_cycles():
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(40):
        phase0 = float(rng.uniform(0, 1))
        step = float(rng.uniform(-0.5, 0.5))  # up to half a cycle per sample
        n0 = int(rng.integers(0, 240)) * 512
        n = 4096
        cr = np.zeros(n, dtype=np.float32)
        ci = np.zeros(n, dtype=np.float32)
        L.orc_rotor_values(phase0, step, n0, n, cr.ctypes.data, ci.ctypes.data)
        k = np.arange(n0, n0 + n, dtype=np.longdouble)
        ph = (np.longdouble(phase0) + k * np.longdouble(step)) % 1
        got = np.angle(cr.astype(np.float64) + 1j * ci.astype(np.float64)) / (2 * np.pi)
        err = (got - ph.astype(np.float64) + 0.5) % 1 - 0.5
        worst = max(worst, float(np.abs(err).max()))
        assert np.abs(np.abs(cr.astype(np.float64) + 1j * ci) - 1).max() < 2e-6
    assert worst < 1e-6, worst


def test_rotor_reseeds_on_the_epoch_grid():
    # At every multiple of 512 samples from the epoch start the rotor equals f32 of the f64 seed.
    phase0, step = 0.3141, 0.2718
    cr = np.zeros(4096, dtype=np.float32)
    ci = np.zeros(4096, dtype=np.float32)
    L.orc_rotor_values(phase0, step, 1024, 4096, cr.ctypes.data, ci.ctypes.data)
    c, s = C.c_double(), C.c_double()
    for j in range(0, 4096, 512):
        L.orc_sincos64(phase0 + (1024 + j) * step, C.byref(c), C.byref(s))
        assert cr[j] == np.float32(c.value) and ci[j] == np.float32(s.value)


def test_sincos64_is_accurate():
    c, s = C.c_double(), C.c_double()
    rng = np.random.default_rng(1)
    for phi in list(rng.uniform(-50, 50, 2000)) + [0.0, 0.25, 0.5, 1 - 2**-52]:
        L.orc_sincos64(phi, C.byref(c), C.byref(s))
        a = 2 * math.pi * (phi - math.floor(phi))
        assert abs(c.value - math.cos(a)) < 2e-15 and abs(s.value - math.sin(a)) < 2e-15


def test_splitmix64_published_outputs():
    # SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c) from state 0: the first outputs.
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [L.orc_splitmix(0, k) for k in (1, 2, 3)] == expect


def test_log_polynomial_error_bound():
    xs = np.concatenate([np.geomspace(2.0**-30, 2.0**32, 20000), [1.0, 2.0, 0.5, 2.0**32 - 256]]).astype(np.float32)
    err = max(abs(L.orc_logf(C.c_float(x)) - math.log(float(x))) for x in xs)
    assert err < 2e-5, err  # the reference fits it to 1.3e-5 (noise.zig:24) plus f32 rounding


def test_noise_is_unit_gaussian_and_chunk_free():
    re4 = (C.c_float * 4)()
    im4 = (C.c_float * 4)()
    vals = []
    for g in range(50_000):
        L.orc_noise_group(12345, g, re4, im4)
        vals.extend(complex(re4[k], im4[k]) for k in range(4))
    v = np.array(vals)
    n = len(v)
    assert abs(v.real.mean()) < 4 / math.sqrt(n) and abs(v.imag.mean()) < 4 / math.sqrt(n)
    assert abs(v.real.var() - 1) < 0.02 and abs(v.imag.var() - 1) < 0.02
    assert abs(np.mean(v.real * v.imag)) < 0.02
    # Kurtosis of a Gaussian: 3.
    assert abs(np.mean(v.real**4) / v.real.var()**2 - 3) < 0.1


def _doc_quantise(fmt, re, im, gain):
    """The documented quantisers (docs/guide/samples.md:120-136), one component at a time, in float32."""
    g = np.float32(gain)
    out, clips = [], 0
    for a, b in zip(re.astype(np.float32), im.astype(np.float32)):
        comps = []
        for x in (a, b):
            if fmt in (2, 3):
                scale = np.float32(25.0 if fmt == 2 else 3277.0)
                lo, hi = (-128, 127) if fmt == 2 else (-32768, 32767)
                y = np.float32(x * np.float32(g * scale))
                t = math.trunc(float(y))
                clips += t > hi or t < lo
                comps.append(max(lo, min(hi, t)))
            elif fmt == 1:
                m = np.float32(np.float32(abs(x)) * np.float32(g * np.float32(3.0)))
                t = math.trunc(float(m))
                clips += t > 7
                comps.append(min(t, 7) | (8 if x < 0 else 0))
            else:
                thr = np.float32(np.float32(1.1) / g)
                clips += abs(x) >= np.float32(np.float32(5.0) * thr)
                comps.append((2 if x < 0 else 0) | (1 if abs(x) >= thr else 0))
        out.append(comps)
    if fmt == 2:
        return bytes(v & 0xFF for p in out for v in p), clips
    if fmt == 3:
        return b"".join(int(v).to_bytes(2, "little", signed=True) for p in out for v in p), clips
    if fmt == 1:
        return bytes((p[0] << 4) | p[1] for p in out), clips
    codes = [p[0] | (p[1] << 2) for p in out]
    return bytes(codes[2 * i] | (codes[2 * i + 1] << 4) for i in range(len(codes) // 2)), clips


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
@pytest.mark.parametrize("gain", [1.0, 0.35, 2.5])
def test_quantisers_follow_the_documented_tables(fmt, gain):
    rng = np.random.default_rng(fmt * 10 + int(gain * 10))
    g = np.float32(gain)
    scale = {0: np.float32(1.0), 1: np.float32(3.0), 2: np.float32(25.0), 3: np.float32(3277.0)}[fmt]
    # Values on and around every code boundary, plus zeros of both signs and wide random values.
    edges = np.arange(-130, 131, dtype=np.float32) / np.float32(g * scale)
    if fmt == 0:
        thr = np.float32(np.float32(1.1) / g)
        edges = np.array([thr, -thr, 5 * thr, -5 * thr, np.float32(5.0) * thr], dtype=np.float32)
    near = np.concatenate([edges, np.nextafter(edges, np.float32(np.inf)), np.nextafter(edges, np.float32(-np.inf))])
    vals = np.concatenate([near, [0.0, -0.0], rng.normal(0, 4, 4000).astype(np.float32),
                           rng.normal(0, 60, 500).astype(np.float32)]).astype(np.float32)
    vals = vals[: len(vals) // 8 * 8]
    re = vals
    im = np.roll(vals, 3)
    out = np.zeros(len(vals) * 4, dtype=np.uint8)
    clipped = L.orc_pack(fmt, re.ctypes.data, im.ctypes.data, len(vals), C.c_float(gain), out.ctypes.data)
    expect, exp_clips = _doc_quantise(fmt, re, im, gain)
    assert bytes(out[:len(expect)]) == expect
    assert clipped == exp_clips


def test_pack_tail_is_zero_padded():
    re = np.array([0.5, -0.7, 0.1], dtype=np.float32)
    im = np.array([-0.2, 0.3, -0.9], dtype=np.float32)
    for fmt in (1, 2, 3):
        out = np.zeros(32, dtype=np.uint8)
        L.orc_pack(fmt, re.ctypes.data, im.ctypes.data, 3, C.c_float(1.0), out.ctypes.data)
        expect, _ = _doc_quantise(fmt, re, im, 1.0)
        assert bytes(out[:len(expect)]) == expect and not out[len(expect):].any()


def test_agc_trace_follows_the_documented_rule():
    # docs/design/synthesis.md: every 100 epochs; rate > 1 % -> x0.95; rate < 0.1 % and gain < 1 -> x1.02 (<= 1).
    agc = O.Agc(1.0, 0, 0)
    comps = 100 * 4000 * 2
    rate = C.c_double()
    trace = []
    for epoch_after, clipped in [(50, 1000), (100, 9000), (200, 100), (300, 0), (400, 80000), (500, 200)]:
        changed = L.orc_agc_account(C.byref(agc), epoch_after, clipped, comps // (2 if epoch_after == 50 or epoch_after == 100 else 1), C.byref(rate))
        trace.append((changed, round(agc.gain, 12)))
    g1 = 0.95
    g2 = g1 * 1.02
    g3 = g2 * 1.02
    g4 = g3 * 0.95
    g5 = min(g4 * 1.02, 1.0)
    assert trace == [(0, 1.0), (1, round(g1, 12)), (1, round(g2, 12)), (1, round(g3, 12)), (1, round(g4, 12)),
                     (1, round(g5, 12))]


def test_db_linear():
    for db in (-60.0, -6.0, 0.0, 3.0, 60.0):
        assert L.orc_db_linear(db) == np.float32(10 ** (db / 20))


def test_derive_closed_form():
    # Exact rational check of every field of parameters.derive on synthetic endpoints (parameters.zig:51-70).
    rng = np.random.default_rng(3)
    for _ in range(200):
        if_hz = int(rng.integers(-30_000_000, 30_000_000))
        spe = int(rng.choice([4092, 16368, 61440, 122880]))
        epm, code_ms = (1023, 1) if rng.random() < 0.5 else (12276, 4)
        ch = O.ChannelConstants(if_hz, code_ms, epm, 0, 3)
        t_rx = int(rng.integers(10**11, 10**12))
        tx_ms = t_rx - int(rng.integers(65, 90))
        a = O.Endpoint(float(rng.uniform(-1e5, 1e5)), tx_ms, float(rng.uniform(0, 1)))
        span = float(rng.uniform(-2e-5, 2e-5))
        sub_b = a.tx_sub_ms + span
        b = O.Endpoint(a.phase_cycles - float(rng.uniform(-6, 6)), tx_ms + 1 + int(math.floor(sub_b)), sub_b - math.floor(sub_b))
        row = np.zeros(1, dtype=O.ROW_DTYPE)
        cn0 = float(rng.uniform(30, 55))
        assert L.orc_derive(C.byref(ch), tx_ms, C.byref(a), C.byref(b), t_rx, 0.0, spe, cn0, row.ctypes.data) == 1
        r = row[0]
        # carrier phase: IF phase of the receiver time, exact on whole ms, less the endpoint phase
        ifp = Fraction((if_hz % 1000) * (t_rx % 1000) % 1000, 1000)
        ph = float(ifp) - a.phase_cycles
        assert r["carrier_phase0"] == ph - math.floor(ph)
        assert r["carrier_step"] == if_hz / (spe * 1000.0) + (a.phase_cycles - b.phase_cycles) / spe
        span_ms = (b.tx_ms - a.tx_ms) + (b.tx_sub_ms - a.tx_sub_ms)
        assert r["entry_phase0"] == ((tx_ms % code_ms) + a.tx_sub_ms) * epm
        assert r["entries_per_sample"] == span_ms * epm / spe
        assert r["period0"] == tx_ms // code_ms
        amp = float(r["amp"])
        assert abs(amp - math.sqrt(2 * 10 ** (cn0 / 10) / (1000 * spe))) <= 1e-7 * amp


def test_derive_rejects_rates_beyond_three_percent_and_negative_scale():
    ch = O.ChannelConstants(0, 1, 1023, 0, 0)
    row = np.zeros(1, dtype=O.ROW_DTYPE)
    a = O.Endpoint(0.0, 1000, 0.5)
    for span, ok in ((1.0299, 1), (1.0301, 0), (0.9701, 1), (0.9699, 0)):
        b = O.Endpoint(0.0, 1000 + int(0.5 + span), (0.5 + span) % 1)
        assert L.orc_derive(C.byref(ch), 1000, C.byref(a), C.byref(b), 1100, 0.0, 4092, 45.0, row.ctypes.data) == ok
    b = O.Endpoint(0.0, 1001, 0.5)
    assert L.orc_derive(C.byref(ch), -1, C.byref(a), C.byref(b), 1100, 0.0, 4092, 45.0, row.ctypes.data) == 0


def test_window_is_thread_count_invariant():
    from pharos_amd import scenario as S
    sc = S.c1(seed=4, epochs=3)
    rows, masks = sc.window(0, 3)
    tabs = O.Tables([(c.signal, c.prn) for c in sc.channels])
    a = O.synth_window(tabs, rows, masks, sc.samples_per_epoch, sc.format_id, 0, 9, 0, 1.0, threads=1)
    b = O.synth_window(tabs, rows, masks, sc.samples_per_epoch, sc.format_id, 0, 9, 0, 1.0, threads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_code_position_error_against_fp64_over_an_epoch():
    """Q32.32 code position (chip.zig:12-29) against the exact position entry_phase0 + n * eps over a whole
    122 880-sample epoch: within (n + 1) * 2^-33 entries, so at most 1.43e-5 entries (chips for one entry per
    chip); the entry index equals floor of the exact position away from such ties."""
    rng = np.random.default_rng(7)
    worst, ties = 0.0, 0
    for _ in range(100):
        ep0, eps = float(rng.uniform(0, 10230)), float(rng.uniform(0.001, 3.0))
        fl = math.floor(ep0)
        phase_q = min(int(round((ep0 - fl) * 2**32)), 2**32 - 1)
        eps_q = int(round(eps * 2**32))
        for k in list(rng.integers(0, 122_880, 200)) + [122_879]:
            k = int(k)
            exact = np.longdouble(ep0) + np.longdouble(k) * np.longdouble(eps)
            q = fl + (phase_q + k * eps_q) / 2**32
            err = abs(float(np.longdouble(q) - exact))
            worst = max(worst, err)
            assert err <= (k + 1) * 2.0**-33 + 1e-12
            index = fl + ((phase_q + k * eps_q) >> 32)
            if abs(float(exact - np.floor(exact))) > 2e-5:
                assert index == int(np.floor(exact))
            else:
                ties += 1
    assert worst < 1.5e-5
