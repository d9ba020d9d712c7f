"""Pins of the truncated Gold-family primary codes (E5a, E5b, B1I/B2I, B3I, B2a, B2b) and of the Galileo memory-code
bit order, from the ICDs' own descriptions rather than the generators' constants.

Each ICD defines these codes by two shift registers given as polynomials, a start state per code, an output
stage and (for the 13-stage GPS/BDS registers) a short cycle of the first register at 8190 chips.  The tests
below write each register from its polynomial exactly as printed in the ICD, then check:
- the oracle's and the product's codes equal that generator chip for chip for every PRN;
- every polynomial is primitive: its register runs a maximal-length sequence (period 2^n - 1 from all ones and
  2^(n-1) ones per period), which a mistyped tap almost never keeps;
- the first n chips are the start state read from the output stage down (closed form of a Fibonacci register),
  which fixes the output stage and the shift direction;
- the ICDs' published start values where they are known here (E5a-I PRN 1, octal 30305), B1I's two-valued
  Gold structure, and the first 192 chips of E1-B PRN 1 from the ICD's hexadecimal listing.
The per-code start-state tables themselves are ICD data both generators read (pharos_amd/csrc/icd_tables.h); the
L5 tables are pinned by the XB code advances (tests/test_oracle_codes.py), the others only by these properties.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ICD_HEADER = os.path.join(ROOT, "pharos_amd", "csrc", "icd_tables.h")

# Register polynomials as the ICDs print them (exponents of the non-constant terms).
POLY = {
    # Galileo OS SIS ICD, Table 15 (octal 40503 / 50661 / 64021 / 51445 / 43143)
    "E5A_1": (14, [1, 6, 8, 14]), "E5A_2": (14, [4, 5, 7, 8, 12, 14]),
    "E5B_1": (14, [4, 11, 13, 14]), "E5BI_2": (14, [2, 5, 8, 9, 12, 14]), "E5BQ_2": (14, [1, 5, 6, 9, 10, 14]),
    # BDS-SIS-ICD-B1I: G1 = 1+x+x^7+x^8+x^9+x^10+x^11, G2 = 1+x+x^2+x^3+x^4+x^5+x^8+x^9+x^11
    "B1I_1": (11, [1, 7, 8, 9, 10, 11]), "B1I_2": (11, [1, 2, 3, 4, 5, 8, 9, 11]),
    # BDS-SIS-ICD-B3I: G1 = 1+x+x^3+x^4+x^13, G2 = 1+x+x^5+x^6+x^7+x^9+x^10+x^12+x^13
    "B3I_1": (13, [1, 3, 4, 13]), "B3I_2": (13, [1, 5, 6, 7, 9, 10, 12, 13]),
    # BDS-SIS-ICD-B2a: data 1+x+x^5+x^11+x^13 / 1+x^3+x^5+x^9+x^11+x^12+x^13; pilot 1+x^3+x^6+x^7+x^13 /
    # 1+x+x^5+x^7+x^8+x^12+x^13
    "B2AD_1": (13, [1, 5, 11, 13]), "B2AD_2": (13, [3, 5, 9, 11, 12, 13]),
    "B2AP_1": (13, [3, 6, 7, 13]), "B2AP_2": (13, [1, 5, 7, 8, 12, 13]),
    # BDS-SIS-ICD-B2b: G1 = 1+x+x^9+x^10+x^13, G2 = 1+x^3+x^4+x^6+x^9+x^12+x^13
    "B2B_1": (13, [1, 9, 10, 13]), "B2B_2": (13, [3, 4, 6, 9, 12, 13]),
}

# (signal, component, register 1, register 2, start-state table, first-register short cycle or None, PRNs)
GOLD = [
    ("E5a", "data", "E5A_1", "E5A_2", "ICD_E5A_REG2_I", None, 50),
    ("E5a", "pilot", "E5A_1", "E5A_2", "ICD_E5A_REG2_Q", None, 50),
    ("E5b", "data", "E5B_1", "E5BI_2", "ICD_E5B_REG2_I", None, 50),
    ("E5b", "pilot", "E5B_1", "E5BQ_2", "ICD_E5B_REG2_Q", None, 50),
    ("B3I", "data", "B3I_1", "B3I_2", "ICD_B3I_G2_INIT", 8190, 63),
    ("B2a", "data", "B2AD_1", "B2AD_2", "ICD_B2A_DATA_G2_INIT", 8190, 63),
    ("B2a", "pilot", "B2AP_1", "B2AP_2", "ICD_B2A_PILOT_G2_INIT", 8190, 63),
    ("B2b", "data", "B2B_1", "B2B_2", "ICD_B2B_G2_INIT", 8190, 63),
]
CHIPS = {"E5a": 10230, "E5b": 10230, "B3I": 10230, "B2a": 10230, "B2b": 10230, "B1I": 2046, "B2I": 2046}


def icd_table(name):
    """An integer table of icd_tables.h (C literals: hexadecimal, octal when they start with 0, else decimal)."""
    text = open(ICD_HEADER).read()
    m = re.search(name + r"\[\d+\]\s*=\s*\{([^}]*)\}", text)
    assert m, name
    return [int(v, 0) if v.startswith("0x") else (int(v, 8) if len(v) > 1 and v.startswith("0") else int(v))
            for v in re.findall(r"0x[0-9A-Fa-f]+|\d+", m.group(1))]


class Register:
    """A Fibonacci shift register as the ICDs draw it: stages 1..n, stage i held in bit i-1 of `state`; each chip
    shifts stage i into stage i+1, feeds stage 1 with the sum of the stages the polynomial names, and outputs
    stage n (or, with `select`, the sum of the selected stages)."""

    def __init__(self, poly, state=None):
        self.n, exps = POLY[poly]
        self.mask = (1 << self.n) - 1
        self.taps = sum(1 << (e - 1) for e in exps)
        self.state = self.mask if state is None else state

    def sequence(self, count, select=None, reset_at=None):
        out = np.empty(count, dtype=np.uint8)
        s = self.state
        for i in range(count):
            if reset_at is not None and i == reset_at:
                s = self.mask
            out[i] = (bin(s & select).count("1") & 1) if select else (s >> (self.n - 1)) & 1
            s = ((s << 1) | (bin(s & self.taps).count("1") & 1)) & self.mask
        return out


def gold_code(r1, r2, start, count, reset_at):
    return Register(r1).sequence(count, reset_at=reset_at) ^ Register(r2, start).sequence(count)


def product_code(signal, prn, pilot):
    buf = (C.c_uint8 * 20000)()
    n = _lib.lib().pharos_primary_code(_lib.SIGNAL[signal], prn, 1 if pilot else 0, buf, len(buf))
    assert n > 0
    return np.frombuffer(buf, dtype=np.uint8, count=n).copy()


@pytest.mark.parametrize("name", sorted(POLY))
def test_register_polynomials_are_primitive(name):
    n = POLY[name][0]
    period = (1 << n) - 1
    seq = Register(name).sequence(period + n)
    assert int(seq[:period].sum()) == 1 << (n - 1)       # an m-sequence has 2^(n-1) ones per period
    assert np.array_equal(seq[period:], seq[:n])          # it repeats after 2^n - 1 chips
    # ... and not before: the n-chip window at the start occurs once per period (each nonzero state once).
    windows = np.lib.stride_tricks.sliding_window_view(seq[:period + n - 1], n)
    keys = windows.dot(1 << np.arange(n, dtype=np.int64))
    assert len(np.unique(keys)) == period


@pytest.mark.parametrize("signal,kind,r1,r2,table,reset,prns", GOLD)
def test_gold_codes_follow_the_icd_registers(signal, kind, r1, r2, table, reset, prns):
    starts = icd_table(table)
    assert len(starts) == prns
    n = POLY[r2][0]
    sats = {"E5a": 36, "E5b": 36}.get(signal, 63)  # satellites the catalogue numbers (Galileo 1-36)
    for prn in range(1, min(prns, sats) + 1):
        expect = gold_code(r1, r2, starts[prn - 1], CHIPS[signal], reset)
        assert np.array_equal(O.code(kind, signal, prn), expect), (signal, kind, prn)
        assert np.array_equal(product_code(signal, prn, kind == "pilot"), expect), (signal, kind, prn)
        # Closed form of the first n chips: register 1 starts all ones, so chip i is 1 xor stage n - i of
        # register 2's start state.
        first = [1 ^ ((starts[prn - 1] >> (n - 1 - i)) & 1) for i in range(n)]
        assert list(expect[:n]) == first


def test_short_cycle_restarts_the_first_register_at_8190():
    # IS-GPS-705 / BDS B3I, B2a, B2b: the first register returns to all ones after 8190 chips, so the code xor the
    # second register repeats its first 2040 chips from chip 8190.
    for signal, kind, r1, r2, table, reset, _ in GOLD:
        if reset is None:
            continue
        start = icd_table(table)[0]
        g1 = O.code(kind, signal, 1) ^ Register(r2, start).sequence(CHIPS[signal])
        free = Register(r1).sequence(8190)           # register 1 from all ones, not yet reset
        assert np.array_equal(g1[:8190], free), signal
        assert np.array_equal(g1[8190:], free[:CHIPS[signal] - 8190]), signal


def test_e5a_i_prn1_start_value_is_the_icds():
    # Galileo OS SIS ICD, E5a-I primary code 1: register 2 start value 30305 (octal).
    assert icd_table("ICD_E5A_REG2_I")[0] == 0o30305


B1I_PHASES = ["1+3", "1+4", "1+5", "1+6", "1+8", "1+9", "1+10", "1+11", "2+7", "3+4", "3+5", "3+6", "3+8", "3+9",
              "3+10", "3+11", "4+5", "4+6", "4+8", "4+9", "4+10", "4+11", "5+6", "5+8", "5+9", "5+10", "5+11",
              "6+8", "6+9", "6+10", "6+11", "8+9", "8+10", "8+11", "9+10", "9+11", "10+11"]


def b1i_select(prn):
    taps = [t for t in icd_table("ICD_B1I_PHASE")[3 * (prn - 1):3 * prn] if t]
    return sum(1 << (t - 1) for t in taps)


def test_b1i_codes_follow_the_icd_registers():
    # BDS-SIS-ICD-B1I: G1 and G2 both start at 01010101010; the code is G1 xor the sum of the G2 stages the
    # phase assignment names, truncated to 2046 chips.  Phase assignments 1..37 are the ICD's two-stage sums.
    for prn, pair in enumerate(B1I_PHASES, 1):
        assert b1i_select(prn) == sum(1 << (int(t) - 1) for t in pair.split("+")), prn
    for prn in range(1, 64):
        g1 = Register("B1I_1", 0b01010101010).sequence(2046)
        g2 = Register("B1I_2", 0b01010101010).sequence(2046, select=b1i_select(prn))
        assert np.array_equal(O.code("data", "B1I", prn), g1 ^ g2), prn
        assert np.array_equal(product_code("B1I", prn, False), g1 ^ g2), prn
        assert np.array_equal(O.code("data", "B2I", prn), g1 ^ g2), prn   # B2I carries the B1I ranging code


def test_b1i_untruncated_codes_are_a_gold_family():
    # A sum of G2 stages is a shift of G2's m-sequence, so with G1 and G2 a preferred pair the 2047-chip codes
    # have the three-valued cross-correlation {-1, -65, 63} of an 11-stage Gold family.
    codes = []
    for prn in range(1, 64, 3):
        g1 = Register("B1I_1", 0b01010101010).sequence(2047)
        g2 = Register("B1I_2", 0b01010101010).sequence(2047, select=b1i_select(prn))
        codes.append(1 - 2 * (g1 ^ g2).astype(np.int64))
    f = np.fft.fft(np.array(codes), axis=1)
    for i in range(len(codes)):
        for j in range(i, len(codes)):
            corr = np.rint(np.fft.ifft(f[i] * np.conj(f[j])).real).astype(int)
            vals = set(corr[1:] if i == j else corr)
            assert vals <= {-1, -65, 63}, (i, j, vals)


# Galileo OS SIS ICD Annex C: E1-B primary code 1 in hexadecimal, first chip in the top bit of the first digit.
E1B_PRN1_HEX = "F5D710130573541B9DBD4FD9E9B20A0D59D144C54BC79355"


def test_e1_memory_code_bit_order_matches_the_icd_listing():
    expect = np.array([int(b) for b in bin(int(E1B_PRN1_HEX, 16))[2:].zfill(4 * len(E1B_PRN1_HEX))], dtype=np.uint8)
    assert np.array_equal(O.code("data", "E1", 1)[:len(expect)], expect)
    assert np.array_equal(product_code("E1", 1, False)[:len(expect)], expect)
    # The reversed bit order inside each byte would give another code.
    rev = expect.reshape(-1, 8)[:, ::-1].reshape(-1)
    assert not np.array_equal(O.code("data", "E1", 1)[:len(expect)], rev)
