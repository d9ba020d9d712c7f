"""The three places where the path's numbers come from a math library rather than from its own arithmetic:

- the f64 turn table of the rotor (`rotor.zig:64-72`), made by Zig's comptime `@cos` / `@sin`;
- `amplitude` (`parameters.zig:78-81`): `f32(sqrt(2 * std.math.pow(f64, 10, cn0 / 10) / fs))`;
- `Db.linear` (`gain.zig:50-52`): `f32(std.math.pow(f64, 10, db / 20))`.

Zig's toolchain is not in this image, so these tests pin the values from outside instead of from a run of the
reference:

- Zig evaluates comptime `@sin` / `@cos` with its compiler_rt (or musl) routines, which are Sun's fdlibm
  algorithm (`__rem_pio2`, `__sin`, `__cos`).  That algorithm is restated below from its published form, with
  its constants checked against their published bit patterns, and bounded against a correctly rounded value
  (decimal arithmetic at 60 digits).  The oracle's table (C, `oracle/synth.c`) must equal this restatement at
  every entry, and so must the engine's (its own C++ restatement, read back through the C ABI).
- `std.math.pow` is Zig's port of Go's `math.Pow`: `exp(yf * log x)` for the fractional part of y times x**yi
  by repeated squaring.  Its f64 result can differ from the host `pow` in the last bits, but only its f32
  rounding reaches the path.  The tests evaluate that algorithm with two exp/log pairs (fdlibm's, restated
  here, and the host's), the host `pow` and a correctly rounded value.  They check that all give the same f32
  over dense grids of the inputs the path takes.  They also check that each f64 result lies far from an f32
  rounding boundary.
"""
import ctypes as C
import math
import struct
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import binding as O

L = O.lib()


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _from_bits(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def _hi(x):
    return _bits(x) >> 32


# ---------------------------------------------------------------- fdlibm sin / cos (Sun Microsystems, 1993)
# Kernel polynomials on [-pi/4, pi/4] and the three-part split of pi/2 (fdlibm k_sin.c, k_cos.c, e_rem_pio2.c).
S1, S2, S3 = -1.66666666666666324348e-01, 8.33333333332248946124e-03, -1.98412698298579493134e-04
S4, S5, S6 = 2.75573137070700676789e-06, -2.50507602534068634195e-08, 1.58969099521155010221e-10
C1, C2, C3 = 4.16666666666666019037e-02, -1.38888888888741095749e-03, 2.48015872894767294178e-05
C4, C5, C6 = -2.75573143513906633035e-07, 2.08757232129817482790e-09, -1.13596475577881948265e-11
PIO2_1, PIO2_1T = 1.57079632673412561417e+00, 6.07710050650619224932e-11
PIO2_2, PIO2_2T = 6.07710050630396597660e-11, 2.02226624879595063154e-21
PIO2_3, PIO2_3T = 2.02226624871116645580e-21, 8.47842766036889956997e-32
INVPIO2 = 6.36619772367581382433e-01
PIO4 = _from_bits(0x3FE921FB54442D18)
TOINT = 1.5 / 2.0 ** -52

# The published high words of the constants, as a check on their decimal forms above.
_CONST_BITS = {S1: 0xBFC5555555555549, S2: 0x3F8111111110F8A6, S3: 0xBF2A01A019C161D5, S4: 0x3EC71DE357B1FE7D,
               S5: 0xBE5AE5E68A2B9CEB, S6: 0x3DE5D93A5ACFD57C, C1: 0x3FA555555555554C, C2: 0xBF56C16C16C15177,
               C3: 0x3EFA01A019CB1590, C4: 0xBE927E4F809C52AD, C5: 0x3E21EE9EBDB4B1C4, C6: 0xBDA8FAE9BE8838D4,
               PIO2_1: 0x3FF921FB54400000, PIO2_1T: 0x3DD0B4611A626331, PIO2_2: 0x3DD0B4611A600000,
               PIO2_2T: 0x3BA3198A2E037073, PIO2_3: 0x3BA3198A2E000000, PIO2_3T: 0x397B839A252049C1,
               INVPIO2: 0x3FE45F306DC9C883}


def _k_sin(x, y, iy):
    z = x * x
    w = z * z
    r = S2 + z * (S3 + z * S4) + z * w * (S5 + z * S6)
    v = z * x
    if iy == 0:
        return x + v * (S1 + z * r)
    return x - ((z * (0.5 * y - v * r) - y) - v * S1)


def _k_cos(x, y):
    z = x * x
    w = z * z
    r = z * (C1 + z * (C2 + z * C3)) + w * w * (C4 + z * (C5 + z * C6))
    hz = 0.5 * z
    w = 1.0 - hz
    return w + (((1.0 - w) - hz) + (z * r - x * y))


def _rem_pio2(x):
    """(n, y0, y1) with x = n * pi/2 + y0 + y1, for 0 < x < 2^20 * pi/2 (the cases the table needs)."""
    ix = _hi(x) & 0x7FFFFFFF
    assert x > 0 and ix < 0x413921FB

    def near(k):
        z = x - k * PIO2_1
        y0 = z - k * PIO2_1T
        return k, y0, (z - y0) - k * PIO2_1T

    if ix <= 0x400F6A7A:  # |x| ~<= 5pi/4
        if (ix & 0xFFFFF) != 0x921FB:  # not ~pi/2 or ~pi: no cancellation
            return near(1) if ix <= 0x4002D97C else near(2)
    elif ix <= 0x401C463B:  # |x| ~<= 9pi/4
        if ix <= 0x4015FDBC:
            if ix != 0x4012D97C:  # not ~3pi/2
                return near(3)
        elif ix != 0x401921FB:  # not ~2pi
            return near(4)
    # medium: n = rint(x / (pi/2)), then one to three rounds of the three-part split
    fn = x * INVPIO2 + TOINT - TOINT
    n = int(fn)
    r = x - fn * PIO2_1
    w = fn * PIO2_1T
    if r - w < -PIO4:
        n, fn = n - 1, fn - 1
        r = x - fn * PIO2_1
        w = fn * PIO2_1T
    elif r - w > PIO4:
        n, fn = n + 1, fn + 1
        r = x - fn * PIO2_1
        w = fn * PIO2_1T
    y0 = r - w
    ex = ix >> 20
    if ex - ((_bits(y0) >> 52) & 0x7FF) > 16:
        t = r
        w = fn * PIO2_2
        r = t - w
        w = fn * PIO2_2T - ((t - r) - w)
        y0 = r - w
        if ex - ((_bits(y0) >> 52) & 0x7FF) > 49:
            t = r
            w = fn * PIO2_3
            r = t - w
            w = fn * PIO2_3T - ((t - r) - w)
            y0 = r - w
    return n, y0, (r - y0) - w


def fdlibm_sin(x):
    ix = _hi(x) & 0x7FFFFFFF
    if ix <= 0x3FE921FB:
        return x if ix < 0x3E500000 else _k_sin(x, 0.0, 0)
    n, y0, y1 = _rem_pio2(x)
    return [_k_sin(y0, y1, 1), _k_cos(y0, y1), -_k_sin(y0, y1, 1), -_k_cos(y0, y1)][n & 3]


def fdlibm_cos(x):
    ix = _hi(x) & 0x7FFFFFFF
    if ix <= 0x3FE921FB:
        return 1.0 if ix < 0x3E46A09E else _k_cos(x, 0.0)
    n, y0, y1 = _rem_pio2(x)
    return [_k_cos(y0, y1), -_k_sin(y0, y1, 1), -_k_cos(y0, y1), _k_sin(y0, y1, 1)][n & 3]


# ---------------------------------------------------------------- fdlibm exp / log (e_exp.c, e_log.c)
LN2_HI, LN2_LO, INVLN2 = 6.93147180369123816490e-01, 1.90821492927058770002e-10, 1.44269504088896338700e+00
P1, P2, P3 = 1.66666666666666019037e-01, -2.77777777770155933842e-03, 6.61375632143793436117e-05
P4, P5 = -1.65339022054652515390e-06, 4.13813679705723846039e-08
LG1, LG2, LG3 = 6.666666666666735130e-01, 3.999999999940941908e-01, 2.857142874366239149e-01
LG4, LG5, LG6 = 2.222219843214978396e-01, 1.818357216161805012e-01, 1.531383769920937332e-01
LG7 = 1.479819860511658591e-01
_CONST_BITS.update({LN2_HI: 0x3FE62E42FEE00000, LN2_LO: 0x3DEA39EF35793C76, INVLN2: 0x3FF71547652B82FE,
                    P1: 0x3FC555555555553E, P2: 0xBF66C16C16BEBD93, P3: 0x3F11566AAF25DE2C, P4: 0xBEBBBD41C5D26BF1,
                    P5: 0x3E66376972BEA4D0, LG1: 0x3FE5555555555593, LG2: 0x3FD999999997FA04,
                    LG3: 0x3FD2492494229359, LG4: 0x3FCC71C51D8E78AF, LG5: 0x3FC7466496CB03DE,
                    LG6: 0x3FC39A09D078C69F, LG7: 0x3FC2F112DF3E5244})


def fdlibm_exp(x):
    hx = _hi(x)
    sign = hx >> 31
    hx &= 0x7FFFFFFF
    assert hx < 0x4086232B  # |x| < 708.39: no overflow or underflow handling needed here
    if hx > 0x3FD62E42:  # |x| > ln2 / 2
        k = int(INVLN2 * x + (0.5, -0.5)[sign]) if hx >= 0x3FF0A2B2 else 1 - sign - sign
        hi = x - k * LN2_HI
        lo = k * LN2_LO
        x = hi - lo
    elif hx > 0x3E300000:
        k, hi, lo = 0, x, 0.0
    else:
        return 1 + x
    xx = x * x
    c = x - xx * (P1 + xx * (P2 + xx * (P3 + xx * (P4 + xx * P5))))
    y = 1 + (x * c / (2 - c) - lo + hi)
    return y if k == 0 else math.ldexp(y, k)


def fdlibm_log(x):
    b = _bits(x)
    hx = b >> 32
    assert 0x00100000 <= hx < 0x7FF00000  # positive, normal, finite
    if hx == 0x3FF00000 and (b & 0xFFFFFFFF) == 0:
        return 0.0
    hx += 0x3FF00000 - 0x3FE6A09E
    k = (hx >> 20) - 0x3FF
    hx = (hx & 0x000FFFFF) + 0x3FE6A09E
    x = _from_bits((hx << 32) | (b & 0xFFFFFFFF))
    f = x - 1.0
    hfsq = 0.5 * f * f
    s = f / (2.0 + f)
    z = s * s
    w = z * z
    t1 = w * (LG2 + w * (LG4 + w * LG6))
    t2 = z * (LG1 + w * (LG3 + w * (LG5 + w * LG7)))
    r = t2 + t1
    dk = float(k)
    return s * (hfsq + r) + dk * LN2_LO - hfsq + f + dk * LN2_HI


def go_pow(x, y, exp=fdlibm_exp, log=fdlibm_log):
    """Go's math.Pow for x > 0 and finite y (the form Zig's std.math.pow follows)."""
    if y == 0 or x == 1:
        return 1.0
    if y == 1:
        return x
    if y == 0.5:
        return math.sqrt(x)
    if y == -0.5:
        return 1 / math.sqrt(x)
    yf, yi = math.modf(abs(y))
    a1, ae = 1.0, 0
    if yf != 0:
        if yf > 0.5:
            yf -= 1
            yi += 1
        a1 = exp(yf * log(x))
    x1, xe = math.frexp(x)
    i = int(yi)
    while i != 0:
        if i & 1:
            a1 *= x1
            ae += xe
        x1 *= x1
        xe <<= 1
        if x1 < 0.5:
            x1 += x1
            xe -= 1
        i >>= 1
    if y < 0:
        a1 = 1 / a1
        ae = -ae
    return math.ldexp(a1, ae)


# ---------------------------------------------------------------- correctly rounded references
getcontext().prec = 60


def _dsin_dcos(x):
    """sin and cos of the exact binary value of x, to 60 digits (Taylor series; |x| < 7)."""
    d = Decimal(x)
    d2 = d * d
    s, c = Decimal(0), Decimal(0)
    ts, tc = d, Decimal(1)
    k = 0
    while True:
        s += ts
        c += tc
        ts = -ts * d2 / ((2 * k + 2) * (2 * k + 3))
        tc = -tc * d2 / ((2 * k + 1) * (2 * k + 2))
        k += 1
        if abs(ts) < Decimal(10) ** -70 and abs(tc) < Decimal(10) ** -70:
            return s, c


def _rounding_margin(exact, rounded):
    """Distance of `exact` from the nearest f64 rounding boundary, in ulps of `rounded` (0.5 means it is a
    double)."""
    if rounded == 0.0:
        return 0.5
    up = math.nextafter(rounded, math.inf)
    dn = math.nextafter(rounded, -math.inf)
    hi = (Decimal(rounded) + Decimal(up)) / 2
    lo = (Decimal(rounded) + Decimal(dn)) / 2
    ulp = Decimal(up) - Decimal(rounded)
    return float(min(hi - exact, exact - lo) / ulp)


# ---------------------------------------------------------------- tests
def test_constants_match_their_published_bit_patterns():
    for value, bits in _CONST_BITS.items():
        assert _bits(value) == bits, (value, hex(_bits(value)), hex(bits))


def test_fdlibm_restatement_is_sin_and_cos():
    rng = np.random.default_rng(3)
    for x in list(rng.uniform(1e-6, 40.0, 4000)) + [math.pi / 2, math.pi, 1.5 * math.pi, 2 * math.pi, 0.1]:
        assert abs(fdlibm_sin(x) - math.sin(x)) <= 2 * math.ulp(1.0) * max(1.0, abs(x))
        assert abs(fdlibm_cos(x) - math.cos(x)) <= 2 * math.ulp(1.0) * max(1.0, abs(x))


def test_fdlibm_exp_log_restatement():
    rng = np.random.default_rng(4)
    for x in rng.uniform(-20.0, 20.0, 3000):
        assert abs(fdlibm_exp(x) / math.exp(x) - 1) < 4e-16
    for x in rng.uniform(1e-3, 1e6, 3000):
        assert abs(fdlibm_log(x) - math.log(x)) <= 4e-16 * max(1.0, abs(math.log(x)))
    assert fdlibm_log(1.0) == 0.0 and fdlibm_exp(0.0) == 1.0


def test_turn_table_is_fdlibms():
    """Every entry of the oracle's turn table equals the fdlibm algorithm restated here, and that algorithm stays
    within one ulp of the correctly rounded value (its documented bound).  It is not correctly rounded: about 3 %
    of the entries are the other neighbour, and there the table differs from one the host libm would build."""
    c, s = C.c_double(), C.c_double()
    off = 0
    for i in range(4096):
        a = 2.0 * math.pi * i / 4096.0  # rotor.zig:69, evaluated in f64 as Zig's comptime expression is
        # sincos64 at a whole number of table steps returns the table entry itself (remainder 0).
        L.orc_sincos64(i / 4096.0, C.byref(c), C.byref(s))
        want = (fdlibm_cos(a), fdlibm_sin(a)) if i else (1.0, 0.0)
        assert (c.value, s.value) == want, i
        if i:
            ds, dc = _dsin_dcos(a)
            for got, exact in ((want[0], dc), (want[1], ds)):
                cr = float(exact)
                assert abs(got - cr) <= math.ulp(cr), i
                off += got != cr
    assert 200 < off < 400, off  # 277 of the 8190 nonzero entries are not the correctly rounded neighbour


def test_engine_builds_the_same_turn_table():
    from pharos_amd import _lib
    got = np.zeros(2 * 4096)
    assert _lib.lib().pharos_synth_turn_table(got.ctypes.data_as(C.POINTER(C.c_double)), 4096) == 0
    c, s = C.c_double(), C.c_double()
    for i in range(4096):
        L.orc_sincos64(i / 4096.0, C.byref(c), C.byref(s))
        assert (got[2 * i], got[2 * i + 1]) == (c.value, s.value), i
    assert _lib.lib().pharos_synth_turn_table(got.ctypes.data_as(C.POINTER(C.c_double)), 4097) == 2


def test_fdlibm_known_values():
    """Cases the algorithm fixes exactly: the direct kernels below pi/4 for tiny arguments, and the reduction's
    sign and quadrant over each quarter turn."""
    assert fdlibm_sin(1e-9) == 1e-9 and fdlibm_cos(1e-9) == 1.0
    for k in range(1, 8):
        x = k * math.pi / 4 + 1e-3
        q = round((x - 1e-3) / (math.pi / 2))
        assert math.copysign(1, fdlibm_sin(x)) == math.copysign(1, math.sin(x)), k
        assert math.copysign(1, fdlibm_cos(x)) == math.copysign(1, math.cos(x)), k
        assert _rem_pio2(x)[0] in (q, q - 1, q + 1)
    for x in (0.3, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0):
        s_, c_ = fdlibm_sin(x), fdlibm_cos(x)
        assert abs(s_ * s_ + c_ * c_ - 1) < 4e-16


def _f32_margin(v):
    """Distance of the f64 value v from the nearest f32 rounding boundary, in f64 ulps of v."""
    f = float(np.float32(v))
    up = float(np.nextafter(np.float32(f), np.float32(np.inf)))
    dn = float(np.nextafter(np.float32(f), np.float32(-np.inf)))
    return min(abs((f + up) / 2 - v), abs(v - (f + dn) / 2)) / math.ulp(v)


def _pow_variants(y):
    return (go_pow(10.0, y), go_pow(10.0, y, math.exp, math.log), 10.0 ** y)


def test_db_linear_is_fixed_at_f32():
    """Db.linear over the whole allowed range (-60 to 60 dB, gain.zig:44-49) in 0.01 dB steps."""
    worst = math.inf
    for k in range(-6000, 6001):
        db = k / 100.0
        vs = _pow_variants(db / 20.0)
        f32s = {float(np.float32(v)) for v in vs}
        assert len(f32s) == 1, (db, vs)
        assert L.orc_db_linear(db) == f32s.pop(), db
        worst = min(worst, min(_f32_margin(v) for v in vs))
    assert worst > 8, worst  # every variant is at least this many f64 ulps away from an f32 boundary


@pytest.mark.parametrize("spe", [2046, 4092, 20460, 61440, 122880])
def test_amplitude_is_fixed_at_f32(spe):
    """amplitude over C/N0 from 0 to 70 dB-Hz in 0.005 dB steps at the path's sample rates."""
    fs = 1000.0 * spe
    worst = math.inf
    for k in range(0, 14001):
        cn0 = k * 0.005
        vs = [math.sqrt(2.0 * p / fs) for p in _pow_variants(cn0 / 10.0)]
        f32s = {float(np.float32(v)) for v in vs}
        assert len(f32s) == 1, (cn0, vs)
        assert L.orc_amplitude(cn0, spe) == f32s.pop(), cn0
        worst = min(worst, min(_f32_margin(v) for v in vs))
    assert worst > 2, worst


def test_pow_variants_match_the_correctly_rounded_power():
    rng = np.random.default_rng(5)
    for y in list(rng.uniform(-3.0, 7.0, 400)) + [0.25, 0.75, 2.5, 3.3]:
        exact = Decimal(10) ** Decimal(y)
        want = float(exact)
        for v in _pow_variants(y):
            assert abs(v - want) <= 4 * math.ulp(want), (y, v, want)
