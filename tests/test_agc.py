"""The AGC step (Agc.account, src/products/run/synthesis/gain.zig:19-35): the library's step (the same agc_step
the batched and graph forms run on the device, exported for the host as pharos_synth_agc_account) against the
oracle's orc_agc_account, on random traces and on the rule's edge cases; and the oracle's rule in closed form.
"""
import ctypes as C

import numpy as np

from oracle import binding as O
from pharos_amd import _lib


class LibAgc:
    def __init__(self, gain=1.0):
        self.s = _lib.SynthAgc(gain, 0, 0, 0, 0)
        self.log = (_lib.SynthAdjustment * 64)()

    def account(self, clipped, components, epoch_after):
        assert _lib.lib().pharos_synth_agc_account(C.byref(self.s), clipped, components, epoch_after, self.log,
                                                   len(self.log)) == 0

    def adjustments(self):
        return [(a.epoch, a.gain, a.clip_rate, a.reduced) for a in self.log[:min(self.s.adjustments, 64)]]


class OracleAgc:
    def __init__(self, gain=1.0):
        self.s = O.Agc(gain, 0, 0)
        self.log = []

    def account(self, clipped, components, epoch_after):
        rate = C.c_double(0.0)
        before = self.s.gain
        if O.lib().orc_agc_account(C.byref(self.s), epoch_after, clipped, components, C.byref(rate)):
            self.log.append((epoch_after, self.s.gain, rate.value, int(self.s.gain < before)))


def run_both(trace, gain=1.0):
    a, b = LibAgc(gain), OracleAgc(gain)
    for clipped, components, epoch_after in trace:
        a.account(clipped, components, epoch_after)
        b.account(clipped, components, epoch_after)
        assert a.s.gain == b.s.gain and a.s.clipped == b.s.clipped and a.s.components == b.s.components
    assert a.adjustments() == b.log
    return a


def test_random_traces_agree_with_the_oracle():
    rng = np.random.default_rng(11)
    for _ in range(50):
        epoch, trace = int(rng.integers(0, 1000)), []
        spe = int(rng.choice([4092, 16_368, 61_440]))
        for _ in range(40):
            n = int(min(100 - epoch % 100, rng.integers(1, 101)))   # windows end on the 100-epoch grid
            comps = n * spe * 2
            rate = float(rng.choice([0.0, 0.0005, 0.001, 0.005, 0.01, 0.02, 0.3]))
            trace.append((int(round(rate * comps)), comps, epoch + n))
            epoch += n
        run_both(trace, gain=float(rng.choice([1.0, 0.5, 0.999])))


def test_edge_cases():
    comps = 100 * 4092 * 2
    # A rate of exactly 1 % does not reduce (the rule is rate > 0.01); one component more does.
    a = run_both([(comps // 100, comps, 100)])
    assert a.s.gain == 1.0 and a.s.adjustments == 0
    a = run_both([(comps // 100 + 1, comps, 100)])
    assert a.s.gain == 0.95 and a.adjustments()[0][3] == 1
    # The rise is capped at 1.0, and a gain of 1.0 does not rise.
    a = run_both([(0, comps, 100)], gain=0.99)
    assert a.s.gain == 1.0 and a.adjustments() == [(100, 1.0, 0.0, 0)]
    a = run_both([(0, comps, 100)], gain=1.0)
    assert a.s.adjustments == 0
    # Counts accumulate over windows until an evaluation epoch, then restart.
    a = run_both([(comps, comps, 50), (0, comps, 100), (0, comps, 137)])
    assert a.s.gain == 0.95 and a.s.clipped == 0 and a.s.components == comps


def test_oracle_rule_in_closed_form():
    # Three evaluations clipping 2 % reduce by 0.95 each; then clean windows raise by 1.02 until capped at 1.
    comps = 100 * 61_440 * 2
    b = OracleAgc()
    for w in range(3):
        b.account(comps // 50, comps, 100 * (w + 1))
    assert b.s.gain == 1.0 * 0.95 * 0.95 * 0.95
    g = b.s.gain
    for w in range(3, 11):           # 0.95^3 * 1.02^7 < 1 <= 0.95^3 * 1.02^8
        b.account(0, comps, 100 * (w + 1))
        g = min(g * 1.02, 1.0)
        assert b.s.gain == g
    assert b.s.gain == 1.0
    assert [e for e, *_ in b.log] == [100 * (w + 1) for w in range(11)]
