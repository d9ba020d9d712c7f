"""The device AGC against the oracle's: a clipping run on the batched form (AGC on the device between windows) and
on the plain form (the library's host AGC step between windows), both against the oracle's windows driven by
orc_agc_account (gain.zig:19-35, applied per window as engine/run.zig:238-247 does)."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd.scenario import Channel, Scenario

pytestmark = pytest.mark.gpu


def clipping_scenario(epochs):
    # Eight strong L1 C/A channels plus noise at IQ4: about 1-2 % of components clip at gain 1, so the AGC
    # reduces, then raises again once the rate falls under 0.1 %.
    r = np.random.default_rng(21)
    chans = [Channel("L1CA", int(p), if_hz=int(r.integers(-500_000, 500_000)), range0_m=float(r.uniform(2.0e7, 2.5e7)),
                     doppler_hz=float(r.uniform(-4000, 4000)), cn0_dbhz=float(r.uniform(48, 52)),
                     nav_seed=int(r.integers(1, 2**62))) for p in r.choice(np.arange(1, 33), 8, replace=False)]
    return Scenario("agc_l1ca_8ch_4.092Msps_iq4", 4092, "IQ4", epochs, chans, noise=True, seed=77, gain_db=None)


def oracle_run(sc, first_epoch=0):
    tabs = O.Tables([(c.signal, c.prn) for c in sc.channels])
    agc = O.Agc(1.0, 0, 0)
    out, log, e = [], [], 0
    while e < sc.epochs:
        absolute = first_epoch + e
        n = min(100 - absolute % 100, sc.epochs - e)
        rows, masks = sc.window(absolute, n)
        gain = float(np.float32(agc.gain))
        b, clips = O.synth_window(tabs, rows, masks, sc.samples_per_epoch, sc.format_id, 0, sc.seed, absolute, gain,
                                  threads=8)
        out.append(b)
        rate = C.c_double(0.0)
        if O.lib().orc_agc_account(C.byref(agc), absolute + n, int(clips.sum()), n * sc.samples_per_epoch * 2,
                                   C.byref(rate)):
            log.append((absolute + n, agc.gain, rate.value))
        e += n
    return np.concatenate(out), agc.gain, log


def test_device_and_host_agc_follow_the_oracle():
    from pharos_amd.engine import Job, run_batch, plain_job
    sc = clipping_scenario(900)
    want_bytes, want_gain, want_log = oracle_run(sc)
    assert len(want_log) >= 2 and any(g < 1.0 for _, g, _ in want_log), want_log  # the trace moves the gain
    job = Job(sc)
    run_batch([job])
    assert job.adjustments() == want_log
    assert job.agc.gain == want_gain
    assert np.array_equal(job.result(), want_bytes)
    job2 = Job(sc)
    _, gain, log = plain_job(job2)
    assert log == want_log and gain == want_gain
    assert np.array_equal(job2.result(), want_bytes)
