"""The batched form (persistent kernel, device AGC) against the plain form (per-window launches, host AGC).

Each job's bytes, clip count and AGC adjustments must be the same alone and in any mix (the reference's
determinism contract, docs/design/engine.md:235-243), and the device AGC must follow gain.zig:19-35 exactly.
"""
import dataclasses

import numpy as np
import pytest

from pharos_amd import scenario as S
from pharos_amd.scenario import Scenario, Channel
from pharos_amd.engine import Job, run_batch, run_plain

pytestmark = pytest.mark.gpu

WIDE_TILES = 2  # PHAROS_SYNTH_TUNE_WIDE_TILES


def _strong_iq4(seed, spe=16_368, epochs=700, first=0):
    r = np.random.default_rng(seed)
    chans = [Channel("L1CA", int(p), if_hz=int(r.integers(-3_000_000, 3_000_000)), doppler_hz=float(r.uniform(-4000, 4000)),
                     cn0_dbhz=float(r.uniform(58, 62)), nav_seed=int(r.integers(1, 2**62)))
             for p in r.choice(np.arange(1, 33), size=20, replace=False)]
    return Scenario("strong_l1ca_iq4", spe, "IQ4", epochs, chans, noise=True, seed=seed, gain_db=None)


def test_batch_equals_plain_constant_gain():
    sc = S.c0(seed=11, epochs=250)
    print("case c0 constant gain, 250 epochs", flush=True)
    pb, pc, _, _ = run_plain(sc)
    (job,) = run_batch([Job(sc)])
    assert np.array_equal(job.result(), pb)
    assert job.clipped.value == int(pc.sum())


def test_device_agc_follows_the_host_agc():
    sc = _strong_iq4(5)
    print("case strong IQ4 adaptive gain, 700 epochs", flush=True)
    pb, pc, pgain, padj = run_plain(sc)
    assert len(padj) >= 3, "the scenario must clip enough to move the gain"
    (job,) = run_batch([Job(sc)])
    assert job.adjustments() == padj
    assert job.agc.gain == pgain
    assert np.array_equal(job.result(), pb)
    assert job.clipped.value == int(pc.sum())


def test_uneven_batch_each_job_as_alone():
    jobs_sc = S.c3(seed=2, count=6, epochs_cap=230) + [_strong_iq4(9, spe=30_690, epochs=310)]
    print("case 7 uneven jobs", [(s.samples_per_epoch, s.format, len(s.channels), s.epochs) for s in jobs_sc], flush=True)
    jobs = run_batch([Job(s) for s in jobs_sc])
    for s, j in zip(jobs_sc, jobs):
        pb, pc, pgain, padj = run_plain(s)
        assert np.array_equal(j.result(), pb), s.name
        assert j.clipped.value == int(pc.sum()), s.name
        assert j.adjustments() == padj and j.agc.gain == pgain, s.name


def test_constant_gain_batch_each_job_as_alone():
    """With no gain waits each workgroup starts on the item of its own index and takes the rest from the queue: a
    batch of many more items than workgroups, in one narrow launch and, with wide tiles asked for, split into a
    wide and a narrow launch."""
    jobs_sc = [dataclasses.replace(s, gain_db=-2.0) for s in S.c3(seed=4, count=5, epochs_cap=260)]
    jobs_sc.append(S.c0(seed=12, epochs=300))
    assert any(s.samples_per_epoch > 4096 for s in jobs_sc) and any(s.samples_per_epoch <= 4096 for s in jobs_sc)
    plain = [run_plain(s) for s in jobs_sc]
    for flags in (0, WIDE_TILES):
        print("case constant-gain batch, flags", flags,
              [(s.samples_per_epoch, s.format, len(s.channels), s.epochs) for s in jobs_sc], flush=True)
        jobs = run_batch([Job(s) for s in jobs_sc], flags=flags)
        for s, j, (pb, pc, _, _) in zip(jobs_sc, jobs, plain):
            assert np.array_equal(j.result(), pb), (s.name, flags)
            assert j.clipped.value == int(pc.sum()), (s.name, flags)


def test_batch_from_a_mid_run_epoch():
    sc = _strong_iq4(3, epochs=420)
    print("case adaptive gain from epoch 137", flush=True)
    rows, masks = sc.window(137, 283)
    pb, pc, pgain, padj = run_plain(sc, first_epoch=137, epochs=283)
    (job,) = run_batch([Job(sc, first_epoch=137, epochs=283, rows=rows, masks=masks)])
    assert np.array_equal(job.result(), pb)
    assert job.adjustments() == padj and job.agc.gain == pgain


def test_every_format_in_one_launch_each_as_alone():
    """A batch mixing all four quantised formats runs as one persistent launch that picks each job's quantiser per
    tile; every job must still equal its plain run."""
    jobs_sc = []
    for k, fmt in enumerate(["IQ2", "IQ4", "IQ8", "IQ16"]):
        sc = _strong_iq4(20 + k, spe=16_368, epochs=260 + 30 * k)
        sc.format, sc.name = fmt, f"strong_l1ca_{fmt.lower()}"
        jobs_sc.append(sc)
    jobs_sc.append(S.c0(seed=12, epochs=150))
    print("case every format", [(s.format, s.epochs) for s in jobs_sc], flush=True)
    jobs = run_batch([Job(s) for s in jobs_sc])
    for s, j in zip(jobs_sc, jobs):
        pb, pc, pgain, padj = run_plain(s)
        assert np.array_equal(j.result(), pb), s.name
        assert j.clipped.value == int(pc.sum()), s.name
        if s.gain_db is None:
            assert j.adjustments() == padj and j.agc.gain == pgain, s.name


def _batch_vs_oracle(sc, rows, masks, first=0):
    from tests.helpers import oracle_window, describe_mismatch
    epochs = rows.shape[0]
    job = Job(sc, first_epoch=first, epochs=epochs, rows=rows, masks=masks)
    (job,) = run_batch([job])
    ob, oc, _, _ = oracle_window(sc, first, epochs, 1.0, rows=rows, masks=masks)
    got = job.result()
    assert np.array_equal(got, ob), describe_mismatch(got, ob, "bytes")
    assert job.clipped.value == int(oc.sum())


def test_batch_pre_pass_with_inactive_rows():
    """Several tiles per epoch (the row pre-pass) with inactive rows, against the oracle."""
    from tests.test_gpu_signals import _channels
    from pharos_amd import _lib
    sc = Scenario("inactive_batch", 16_368, "IQ8", 3, _channels("L1CA", 5, 8, 16_368), noise=True, seed=19,
                  gain_db=0.0)
    print("case batch, 5 L1CA channels at 16.368 Msps with inactive rows", flush=True)
    rows, masks = sc.window(0, 3)
    rows["table"][1, 2] = _lib.INACTIVE
    rows["table"][:, 4] = _lib.INACTIVE
    _batch_vs_oracle(sc, rows, masks)


def test_batch_pre_pass_over_several_passes():
    """70 channels at 61.44 Msps: several LDS passes, each reading its channels' pre-pass constants."""
    from tests.test_gpu_signals import _channels
    chans = []
    for k, sig in enumerate(["L1CA", "E1", "L5", "B1I", "G1", "E5a", "B3I"]):
        chans += _channels(sig, 10, 90 + k, 61_440)
    sc = Scenario("many_batch", 61_440, "IQ4", 2, chans, noise=True, seed=33, gain_db=0.0)
    print("case batch, 70 channels at 61.44 Msps from epoch 777", flush=True)
    rows, masks = sc.window(777, 2)
    _batch_vs_oracle(sc, rows, masks, first=777)
