"""The batched kernel's bounded gain wait: with the first window's gain held back by the host (a test flag), the
tiles that wait for it give up after the wait limit, or when the host aborts the batch; every workgroup still exits,
the batch reports status 4, and the next batch on the device runs normally."""
import numpy as np
import pytest

from pharos_amd import _lib, scenario as S
from pharos_amd.engine import Batch, Job, run_batch, plain_job

pytestmark = pytest.mark.gpu

HOLD = 1  # PHAROS_SYNTH_TUNE_HOLD_FIRST_GAIN


def adaptive_c0(epochs):
    sc = S.c0(seed=9, epochs=epochs)
    sc.gain_db = None
    sc.noise = True
    return sc


def test_gain_wait_gives_up_at_its_limit():
    print("wait bound: held first gain, wait limit 2000 polls", flush=True)
    job = Job(adaptive_c0(300))
    with pytest.raises(_lib.PharosError) as e:
        run_batch([job], wait_limit=2000, flags=HOLD)
    assert e.value.status == 4 and "wait limit" in str(e.value)


def test_host_abort_ends_the_wait():
    print("wait bound: held first gain, host abort", flush=True)
    job = Job(adaptive_c0(300))
    b = Batch([job], wait_limit=1 << 21, flags=HOLD)
    b.abort()
    with pytest.raises(_lib.PharosError) as e:
        b.wait()
    assert e.value.status == 4 and "aborted" in str(e.value)


def test_the_device_runs_normally_after_a_wait_gave_up():
    sc = adaptive_c0(300)
    job = Job(sc)
    with pytest.raises(_lib.PharosError):
        run_batch([job], wait_limit=100, flags=HOLD)
    again = Job(sc)
    run_batch([again])
    ref = Job(sc)
    _, gain, log = plain_job(ref)
    assert np.array_equal(again.result(), ref.result())
    assert again.adjustments() == log and again.agc.gain == gain
