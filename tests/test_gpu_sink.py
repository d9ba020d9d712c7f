"""The file sink (pharos_synth_to_file, pharos-run's output/sink.zig:18-47): the file holds exactly the bytes of the
batched form, and the AGC trace and clip total equal it, for a constant gain and for the adaptive gain from a
mid-run epoch (whose first block ends on the 100-epoch grid)."""
import numpy as np
import pytest

from pharos_amd import scenario as S
from pharos_amd.engine import Job, run_batch
from tests.test_gpu_batch import _strong_iq4

pytestmark = pytest.mark.gpu


def test_sink_file_equals_batched_output_constant_gain(tmp_path):
    sc = S.c0(seed=31, epochs=250)
    print("case sink c0 constant gain, 250 epochs", flush=True)
    path = tmp_path / "c0.bin"
    job = Job(sc)
    n = job.to_file(path)
    (ref,) = run_batch([Job(sc)])
    data = np.fromfile(path, dtype=np.uint8)
    assert n == len(data) == sc.epochs * sc.bytes_per_epoch
    assert np.array_equal(data, ref.result())
    assert job.clipped.value == ref.clipped.value


def test_sink_file_equals_batched_output_adaptive_gain_mid_run(tmp_path):
    sc = _strong_iq4(13, epochs=520)
    print("case sink strong IQ4 adaptive gain from epoch 137", flush=True)
    rows, masks = sc.window(137, 383)
    path = tmp_path / "iq4.bin"
    job = Job(sc, first_epoch=137, epochs=383, rows=rows, masks=masks)
    n = job.to_file(path)
    (ref,) = run_batch([Job(sc, first_epoch=137, epochs=383, rows=rows, masks=masks)])
    data = np.fromfile(path, dtype=np.uint8)
    assert n == len(data) == 383 * sc.bytes_per_epoch
    assert np.array_equal(data, ref.result())
    assert len(ref.adjustments()) >= 2, "the scenario must clip enough to move the gain"
    assert job.adjustments() == ref.adjustments()
    assert job.agc.gain == ref.agc.gain and job.agc.clipped == ref.agc.clipped
    assert job.clipped.value == ref.clipped.value


def test_sink_reports_an_unwritable_path(tmp_path):
    sc = S.c0(seed=32, epochs=10)
    print("case sink into a missing directory", flush=True)
    job = Job(sc)
    with pytest.raises(RuntimeError, match="cannot create"):
        job.to_file(tmp_path / "missing" / "out.bin")
