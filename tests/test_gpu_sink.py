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


CHILD = r'''
import ctypes as C, json, os, resource, signal, sys
sys.path.insert(0, os.getcwd())
from pharos_amd import _lib, scenario as S
from pharos_amd.engine import Job
signal.signal(signal.SIGXFSZ, signal.SIG_IGN)            # a write past the limit fails with EFBIG instead
limit = int(sys.argv[2])
job = Job(S.c0(seed=33, epochs=700))
resource.setrlimit(resource.RLIMIT_FSIZE, (limit, limit))
n = C.c_uint64(0)
rec = job.record()
st = _lib.lib().pharos_synth_to_file(C.byref(rec), os.fsencode(sys.argv[1]), C.byref(n))
print(json.dumps({"status": st, "bytes_written": n.value, "error": _lib.last_error()}), flush=True)
'''


def test_sink_reports_the_bytes_written_before_a_failed_write(tmp_path):
    # The file may grow to 2.5 blocks of 100 epochs: the third block's write stops part-way (EFBIG) and later
    # blocks write nothing, so the bytes on file in order are exactly the limit.
    import json
    import os
    import subprocess
    import sys
    sc = S.c0(seed=33, epochs=700)
    block = 100 * sc.bytes_per_epoch
    limit = 2 * block + block // 2
    path = tmp_path / "limited.bin"
    print(f"case sink with a file size limit of {limit} bytes", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", CHILD, str(path), str(limit)], cwd=root, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert res["status"] == 4 and "writing" in res["error"], res
    assert res["bytes_written"] == limit, res
    (ref,) = run_batch([Job(sc)])
    data = np.fromfile(path, dtype=np.uint8)
    assert np.array_equal(data[:limit], ref.result()[:limit])
