"""The benchmark configurations at their full sizes, checked through properties that do not depend on size:

- the batched run (one persistent launch, device AGC) gives the same bytes, clip total and AGC trace as the plain
  form (per-window launches, host AGC, engine/run.zig:225-256);
- epochs sampled across the run, up to the last, equal the oracle's bytes at the gain the AGC trace gives for
  that epoch.  Late epochs exercise large absolute sample indices (noise counters), code periods and carrier
  phases that the short parity windows do not reach.
"""
import numpy as np
import pytest

from pharos_amd import scenario as S
from pharos_amd.engine import Job, run_batch, plain_job
from tests.helpers import oracle_window, describe_mismatch

pytestmark = pytest.mark.gpu


def _gain_at(adjustments, epoch):
    """The gain a window starting at or containing `epoch` runs at: the last adjustment made at or before it
    (an adjustment's epoch is the end of the window it evaluated), as the f32 the kernel takes."""
    g = 1.0
    for e, gain, _ in adjustments:
        if e <= epoch:
            g = gain
    return float(np.float32(g))


def _assert_outputs_equal(a, b, chunk=1 << 28):
    assert a.nbytes == b.nbytes
    for off in range(0, a.nbytes, chunk):
        n = min(chunk, a.nbytes - off)
        x, y = a.to_numpy(n, off), b.to_numpy(n, off)
        assert np.array_equal(x, y), describe_mismatch(x, y, f"bytes at offset {off}")


@pytest.mark.parametrize("name,make,sampled", [
    ("c1", S.c1, (0, 1, 777, 29_999, 30_000, 59_999)),
    ("c2", S.c2, (0, 12_345, 19_999)),
    ("c0", S.c0, (0, 4_321, 9_999)),
])
def test_full_run(name, make, sampled):
    sc = make(seed=1)
    print(f"case {name} full run: {len(sc.channels)} channels, {sc.epochs} epochs, {sc.format}", flush=True)
    rows, masks = sc.window(0, sc.epochs)
    job = Job(sc, rows=rows, masks=masks)
    run_batch([job])
    ref = Job(sc, rows=rows, masks=masks)
    total, gain, adjustments = plain_job(ref)
    assert job.clipped.value == total
    if sc.gain_db is None:
        assert job.adjustments() == adjustments
        assert job.agc.gain == gain
    _assert_outputs_equal(job.out, ref.out)
    bpe = job.engine.bytes_per_epoch
    for e in sampled:
        g = job.constant_gain if job.constant_gain is not None else _gain_at(adjustments, e)
        print(f"  oracle at epoch {e}, gain {g}", flush=True)
        ob, _, _, _ = oracle_window(sc, e, 1, g)
        gb = job.out.to_numpy(bpe, e * bpe)
        assert np.array_equal(gb, ob), describe_mismatch(gb, ob, f"epoch {e}")
