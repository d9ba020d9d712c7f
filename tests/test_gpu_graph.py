"""The graph form (a job's window loop captured as a HIP graph, AGC steps on the device) against the plain form.

A graph is replayed several times, with eager work and synchronisations between the replays, and every replay
must equal the plain form's bytes, clips and AGC trace."""
import numpy as np
import pytest

from pharos_amd import scenario as S
from pharos_amd.engine import Job, Graph, run_plain, Stream, DeviceBuffer, Engine
from tests.test_gpu_batch import _strong_iq4

pytestmark = pytest.mark.gpu


def _reset(job):
    job.agc.gain, job.agc.clipped, job.agc.components, job.agc.adjustments = 1.0, 0, 0, 0
    job.clipped.value = 0


def test_graph_replays_equal_plain_adaptive_gain():
    sc = _strong_iq4(5, epochs=530)
    print("case graph, strong IQ4 adaptive gain, 530 epochs", flush=True)
    pb, pc, pgain, padj = run_plain(sc)
    assert len(padj) >= 3
    job = Job(sc)
    g = Graph(job)
    assert g.windows == 6
    st = Stream()
    for replay in range(3):
        _reset(job)
        g.run(st)
        assert np.array_equal(job.result(), pb), replay
        assert job.clipped.value == int(pc.sum()), replay
        assert job.adjustments() == padj and job.agc.gain == pgain, replay
        # eager work between replays on the same device
        eng = Engine(S.c0(seed=1, epochs=3))
        rows, masks = eng.scenario.window(0, 3)
        d_rows, d_masks = DeviceBuffer.from_array(rows), DeviceBuffer.from_array(masks)
        out = DeviceBuffer(3 * eng.bytes_per_epoch)
        eng.window(0, 3, d_rows.ptr, d_masks.ptr, len(masks), 1.0, out.ptr)
        eng.close()
    g.close()


def test_graph_constant_gain_from_a_mid_run_epoch():
    sc = S.c0(seed=4, epochs=400)
    print("case graph, c0 constant gain from epoch 57", flush=True)
    pb, pc, _, _ = run_plain(sc, first_epoch=57, epochs=343)
    job = Job(sc, first_epoch=57, epochs=343)
    g = Graph(job)
    for replay in range(2):
        g.run()
        assert np.array_equal(job.result(), pb), replay
        assert job.clipped.value == int(pc.sum())
    g.close()


def test_graph_with_the_row_pre_pass_from_a_mid_run_epoch():
    """An engine with several tiles per epoch: the graph's first node computes the rows' rotor constants, and the
    window launches read them at their own row offsets."""
    sc = _strong_iq4(7, epochs=420)
    print("case graph, strong IQ4 (4 tiles per epoch) from epoch 57", flush=True)
    pb, pc, pgain, padj = run_plain(sc, first_epoch=57, epochs=363)
    rows, masks = sc.window(57, 363)
    job = Job(sc, first_epoch=57, epochs=363, rows=rows, masks=masks)
    g = Graph(job)
    for replay in range(2):
        _reset(job)
        g.run()
        assert np.array_equal(job.result(), pb), replay
        assert job.clipped.value == int(pc.sum()), replay
        assert job.adjustments() == padj and job.agc.gain == pgain, replay
    g.close()
