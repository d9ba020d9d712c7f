"""The serving loop: requests that arrive while others run share batches, and each one's bytes, clip count and
AGC trace equal the same run alone (the plain form with the host AGC)."""
import numpy as np
import pytest

from pharos_amd import serve
from pharos_amd.engine import run_plain
from tests.test_gpu_batch import _strong_iq4

pytestmark = pytest.mark.gpu


def test_requests_in_the_mix_equal_alone(monkeypatch):
    # Short slices, so each request crosses several batches and carries its AGC state between them.
    monkeypatch.setattr(serve, "SLICE_MIN", 300)
    monkeypatch.setattr(serve, "SLICE_WORK", 1.0)
    reqs = serve.make_requests(rate_hz=400.0, duration_s=0.02, seed=3, epochs_cap=1400)
    reqs.append(serve.Request(_strong_iq4(8, spe=16_368, epochs=2300), 0.004, len(reqs)))
    print("case serve mix", [(r.scenario.samples_per_epoch, r.scenario.format, len(r.scenario.channels),
                              r.scenario.epochs, round(r.arrival, 4)) for r in reqs], flush=True)
    assert len(reqs) >= 4
    server = serve.Server(threads=4)
    outs = {}
    done = server.serve(reqs, on_complete=lambda r: outs.__setitem__(r.id, r.out.to_numpy(r.out_bytes)), keep_outputs=True)
    assert len(done) == len(reqs)
    assert max(server.batch_sizes) >= 2, "requests must overlap on the device"
    for r in done:
        pb, pc, pgain, padj = run_plain(r.scenario, derive=True)  # the loop derives rows on the device
        assert np.array_equal(outs[r.id], pb), r.scenario.name
        assert r.clipped_total == int(pc.sum()), r.scenario.name
        assert r.adjustments() == padj and r.agc.gain == pgain, r.scenario.name
        assert r.completed >= r.arrival_abs
    strong = [r for r in done if r.scenario.name == "strong_l1ca_iq4"][0]
    assert len(strong.adjustments()) >= 3
    assert server.arena.used == sum(r.out.nbytes for r in done)  # only the kept outputs remain
    server.close()


def test_measure_reports_latency_percentiles():
    m = serve.measure(duration_s=0.05, rate_hz=100.0, seed=5, epochs_cap=600)
    print("serve measure", m, flush=True)
    assert m["requests"] >= 1 and m["p99_ms"] >= m["p50_ms"] > 0
