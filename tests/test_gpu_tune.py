"""Every launch shape the autotuner may choose gives the same bytes, clips and AGC trace."""
import numpy as np
import pytest

from pharos_amd import scenario as S, tune
from pharos_amd.engine import Job, run_batch

pytestmark = pytest.mark.gpu


def _run(scs, w, c):
    jobs = run_batch([Job(s) for s in scs], workgroups_per_cu=w, channels_per_pass=c)
    return [(j.result(), j.clipped.value, j.adjustments(), j.agc.gain) for j in jobs]


def test_launch_shapes_give_identical_results():
    scs = S.c3(seed=7, count=4, epochs_cap=250) + [S.c1(seed=2, epochs=150)]
    print("case tuning shapes", [(s.samples_per_epoch, s.format, len(s.channels)) for s in scs], flush=True)
    ref = _run(scs, 0, 0)
    for w, c in [(1, 4), (2, 16), (4, 8), (3, 12), (4, 32), (2, 24)]:
        got = _run(scs, w, c)
        for (rb, rc, ra, rg), (gb, gc, ga, gg) in zip(ref, got):
            assert np.array_equal(rb, gb), (w, c)
            assert rc == gc and ra == ga and rg == gg, (w, c)


def test_tuner_picks_a_candidate(tmp_path):
    sc = [S.c1(seed=1, epochs=200)]
    res = tune.tune(sc, sample_epochs=100, repeats=1, cache_path=str(tmp_path / "t.json"))
    print("tuner", res["best"], {k: round(v, 3) for k, v in res["times_ms"].items()}, flush=True)
    assert tuple(res["best"]) in tune.CANDIDATES
    assert tune.lookup(sc, str(tmp_path / "t.json")) == tuple(res["best"])
