"""Every launch shape the autotuner may choose gives the same bytes, clips and AGC trace."""
import numpy as np
import pytest

from pharos_amd import scenario as S, tune
from pharos_amd.engine import Job, run_batch

pytestmark = pytest.mark.gpu


def _run(scs, w, c, flags=0):
    jobs = run_batch([Job(s) for s in scs], workgroups_per_cu=w, channels_per_pass=c, flags=flags)
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


WIDE = 2  # PHAROS_SYNTH_TUNE_WIDE_TILES


@pytest.mark.parametrize("case", ["multi_gnss", "wideband", "uneven_above_4096", "uneven_with_l1ca"])
def test_wide_tiles_give_identical_results(case):
    """8192-sample work items of 256 threads give the bytes, clips and AGC trace of 4096-sample ones: the multi-GNSS
    run (7.5 wide tiles per epoch, so each epoch ends on a half-empty one), the wideband run, uneven jobs above 4096
    samples per epoch (30.69 Msps: the last rotor block of an epoch is partial), and a batch holding 4.092 Msps jobs,
    which has no wide form and runs narrow."""
    if case == "multi_gnss":
        scs = [S.c1(seed=4, epochs=160)]
    elif case == "wideband":
        scs = [S.c2(seed=4, epochs=120)]
    else:
        pool = S.c3(seed=11, count=24, epochs_cap=130)
        scs = [s for s in pool if s.samples_per_epoch > 4096][:5]
        if case == "uneven_with_l1ca":
            scs = scs[:3] + [s for s in pool if s.samples_per_epoch <= 4096][:2]
        assert any(s.samples_per_epoch == 30690 for s in pool)
    print("case wide tiles", case, [(s.samples_per_epoch, s.format, len(s.channels), s.epochs) for s in scs], flush=True)
    ref = _run(scs, 0, 0)
    for w, c in [(0, 0), (2, 32), (1, 16), (2, 8)]:
        got = _run(scs, w, c, WIDE)
        for (rb, rc, ra, rg), (gb, gc, ga, gg) in zip(ref, got):
            assert np.array_equal(rb, gb), (case, w, c)
            assert rc == gc and ra == ga and rg == gg, (case, w, c)


def test_tuner_picks_a_candidate(tmp_path):
    sc = [S.c1(seed=1, epochs=200)]
    res = tune.tune(sc, sample_epochs=100, repeats=1, cache_path=str(tmp_path / "t.json"))
    print("tuner", res["best"], {k: round(v, 3) for k, v in res["times_ms"].items()}, flush=True)
    assert tuple(res["best"]) in tune.CANDIDATES
    assert tune.lookup(sc, str(tmp_path / "t.json")) == tuple(res["best"])
