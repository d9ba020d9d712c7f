"""The autotuner's shape keys and cache (CPU)."""
from pharos_amd import scenario as S
from pharos_amd import tune


def test_shape_key_buckets_channels_and_jobs():
    c1 = S.c1(seed=1, epochs=10)
    assert tune.shape_key([c1]) == "61440/IQ4/40/1"
    c3 = S.c3(seed=1, count=5, epochs_cap=10)
    key = tune.shape_key(c3)
    assert key.endswith("/8") and key.count("/") == 3


def test_cache_store_and_lookup(tmp_path):
    path = str(tmp_path / "tuning.json")
    sc = [S.c0(seed=1, epochs=10)]
    assert tune.lookup(sc, path) == (0, 0, 0)
    tune.store(path, tune.shape_key(sc), (3, 12, 0), 1.5)
    assert tune.lookup(sc, path) == (3, 12, 0)
    tune.store(path, "other", (1, 4, tune.WIDE_TILES), 2.0)
    assert tune.lookup(sc, path) == (3, 12, 0)
    assert len(tune.CANDIDATES) == 24


def test_wide_candidates_only_for_long_epochs():
    """Wide tiles are timed only when some scenario's epoch is longer than a narrow tile: the library runs the jobs
    with shorter epochs narrow, in a second launch."""
    c0, c1 = S.c0(seed=1, epochs=10), S.c1(seed=1, epochs=10)
    assert len(tune.candidates([c1])) == 24 and any(f for _, _, f in tune.candidates([c1]))
    assert len(tune.candidates([c0])) == 16 and not any(f for _, _, f in tune.candidates([c0]))
    assert len(tune.candidates([c0, c1])) == 24
    assert tune.label(2, 32, tune.WIDE_TILES) == "2x32w" and tune.label(4, 24, 0) == "4x24"
