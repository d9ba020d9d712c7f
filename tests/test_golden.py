"""The oracle and the scenario stage against the committed golden fixtures (CPU)."""
import numpy as np
import pytest

from tests.golden_cases import manifest, load, sha
from tests.helpers import oracle_window
from pharos_amd import scenario as S

CASES = sorted(manifest())


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    sc, meta, rows, masks, arrays = load(name)
    out, clips, sre, sim = oracle_window(sc, meta["first_epoch"], meta["epochs"], meta["gain"], rows=rows, masks=masks)
    assert sha(out) == meta["sha256_bytes"]
    assert sha(sre) == meta["sha256_sums_re"] and sha(sim) == meta["sha256_sums_im"]
    assert [int(c) for c in clips] == meta["clips"]
    if meta["full_outputs"]:
        assert np.array_equal(out, arrays["bytes"])
        assert np.array_equal(sre, arrays["sums_re"]) and np.array_equal(sim, arrays["sums_im"])


def test_fixture_digests_match_stored_outputs():
    for name in CASES:
        sc, meta, rows, masks, arrays = load(name)
        if meta["full_outputs"]:
            assert sha(arrays["bytes"]) == meta["sha256_bytes"], name
            assert sha(arrays["sums_re"]) == meta["sha256_sums_re"], name


@pytest.mark.parametrize("config,name", [(S.c0, "c0_iq8_noisefree"), (S.c1, "c1_iq4"), (S.c2, "c2_iq16")])
def test_scenario_stage_reproduces_fixture_rows(config, name):
    """The host stage (derive + masks) is deterministic: the same seed gives the fixture's rows and masks."""
    _, meta, rows, masks, _ = load(name)
    sc = config(seed=3)
    got_rows, got_masks = sc.window(meta["first_epoch"], meta["epochs"])
    assert np.array_equal(got_rows.view(np.uint8), rows.view(np.uint8))
    assert np.array_equal(got_masks, masks)
