"""The HIP engine against the committed golden fixtures: bytes, clips and pre-quantisation sums, bit-exact."""
import numpy as np
import pytest

from tests.golden_cases import manifest, load, sha
from tests.helpers import gpu_window, describe_mismatch

pytestmark = pytest.mark.gpu
CASES = sorted(manifest())


@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_fixture(name):
    sc, meta, rows, masks, arrays = load(name)
    print(f"golden {name}: {len(sc.channels)} ch, spe {sc.samples_per_epoch}, {sc.format}", flush=True)
    out, clips, sre, sim = gpu_window(sc, meta["first_epoch"], meta["epochs"], meta["gain"], rows=rows, masks=masks)
    if meta["full_outputs"]:
        assert np.array_equal(out, arrays["bytes"]), describe_mismatch(out, arrays["bytes"], "bytes")
        assert np.array_equal(sre, arrays["sums_re"]), describe_mismatch(sre, arrays["sums_re"], "I sums")
        assert np.array_equal(sim, arrays["sums_im"]), describe_mismatch(sim, arrays["sums_im"], "Q sums")
    assert sha(out) == meta["sha256_bytes"]
    assert sha(sre) == meta["sha256_sums_re"] and sha(sim) == meta["sha256_sums_im"]
    assert [int(c) for c in clips] == meta["clips"]
