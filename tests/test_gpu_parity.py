"""GPU parity of the HIP engine against the CPU oracle on the same seeded rows and masks.

The bar: quantised bytes, clip counts and pre-quantisation sums bit-identical to the oracle (a stricter bar than
the 1e-5 relative / +-1 LSB the engine is allowed), because every per-sample quantity is computed by the same IEEE
operations in the same order.
"""
import numpy as np
import pytest

from pharos_amd import scenario as S
from tests.helpers import oracle_window, gpu_window, describe_mismatch

pytestmark = pytest.mark.gpu


def _check(sc, first_epoch, epochs, gain=1.0):
    print(f"case {sc.name} first={first_epoch} epochs={epochs} spe={sc.samples_per_epoch} fmt={sc.format}", flush=True)
    ob, oc, osr, osi = oracle_window(sc, first_epoch, epochs, gain)
    gb, gc, gsr, gsi = gpu_window(sc, first_epoch, epochs, gain)
    assert np.array_equal(gsr, osr), describe_mismatch(gsr, osr, "I sums")
    assert np.array_equal(gsi, osi), describe_mismatch(gsi, osi, "Q sums")
    assert np.array_equal(gb, ob), describe_mismatch(gb, ob, "bytes")
    assert np.array_equal(gc, oc), describe_mismatch(gc, oc, "clips")


def test_c0_l1ca_iq8_noise_free():
    _check(S.c0(seed=3, epochs=30), 0, 30)


def test_c0_mid_run_window():
    sc = S.c0(seed=5, epochs=3000)
    _check(sc, 2480, 40)


def test_c1_multi_gnss_iq4():
    _check(S.c1(seed=2, epochs=4), 0, 4)


def test_c2_wideband_iq16():
    _check(S.c2(seed=2, epochs=3), 0, 3)
