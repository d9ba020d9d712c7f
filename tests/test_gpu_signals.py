"""GPU parity over every signal, shape, format and the awkward rates, against the CPU oracle (bit-exact)."""
import numpy as np
import pytest

from pharos_amd import _lib
from pharos_amd.scenario import Scenario, Channel
from tests.helpers import oracle_window, gpu_window, describe_mismatch

pytestmark = pytest.mark.gpu

SATS = {"L1CA": 32, "L1C": 32, "L2C": 32, "L5": 32, "B1C": 63, "B1I": 63, "B2I": 63, "B3I": 63, "B2a": 63, "B2b": 63,
        "E1": 36, "E5a": 36, "E5b": 36, "E6": 36, "G1": 24, "G2": 24}


def _channels(signal, n, seed, spe, lead=None):
    r = np.random.default_rng(seed)
    out = []
    for p in r.choice(np.arange(1, SATS[signal] + 1), size=n, replace=False):
        ch = Channel(signal, int(p), freq_k=int(r.integers(-7, 7)) if signal in ("G1", "G2") else 0,
                     if_hz=int(r.integers(-spe * 350, spe * 350)), range0_m=float(r.uniform(2.0e7, 2.6e7)),
                     doppler_hz=float(r.uniform(-5000, 5000)), doppler_rate_hz_s=float(r.uniform(-400, 400)),
                     cn0_dbhz=float(r.uniform(40, 56)), nav_seed=int(r.integers(1, 2**62)))
        if lead:
            ch.lead_component, ch.lead_entries = lead[0], float(r.uniform(0.05, lead[1]))
        out.append(ch)
    return out


def _check(sc, first=0, epochs=3, gain=1.0, rows=None, masks=None):
    print(f"case {sc.name} spe={sc.samples_per_epoch} fmt={sc.format} first={first} epochs={epochs}", flush=True)
    if rows is None:
        rows, masks = sc.window(first, epochs)
    ob, oc, osr, osi = oracle_window(sc, first, epochs, gain, rows=rows, masks=masks)
    gb, gc, gsr, gsi = gpu_window(sc, first, epochs, gain, rows=rows, masks=masks)
    assert np.array_equal(gsr, osr), describe_mismatch(gsr, osr, "I sums")
    assert np.array_equal(gsi, osi), describe_mismatch(gsi, osi, "Q sums")
    assert np.array_equal(gb, ob), describe_mismatch(gb, ob, "bytes")
    assert np.array_equal(gc, oc), describe_mismatch(gc, oc, "clips")


RATES = [4092, 16_368, 30_690, 61_440]
FORMATS = ["IQ2", "IQ4", "IQ8", "IQ16"]


@pytest.mark.parametrize("i,signal", list(enumerate(_lib.SIGNALS)))
def test_every_signal(i, signal):
    spe = RATES[i % 4]
    sc = Scenario(f"sig_{signal}", spe, FORMATS[(i // 4) % 4], 3, _channels(signal, 3, 100 + i, spe), noise=True,
                  seed=17 + i)
    _check(sc, first=1000 * i + 37, epochs=3, gain=0.8)


@pytest.mark.parametrize("signal,component,max_lead,spe", [("L5", 1, 2.5, 16_368), ("L5", 2, 0.9, 61_440),
                                                             ("B1C", 1, 3.5, 4092), ("L1C", 1, 1.8, 30_690),
                                                             ("B2a", 2, 1.2, 122_880), ("E5a", 1, 0.5, 4092)])
def test_lead_shapes(signal, component, max_lead, spe):
    sc = Scenario(f"lead_{signal}", spe, "IQ16", 2, _channels(signal, 3, 7, spe, lead=(component, max_lead)),
                  noise=False, seed=3)
    rows, masks = sc.window(500, 2)
    assert (rows["lead_bits"] != 0).all()
    _check(sc, first=500, epochs=2, rows=rows, masks=masks)


@pytest.mark.parametrize("spe", [4092, 16_368, 1000])
def test_tdm_l2c(spe):
    sc = Scenario("l2c_tdm", spe, "IQ8", 4, _channels("L2C", 4, 11, spe), noise=True, seed=5)
    _check(sc, first=1234 * 20 - 2, epochs=4)  # crosses a 20 ms code period


def test_mixed_everything_iq2_odd_slots():
    chans = []
    for k, sig in enumerate(["L1CA", "E1", "L5", "B1C", "G1", "L2C", "E6", "B3I"]):
        chans += _channels(sig, 2, 40 + k, 30_690)
    sc = Scenario("mixed_iq2", 30_690, "IQ2", 3, chans, noise=True, seed=77)
    _check(sc, first=3, epochs=3, gain=1.7)  # 15 345-byte epochs: slots start at odd addresses


def test_inactive_rows_and_zero_channels():
    sc = Scenario("inactive", 16_368, "IQ8", 3, _channels("L1CA", 5, 8, 16_368), noise=True, seed=9)
    rows, masks = sc.window(0, 3)
    rows["table"][1, 2] = _lib.INACTIVE
    rows["table"][:, 4] = _lib.INACTIVE
    _check(sc, 0, 3, rows=rows, masks=masks)
    empty = Scenario("noise_only", 4092, "IQ4", 2, [], noise=True, seed=12)
    _check(empty, first=99, epochs=2)


@pytest.mark.parametrize("spe,fmt", [(512, "IQ8"), (1000, "IQ2"), (2046, "IQ16"), (4096, "IQ4")])
def test_low_and_round_rates(spe, fmt):
    sc = Scenario(f"rate_{spe}", spe, fmt, 4, _channels("L1CA", 4, spe, spe) + _channels("G1", 2, 1, spe), noise=True,
                  seed=spe)
    _check(sc, first=17, epochs=4)


def test_noise_index_beyond_32_bits():
    sc = Scenario("late_epoch", 61_440, "IQ4", 2, _channels("E5a", 2, 4, 61_440), noise=True, seed=2**63 + 5)
    _check(sc, first=1_000_003, epochs=2)  # absolute sample index ~6.1e10


def test_many_channels_take_several_passes():
    """70 channels: passes of up to 32 channels (4 setup threads each) and a short last pass (8 threads each)."""
    chans = []
    for k, sig in enumerate(["L1CA", "E1", "L5", "B1I", "G1", "E5a", "B3I"]):
        chans += _channels(sig, 10, 90 + k, 61_440)
    sc = Scenario("many", 61_440, "IQ4", 2, chans, noise=True, seed=31)
    _check(sc, first=777, epochs=2, gain=0.6)


def test_highest_rate():
    sc = Scenario("rate_200000", 200_000, "IQ8", 2, _channels("L5", 3, 5, 200_000) + _channels("E1", 3, 6, 200_000),
                  noise=True, seed=8)
    _check(sc, first=5, epochs=2)
