"""The device derive (pharos_synth_derive): rows evaluated on the GPU from per-block segments equal the oracle's
segment derive (parameters.derive on the cubics' endpoints, oracle/synth.c) byte for byte, and a run synthesised
from them equals the oracle's synthesis of the oracle's rows."""
import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib, scenario as S
from pharos_amd.engine import Job, derive_rows, run_batch
from pharos_amd.scenario import Channel, Scenario
from tests.helpers import oracle_window

pytestmark = pytest.mark.gpu


def lead_scenario(epochs):
    r = np.random.default_rng(8)
    chans = [Channel(sig, int(p), if_hz=int(r.integers(-2_000_000, 2_000_000)), doppler_hz=float(r.uniform(-3000, 3000)),
                     doppler_rate_hz_s=float(r.uniform(-1, 1)), cn0_dbhz=45.0, lead_component=1 + (p % 2),
                     lead_entries=float(r.uniform(0.5, 20.0)), nav_seed=int(r.integers(1, 2**62)))
             for sig, p in (("L5", 3), ("E5a", 11), ("B2a", 20), ("E1", 4), ("L1C", 7))]
    return Scenario("lead_mix", 30_690, "IQ8", epochs, chans, noise=True, seed=3, gain_db=0.0)


CASES = [("c1", lambda: S.c1(seed=6, epochs=1500), 0.0, 100, 300, 1000),
         ("c2", lambda: S.c2(seed=6, epochs=900), 0.37, 100, 150, 600),
         ("c0", lambda: S.c0(seed=6, epochs=900), 0.999, 37, 0, 333),
         ("lead", lambda: lead_scenario(700), 0.5, 100, 50, 400)]


@pytest.mark.parametrize("name,make,sub,block,first,epochs", CASES)
def test_device_rows_equal_the_oracle(name, make, sub, block, first, epochs):
    sc = make()
    sc.start_gps_sub_ms = sub
    print(f"derive: {name}, {epochs} epochs from {first}, blocks of {block}", flush=True)
    segs, masks = sc.segments(first, epochs, block)
    d_rows, d_masks, _ = derive_rows(sc, first, epochs, block, segments=(segs, masks))
    got = d_rows.to_numpy(epochs * len(sc.channels) * 64).view(_lib.ROW_DTYPE).reshape(epochs, -1)
    want = O.derive_segments(segs, sc.start_gps_ms, sub, first, epochs, block, sc.samples_per_epoch)
    assert (want["table"] != _lib.INACTIVE).all()
    assert got.tobytes() == want.tobytes()


def test_run_from_device_rows_equals_the_oracle():
    sc = lead_scenario(300)
    sc.start_gps_sub_ms = 0.25
    job = Job(sc, derive=True)
    (job,) = run_batch([job])
    segs, masks = sc.segments(0, sc.epochs, 100)
    rows = O.derive_segments(segs, sc.start_gps_ms, 0.25, 0, sc.epochs, 100, sc.samples_per_epoch)
    for e in (0, 149, 299):
        want, _, _, _ = oracle_window(sc, e, 1, gain=1.0, rows=rows[e:e + 1], masks=masks)
        got = job.out.to_numpy(sc.bytes_per_epoch, e * sc.bytes_per_epoch)
        assert np.array_equal(got, want), e


def level_steps(nch, spe):
    """Power-level steps as the run API emits them: amplitudes of C/N0 values at mid-block and block-edge epochs,
    a return to the segment's amplitude (NaN), a step before the derived range and a channel with none."""
    amp = lambda cn0: np.float32(_lib.lib().pharos_scenario_amplitude(cn0, spe))
    steps = [(40, 0, amp(38.0)), (137, 0, amp(52.5)), (250, 0, np.float32(np.nan)), (251, 0, amp(41.0)),
             (150, 3, amp(30.0)), (999, 3, amp(60.0)), (100, nch - 1, amp(47.25)), (1234, nch - 1, amp(35.0))]
    lv = np.array(steps, dtype=_lib.LEVEL_DTYPE)
    return lv[np.lexsort((lv["epoch"], lv["channel"]))]


def apply_levels(rows, levels, first):
    """The rows' amplitudes under the steps: the last step of the channel at or before each row's epoch."""
    out = rows.copy()
    for e in range(rows.shape[0]):
        for c in range(rows.shape[1]):
            if out[e, c]["table"] == _lib.INACTIVE:
                continue
            mine = levels[(levels["channel"] == c) & (levels["epoch"] <= first + e)]
            if len(mine) and not np.isnan(mine[-1]["amp"]):
                out[e, c]["amp"] = mine[-1]["amp"]
    return out


def test_device_levels_apply_at_their_epochs():
    sc = S.c1(seed=9, epochs=1500)
    first, epochs, block = 100, 1300, 100
    nch = len(sc.channels)
    levels = level_steps(nch, sc.samples_per_epoch)
    print(f"derive levels: c1, {len(levels)} steps, {epochs} epochs from {first}", flush=True)
    segs, masks = sc.segments(first, epochs, block)
    d_rows, _, _ = derive_rows(sc, first, epochs, block, segments=(segs, masks), levels=levels)
    got = d_rows.to_numpy(epochs * nch * 64).view(_lib.ROW_DTYPE).reshape(epochs, -1)
    plain = O.derive_segments(segs, sc.start_gps_ms, 0.0, first, epochs, block, sc.samples_per_epoch)
    want = apply_levels(plain, levels, first)
    assert got.tobytes() == want.tobytes()
    changed = want["amp"] != plain["amp"]
    assert changed[:, 0].any() and changed[:, 3].any() and changed[:, nch - 1].any() and not changed[:, 1].any()
    # A run derived with levels synthesises as the oracle does from the same rows (under a constant gain).
    sc.gain_db = 0.0
    job = Job(sc, first_epoch=first, epochs=epochs, derive=True, block_epochs=block, levels=levels)
    (job,) = run_batch([job])
    for e in (37, 151, 1133):
        want_b, _, _, _ = oracle_window(sc, first + e, 1, gain=1.0, rows=want[e:e + 1], masks=masks)
        got_b = job.out.to_numpy(sc.bytes_per_epoch, e * sc.bytes_per_epoch)
        assert np.array_equal(got_b, want_b), e
