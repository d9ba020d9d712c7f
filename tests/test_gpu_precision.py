"""The split-precision report (pharos_synth_precision): the device diagnostic against the same error measured on the
CPU from the oracle's rotor (orc_rotor_values) with numpy FP64, and the bound the design states for the multi-GNSS
and wideband configurations: carrier phase error at most 1e-6 cycles over the run."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib, scenario as S
from pharos_amd.engine import DeviceBuffer, Engine

pytestmark = pytest.mark.gpu


def device_report(sc, rows):
    eng = Engine(sc)
    d_rows = DeviceBuffer.from_array(rows)
    rep = _lib.PrecisionReport()
    per_table = (C.c_double * max(1, len(sc.channels)))()
    _lib.check(_lib.lib().pharos_synth_precision(eng.handle, d_rows.ptr, rows.size, C.byref(rep), per_table, None))
    eng.close()
    return rep, np.array(per_table[:len(sc.channels)])


def cpu_rotor_error(row, spe):
    """Largest phase error (cycles) and magnitude error of the oracle's f32 rotor over one epoch, in FP64."""
    cre = np.zeros(spe, dtype=np.float32)
    cim = np.zeros(spe, dtype=np.float32)
    O.lib().orc_rotor_values(float(row["carrier_phase0"]), float(row["carrier_step"]), 0, spe,
                             cre.ctypes.data, cim.ctypes.data)
    n = np.arange(spe, dtype=np.float64)
    ph = float(row["carrier_phase0"]) + n * float(row["carrier_step"])
    ph -= np.floor(ph)
    e = np.exp(2j * np.pi * ph)
    z = (cre.astype(np.float64) + 1j * cim.astype(np.float64)) * np.conj(e)
    return float(np.max(np.abs(np.angle(z)))) / (2 * np.pi), float(np.max(np.abs(np.abs(z) - 1.0)))


def test_device_report_matches_the_oracle_rotor_on_the_cpu():
    sc = S.c2(seed=5, epochs=3)
    rows, _ = sc.window(0, 3)
    flat = rows.reshape(-1)
    for r in flat[:6]:
        single = np.array([r], dtype=_lib.ROW_DTYPE)
        rep, _ = device_report(sc, single)
        want_ph, want_mag = cpu_rotor_error(r, sc.samples_per_epoch)
        assert want_ph > 0
        assert abs(rep.carrier_cycles - want_ph) <= 1e-3 * want_ph + 1e-13, (rep.carrier_cycles, want_ph)
        assert abs(rep.magnitude - want_mag) <= 1e-2 * want_mag + 1e-12, (rep.magnitude, want_mag)


@pytest.mark.parametrize("make,epochs", [(S.c1, 2000), (S.c2, 1000), (S.c0, 2000)])
def test_phase_error_is_below_1e6_cycles(make, epochs):
    sc = make(seed=1, epochs=epochs)
    rows, _ = sc.window(0, epochs)
    print(f"precision: {sc.name} over {epochs} epochs", flush=True)
    rep, per_table = device_report(sc, rows.reshape(-1))
    assert rep.rows == rows.size
    assert 0 < rep.carrier_cycles <= 1e-6, rep.carrier_cycles
    assert rep.magnitude <= 2e-6, rep.magnitude
    # Q32.32: (n + 1) 2^-33 entries at most over an epoch's n samples.
    assert rep.code_entries <= (sc.samples_per_epoch + 1) * 2.0 ** -33, rep.code_entries
