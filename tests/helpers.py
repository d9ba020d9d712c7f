"""Shared helpers of the parity tests: run a scenario window on the HIP engine and on the CPU oracle."""
import numpy as np

from oracle import binding as O

F32 = 4


def oracle_tables(scenario):
    return O.Tables([(c.signal, c.prn) for c in scenario.channels])


def oracle_window(scenario, first_epoch, epochs, gain=1.0, threads=8, rows=None, masks=None):
    """Oracle bytes, clips and pre-quantisation sums of a window."""
    if rows is None:
        rows, masks = scenario.window(first_epoch, epochs)
    tabs = oracle_tables(scenario)
    out, clips, sre, sim = O.synth_window(tabs, rows, masks, scenario.samples_per_epoch, scenario.format_id,
                                          0 if scenario.noise else 1, scenario.seed, first_epoch, gain, threads,
                                          want_sums=True)
    return out, clips, sre, sim


def gpu_window(scenario, first_epoch, epochs, gain=1.0, sums=True, rows=None, masks=None):
    """GPU bytes and clips of a window (plain launch), and its pre-quantisation sums from the F32 format."""
    from pharos_amd.engine import Engine, DeviceBuffer
    if rows is None:
        rows, masks = scenario.window(first_epoch, epochs)
    eng = Engine(scenario)
    d_rows, d_masks = DeviceBuffer.from_array(rows), DeviceBuffer.from_array(masks)
    out = DeviceBuffer(epochs * eng.bytes_per_epoch)
    clips = eng.window(first_epoch, epochs, d_rows.ptr, d_masks.ptr, len(masks), gain, out.ptr)
    result = out.to_numpy()
    eng.close()
    sre = sim = None
    if sums:
        eng32 = Engine(scenario, format_override=F32)
        out32 = DeviceBuffer(epochs * eng32.bytes_per_epoch)
        eng32.window(first_epoch, epochs, d_rows.ptr, d_masks.ptr, len(masks), gain, out32.ptr)
        pairs = out32.to_numpy().view(np.float32).reshape(-1, 2)
        sre, sim = pairs[:, 0].copy(), pairs[:, 1].copy()
        eng32.close()
    return result, clips, sre, sim


def describe_mismatch(a, b, label):
    bad = np.nonzero(a != b)[0]
    if len(bad) == 0:
        return f"{label}: equal"
    return f"{label}: {len(bad)} of {len(a)} differ, first at {bad[:8]}: {a[bad[:8]]} vs {b[bad[:8]]}"
