"""Invalid input on the GPU: every case is one the kernel's own checks turn into an error status (no access
outside its buffers), after which the engine keeps working on valid input."""
import ctypes as C

import numpy as np
import pytest

from pharos_amd import _lib, scenario as S
from pharos_amd.engine import Engine, DeviceBuffer, Job, run_batch
from tests.helpers import oracle_window

pytestmark = pytest.mark.gpu


def _run(eng, rows, masks, first=0, mask_len=None):
    d_rows, d_masks = DeviceBuffer.from_array(rows), DeviceBuffer.from_array(masks)
    out = DeviceBuffer(rows.shape[0] * eng.bytes_per_epoch)
    clips = eng.window(first, rows.shape[0], d_rows.ptr, d_masks.ptr, len(masks) if mask_len is None else mask_len,
                       1.0, out.ptr)
    return out.to_numpy(), clips


def _case(mutate, expect):
    sc = S.c0(seed=21, epochs=4)
    eng = Engine(sc)
    rows, masks = sc.window(10, 4)
    bad = rows.copy()
    mask_len = mutate(bad)
    print(f"case {expect}", flush=True)
    with pytest.raises(_lib.PharosError) as exc:
        _run(eng, bad, masks, first=10, mask_len=mask_len)
    assert expect in str(exc.value) or expect in _lib.lib().pharos_synth_last_error().decode()
    # the engine is still good: valid rows give the oracle's bytes
    out, clips = _run(eng, rows, masks, first=10)
    ob, oc, _, _ = oracle_window(sc, 10, 4, 1.0, rows=rows, masks=masks)
    assert np.array_equal(out, ob) and np.array_equal(clips, oc)
    eng.close()


def test_table_index_out_of_range():
    def m(r):
        r["table"][1, 2] = 99
    _case(m, "bad table")


def test_zero_code_rate():
    def m(r):
        r["entries_per_sample"][0, 0] = 0.0
    _case(m, "bad table")


def test_code_rate_beyond_the_staged_range():
    def m(r):
        r["entries_per_sample"][2, 3] *= 1.5  # +50 %: more entries than a tile may stage
    _case(m, "staged range")


def test_mask_bytes_missing():
    def m(r):
        r["mask_count"][3, 1] = 0
    _case(m, "mask bytes missing")


def test_mask_buffer_too_short():
    def m(r):
        return 1  # mask_len of one byte: every row past its first period lacks masks
    _case(m, "mask bytes missing")


def test_lead_too_long():
    def m(r):
        r["lead_bits"][0, 5] = 1
        r["lead_entries"][0, 5] = 100.0
    _case(m, "staged range")


def test_argument_errors():
    L = _lib.lib()
    sc = S.c0(seed=2, epochs=2)
    eng = Engine(sc)
    rows, masks = sc.window(0, 2)
    d_rows, d_masks = DeviceBuffer.from_array(rows), DeviceBuffer.from_array(masks)
    out = DeviceBuffer(2 * eng.bytes_per_epoch)
    nch = len(sc.channels)
    assert L.pharos_synth_window(eng.handle, 0, 0, nch, d_rows.ptr, d_masks.ptr, len(masks), C.c_float(1.0), out.ptr,
                                 None, None) == 2
    assert L.pharos_synth_window(eng.handle, 0, 2, nch, d_rows.ptr, d_masks.ptr, len(masks), C.c_float(0.0), out.ptr,
                                 None, None) == 2
    assert L.pharos_synth_window(eng.handle, 0, 2, nch, d_rows.ptr, d_masks.ptr, len(masks), C.c_float(float("nan")),
                                 out.ptr, None, None) == 2
    tabs = sc.tables()
    h = C.c_void_p()
    for spe, fmt in [(511, 2), (200_001, 2), (4093, 0), (4092, 9)]:
        cfg = _lib.SynthConfig(spe, fmt, 1, 0, 1)
        assert L.pharos_synth_create(C.byref(cfg), tabs.array, tabs.count, 0, C.byref(h)) == 2, (spe, fmt)
    eng.close()


def test_batch_reports_a_bad_row_and_keeps_the_others():
    good = S.c0(seed=5, epochs=200)
    bad = S.c0(seed=6, epochs=200)
    rows, masks = bad.window(0, 200)
    rows["table"][150, 0] = 77
    jobs = [Job(good), Job(bad, rows=rows, masks=masks)]
    print("case batch with a bad row", flush=True)
    with pytest.raises(_lib.PharosError):
        run_batch(jobs)
    (again,) = run_batch([Job(good)])
    ob, _, _, _ = oracle_window(good, 0, 200, 1.0)
    assert np.array_equal(again.result(), ob)
