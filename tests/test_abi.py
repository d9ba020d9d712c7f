"""The C ABI without a GPU: the library loads, exports every function include/*.h declares, and its records have
the layout the headers give (checked against the C compiler's own sizeof/offsetof)."""
import ctypes as C
import glob
import json
import os
import re
import subprocess

import numpy as np
import pytest

from pharos_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pharos_\w+)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in DECL.finditer(text):
            line = text[m.start():text.index("\n", m.start())]
            if line.lstrip().startswith(("typedef", "#", "enum")):
                continue
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "pharos_synth_window" in names and "pharos_synth_batch_submit" in names
    assert len(names) >= 25


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = C.CDLL(_lib.LIB_PATH)
    assert hasattr(lib, name), f"{name} is declared in include/ but not exported"


def test_abi_version_and_error_text():
    L = _lib.lib()
    assert L.pharos_synth_abi_version() == 0x00010000
    assert isinstance(L.pharos_synth_last_error(), bytes)


def test_argument_errors_need_no_device():
    """Null arguments are refused with the argument status before any device call."""
    L = _lib.lib()
    out = C.c_void_p()
    assert L.pharos_synth_create(None, None, 0, 0, C.byref(out)) == 2
    assert b"null" in L.pharos_synth_last_error().lower() or L.pharos_synth_last_error()
    assert L.pharos_synth_window(None, 0, 1, 0, None, None, 0, C.c_float(1.0), None, None, None) == 2
    assert L.pharos_synth_batch_submit(None, 0, None, None, C.byref(out)) == 2
    n = C.c_uint64(7)
    assert L.pharos_synth_to_file(None, b"/nonexistent/x.bin", C.byref(n)) == 2 and n.value == 0


C_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "pharos_synth.h"
#define S(T) printf("\"%s\": [%zu", #T, sizeof(T));
#define F(T, f) printf(", [\"%s\", %zu]", #f, offsetof(T, f));
#define E printf("],\n");
int main(void) {
  printf("{\n");
  S(pharos_synth_config) F(pharos_synth_config, samples_per_epoch) F(pharos_synth_config, seed) E
  S(pharos_synth_table) F(pharos_synth_table, states) F(pharos_synth_table, tdm_pilot) F(pharos_synth_table, values_re)
      F(pharos_synth_table, values_im) E
  S(pharos_synth_row) F(pharos_synth_row, carrier_phase0) F(pharos_synth_row, period0) F(pharos_synth_row, table)
      F(pharos_synth_row, amp) F(pharos_synth_row, mask_offset) F(pharos_synth_row, mask_count)
      F(pharos_synth_row, lead_bits) F(pharos_synth_row, flags) E
  S(pharos_synth_agc) F(pharos_synth_agc, gain) F(pharos_synth_agc, adjustments) E
  S(pharos_synth_adjustment) F(pharos_synth_adjustment, epoch) F(pharos_synth_adjustment, reduced) E
  S(pharos_synth_job) F(pharos_synth_job, engine) F(pharos_synth_job, rows) F(pharos_synth_job, mask_len)
      F(pharos_synth_job, gain_mode) F(pharos_synth_job, constant_gain) F(pharos_synth_job, agc)
      F(pharos_synth_job, adjustments_cap) F(pharos_synth_job, clipped) E
  S(pharos_synth_level) F(pharos_synth_level, epoch) F(pharos_synth_level, channel) F(pharos_synth_level, amp) E
  printf("\"pharos_synth_tuning\": [%zu]\n}\n", sizeof(pharos_synth_tuning));
  return 0;
}
"""


def test_record_layouts_match_the_c_compiler(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(C_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    layout = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    ctypes_of = {"pharos_synth_config": _lib.SynthConfig, "pharos_synth_table": _lib.SynthTable,
                 "pharos_synth_agc": _lib.SynthAgc, "pharos_synth_adjustment": _lib.SynthAdjustment,
                 "pharos_synth_job": _lib.SynthJob, "pharos_synth_tuning": _lib.SynthTuning}
    py_name = {"samples_per_epoch": "samples_per_epoch"}
    for cname, entry in layout.items():
        size, fields = entry[0], entry[1:]
        if cname in ("pharos_synth_row", "pharos_synth_level"):
            dt = _lib.ROW_DTYPE if cname == "pharos_synth_row" else _lib.LEVEL_DTYPE
            assert dt.itemsize == size, cname
            for f, off in fields:
                assert dt.fields[f][1] == off, (cname, f)
            continue
        T = ctypes_of[cname]
        assert C.sizeof(T) == size, cname
        for f, off in fields:
            assert getattr(T, py_name.get(f, f)).offset == off, (cname, f)


def test_row_dtype_is_the_header_row():
    assert _lib.ROW_DTYPE.itemsize == 64
    r = np.zeros(1, _lib.ROW_DTYPE)
    r["table"] = _lib.INACTIVE
    assert r.view(np.uint8)[48:52].tolist() == [255, 255, 255, 255]
