"""pharos_run.h without a GPU: the record layouts the reference's header generator gives (c_api/run.zig:25-61,
interface/c/command.zig:9-17, c_api/convert.zig:7-19), the status values (interface/c/status.zig:5 and the run's
io/bundle/busy), null-argument rules, bundle refusal before any device call, and the reference sources' digest a
.phb header carries (build.zig:38-50) with its hash pinned by Zig's published Wyhash test vectors."""
import ctypes as C
import os

import pytest

from pharos_amd import _lib, scenario as S
from pharos_amd.source_digest import source_digest, wyhash


def test_record_layouts():
    assert C.sizeof(_lib.RunBlock) == 48
    assert [getattr(_lib.RunBlock, f).offset for f in ("epoch_first", "epochs", "bytes", "clipped", "applied",
                                                       "applied_count")] == [0, 8, 16, 24, 32, 40]
    assert C.sizeof(_lib.RunOptions) == 32 and _lib.RunOptions.gain_db.offset == 24
    assert C.sizeof(_lib.RunCommand) == 40 and _lib.RunCommand.kind.offset == 16
    assert C.sizeof(_lib.RunApplied) == 48
    assert C.sizeof(_lib.RunEvent) == 56 and _lib.RunEvent.kind.offset == 32
    assert C.sizeof(_lib.RunInfo) == 80 and _lib.RunInfo.leap_seconds_after.offset == 76


def test_null_arguments_and_versions():
    L = _lib.lib()
    assert L.pharos_run_abi_version() == 0x00010000
    b = _lib.RunBlock()
    assert L.pharos_run_pull(None, None, 0, 1, C.byref(b)) == 2
    assert L.pharos_run_seek(None, 0) == 2 and L.pharos_run_submit(None, None) == 2
    assert L.pharos_run_next_event(None, None) == 0 and L.pharos_run_events_dropped(None) == 0
    assert L.pharos_run_error(None) == b""
    assert L.pharos_run_info_get(None, None) == 2
    h = C.c_void_p()
    assert L.pharos_run_create(None, 0, None, C.byref(h)) == 2 and not h.value
    L.pharos_run_destroy(None)


def test_bundles_are_checked_before_any_device_call():
    L = _lib.lib()
    h = C.c_void_p()
    good = S.c0(epochs=10).bundle()
    assert L.pharos_run_create(b"JUNK" + good[4:], len(good), None, C.byref(h)) == 5
    assert L.pharos_run_create(good[:-1], len(good) - 1, None, C.byref(h)) == 5
    digest = L.pharos_run_source_digest()
    phb = b"PHB1" + (2).to_bytes(4, "little") + digest.to_bytes(8, "little")
    assert L.pharos_run_create(phb, 16, None, C.byref(h)) == 7            # a real bundle: refused, E_BUNDLE
    assert b"epoch-rate stage" in L.pharos_run_last_error()
    old = b"PHB1" + (1).to_bytes(4, "little") + digest.to_bytes(8, "little")
    assert L.pharos_run_create(old, 16, None, C.byref(h)) == 5
    other = b"PHB1" + (2).to_bytes(4, "little") + (digest ^ 1).to_bytes(8, "little")
    assert L.pharos_run_create(other, 16, None, C.byref(h)) == 5
    assert b"other sources" in L.pharos_run_last_error()
    opts = _lib.RunOptions(0, 0, 2, 0, 0.0)
    assert L.pharos_run_create(good, len(good), C.byref(opts), C.byref(h)) == 2   # noise mode 2
    opts = _lib.RunOptions(0, 0, 0, 0, 3.0)
    assert L.pharos_run_create(good, len(good), C.byref(opts), C.byref(h)) == 2   # gain_db under adaptive gain
    opts = _lib.RunOptions(0, 0, 0, 1, 61.0)
    assert L.pharos_run_create(good, len(good), C.byref(opts), C.byref(h)) == 2   # beyond +-60 dB


def test_wyhash_matches_zigs_test_vectors():
    vectors = [(0, 0x409638EE2BDE459, b""), (1, 0xA8412D091B5FE0A9, b"a"), (2, 0x32DD92E4B2915153, b"abc"),
               (3, 0x8619124089A3A16B, b"message digest"), (4, 0x7A43AFB61D7F5F40, b"abcdefghijklmnopqrstuvwxyz"),
               (5, 0xFF42329B90E50D58, b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789"),
               (6, 0xC39CAB13B115AAD3, b"1234567890" * 8)]
    for seed, want, data in vectors:
        assert wyhash(data, seed) == want, data


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"), reason="the reference tree is not on this host")
def test_source_digest_is_the_reference_trees():
    assert source_digest("/root/reference") == _lib.lib().pharos_run_source_digest()


C_RUN_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "pharos_run.h"
#include "pharos_synth.h"
#define F(T, f) printf("[\"%s\", \"%s\", %zu],", #T, #f, offsetof(T, f));
int main(void) {
  printf("{\"sizes\": {\"pharos_run_options\": %zu, \"pharos_run_info\": %zu, \"pharos_run_command\": %zu, "
         "\"pharos_run_applied\": %zu, \"pharos_run_block\": %zu, \"pharos_run_event\": %zu, "
         "\"pharos_synth_segment\": %zu}, \"fields\": [",
         sizeof(pharos_run_options), sizeof(pharos_run_info), sizeof(pharos_run_command), sizeof(pharos_run_applied),
         sizeof(pharos_run_block), sizeof(pharos_run_event), sizeof(pharos_synth_segment));
  F(pharos_run_options, seed) F(pharos_run_options, gain_db) F(pharos_run_info, sample_rate_hz)
  F(pharos_run_info, start_gps_sub_ms) F(pharos_run_info, leap_seconds_after) F(pharos_run_command, kind)
  F(pharos_run_command, use_default) F(pharos_run_block, bytes) F(pharos_run_block, applied_count)
  F(pharos_run_event, kind) F(pharos_run_event, signal) F(pharos_synth_segment, lead_entries)
  F(pharos_synth_segment, mask_first) F(pharos_synth_segment, table)
  printf("[\"end\", \"\", 0]]}\n");
  return 0;
}
"""


def test_header_layouts_match_the_c_compiler(tmp_path):
    """pharos_run.h and the segment record as a C compiler lays them out, against the ctypes binding and the
    reference's records (c_api/run.zig:25-61, interface/c/command.zig:9-17, c_api/convert.zig:7-19)."""
    import json
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "probe.c"
    src.write_text(C_RUN_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    ctypes_of = {"pharos_run_options": _lib.RunOptions, "pharos_run_info": _lib.RunInfo,
                 "pharos_run_command": _lib.RunCommand, "pharos_run_applied": _lib.RunApplied,
                 "pharos_run_block": _lib.RunBlock, "pharos_run_event": _lib.RunEvent}
    for name, size in got["sizes"].items():
        if name == "pharos_synth_segment":
            assert size == _lib.SEGMENT_DTYPE.itemsize == 136
        else:
            assert size == C.sizeof(ctypes_of[name]), name
    assert got["sizes"]["pharos_run_block"] == 48
    for name, field, off in got["fields"][:-1]:
        if name == "pharos_synth_segment":
            assert _lib.SEGMENT_DTYPE.fields[field][1] == off, field
        else:
            assert getattr(ctypes_of[name], field).offset == off, (name, field)
