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
