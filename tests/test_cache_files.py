"""The reference's cache files, .pht trajectories (src/pht/read.zig, write.zig), through the library's header check
and codec (pharos_amd/csrc/cache_codec.h, trajectory.cpp), against the five trajectory files the reference ships
(examples/trajectories/*.pht, written by its own encoder; copied to tests/golden/pht by copy_pht.py).

- Each file decodes, and encoding what was decoded gives the file back byte for byte: the header layout and
  version, the sources' digest, the u64 slice lengths, little-endian f64 bits and the field order are all as the
  reference wrote them.
- Every file carries the digest this library computes for the reference's sources (pharos_run_source_digest).
- The knots obey the trajectory's domain bounds (trajectory.zig:19-23), start at t = 0 and increase; the first
  position is the plan's start point (examples/trajectories/wuhan_north.json) in WGS84 ECEF x, y, z.
- Refusals follow header.zig:16-30 and codec.zig:96-104: a file cut anywhere is "truncated" at the start of the
  innermost value the cut falls in, or an "invalid value" where a slice counts more items than bytes are left;
  extra bytes are "trailing bytes".
"""
import glob
import math
import os
import struct

import numpy as np
import pytest

from pharos_amd import _lib, trajectory

HERE = os.path.join(os.path.dirname(__file__), "golden", "pht")
FILES = sorted(glob.glob(os.path.join(HERE, "*.pht")))


def _blob(path):
    with open(path, "rb") as f:
        return f.read()


def test_fixtures_are_present():
    assert len(FILES) == 5


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(p) for p in FILES])
def test_files_decode_and_encode_back_byte_for_byte(path):
    blob = _blob(path)
    name, knots = trajectory.read(blob)
    assert name == os.path.splitext(os.path.basename(path))[0]
    assert trajectory.write(name, knots) == blob
    assert blob[:4] == b"PHT1" and struct.unpack("<I", blob[4:8])[0] == 2
    assert struct.unpack("<Q", blob[8:16])[0] == _lib.lib().pharos_run_source_digest()
    # Domain bounds (trajectory.zig:19-23) and ordering.
    assert knots["t"][0] == 0.0 and np.all(np.diff(knots["t"]) > 0) and knots["t"][-1] <= 1e7
    r = np.linalg.norm(knots["pos"], axis=1)
    assert np.all((r >= 6_350_000) & (r <= 2e8))
    assert np.all(np.linalg.norm(knots["vel"], axis=1) <= 20_000)
    assert np.all(np.linalg.norm(knots["acc"], axis=1) <= 1_000)


def test_static_trajectory_stands_still():
    name, knots = trajectory.read(_blob(os.path.join(HERE, "wuhan_static.pht")))
    assert len(knots) >= 2 and np.all(knots["vel"] == 0) and np.all(knots["acc"] == 0)
    assert np.all(knots["pos"] == knots["pos"][0])


def _wgs84_ecef(lat_deg, lon_deg, h):
    a, f = 6378137.0, 1 / 298.257223563
    e2 = f * (2 - f)
    lat, lon = math.radians(lat_deg), math.radians(lon_deg)
    n = a / math.sqrt(1 - e2 * math.sin(lat) ** 2)
    return np.array([(n + h) * math.cos(lat) * math.cos(lon), (n + h) * math.cos(lat) * math.sin(lon),
                     (n * (1 - e2) + h) * math.sin(lat)])


def test_first_position_is_the_plans_start_point():
    # wuhan_north.json: start at place A, latitude 30.5292, longitude 114.3642, altitude 40 m above the EGM96
    # geoid, which lies within 110 m of the WGS84 ellipsoid everywhere: the knot must be within that of the point.
    _, knots = trajectory.read(_blob(os.path.join(HERE, "wuhan_north.pht")))
    want = _wgs84_ecef(30.5292, 114.3642, 40.0)
    assert np.linalg.norm(knots["pos"][0] - want) < 110.0
    # and it moves north: along +z, by about 1.8 km in its 60 s (to latitude 30.545436)
    assert knots["t"][-1] == 60.0 and 1500 < knots["pos"][-1][2] - knots["pos"][0][2] < 2000


def _expected_refusal(blob, cut):
    """(code, offset) of a file cut at `cut`: truncated at the start of the innermost value the cut falls in
    (codec.zig:11-17); but a slice whose item count exceeds the bytes left is an invalid value at the slice, before
    any item is read (codec.zig:39)."""
    if cut < 16:
        return 602, cut               # header.zig:24: truncated at the file's length
    name_len = struct.unpack("<Q", blob[16:24])[0]
    if cut < 24 + name_len:
        return 602, 16                # inside the name's length or its bytes: the name slice
    knots_at = 24 + name_len
    if cut < knots_at + 8:
        return 602, knots_at          # inside the knots' length
    count = struct.unpack("<Q", blob[knots_at:knots_at + 8])[0]
    if count > cut - knots_at - 8:
        return 603, knots_at          # more knots than bytes left
    return 602, knots_at + 8 + (cut - knots_at - 8) // 8 * 8  # inside one f64 of a knot


def test_a_file_cut_anywhere_is_refused_at_the_innermost_value():
    blob = _blob(os.path.join(HERE, "wuhan_drive.pht"))
    seen = set()
    for cut in range(len(blob)):
        with pytest.raises(trajectory.CacheRefused) as e:
            trajectory.read(blob[:cut])
        assert (e.value.code, e.value.offset) == _expected_refusal(blob, cut), cut
        seen.add(e.value.code)
    assert seen == {602, 603}


def test_invalid_values_trailing_bytes_and_headers():
    blob = bytearray(_blob(os.path.join(HERE, "wuhan_drive.pht")))
    name_len = struct.unpack("<Q", blob[16:24])[0]
    knots_at = 24 + name_len
    # More knots than bytes left: invalid value at the slice (codec.zig:39).
    bad = bytearray(blob)
    bad[knots_at:knots_at + 8] = struct.pack("<Q", 1 << 40)
    with pytest.raises(trajectory.CacheRefused) as e:
        trajectory.read(bytes(bad))
    assert e.value.code == 603 and e.value.offset == knots_at
    # A name longer than the file: its bytes run out (u8 slices take their bytes at once, codec.zig:38).
    bad = bytearray(blob)
    bad[16:24] = struct.pack("<Q", 1 << 40)
    with pytest.raises(trajectory.CacheRefused) as e:
        trajectory.read(bytes(bad))
    assert e.value.code == 602 and e.value.offset == 16
    with pytest.raises(trajectory.CacheRefused) as e:
        trajectory.read(bytes(blob) + b"\0")
    assert e.value.code == 604 and e.value.offset == len(blob)
    for at, code, patch in ((0, 600, b"PHB1"), (4, 600, struct.pack("<I", 3)), (8, 601, struct.pack("<Q", 1))):
        bad = bytearray(blob)
        bad[at:at + len(patch)] = patch
        with pytest.raises(trajectory.CacheRefused) as e:
            trajectory.read(bytes(bad))
        assert e.value.code == code and e.value.offset == at, code


def test_encoder_round_trips_edge_values():
    knots = np.zeros(3, dtype=_lib.KNOT_DTYPE)
    knots["t"] = [0.0, 1e-300, 5e6]
    knots["pos"][1] = [-0.0, float("inf"), 2.5e-310]     # bits survive, including a denormal and -0
    blob = trajectory.write("é path", knots)
    name, got = trajectory.read(blob)
    assert name == "é path" and got.tobytes() == knots.tobytes()
    assert trajectory.read(trajectory.write("", knots[:0]))[1].size == 0
