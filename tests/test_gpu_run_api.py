"""pharos_run.h on the GPU: pull, seek, submit, events, to_file, tag and manifest over the engine, with the
reference's rules (src/products/run/c_api/run.zig, engine/run.zig:225-331, engine/commands.zig, engine/events.zig).

A run's bytes equal the batched form's on the same segment-derived rows, in any pull sizes; sampled epochs equal the
oracle's; seek(T) lands on the sequential run's bytes; a signal-power command changes exactly the epochs from its
apply epoch on, as the same run with that level from the start does."""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib, scenario as S
from pharos_amd.engine import Job, run_batch
from pharos_amd.run import Run, STATUS
from tests.helpers import oracle_window

pytestmark = pytest.mark.gpu


def small(epochs=450, gain_db=None, fmt="IQ8", spe=16_368, seed=5):
    r = np.random.default_rng(seed)
    plan = [("L1CA", 3), ("L1CA", 17), ("L1CA", 29), ("E1", 4), ("E1", 21), ("B1I", 9), ("G1", 2), ("L1C", 11)]
    chans = [S.Channel(sig, prn, freq_k=(-3 if sig == "G1" else 0), if_hz=int(r.integers(-2_000_000, 2_000_000)),
                       range0_m=float(r.uniform(2.0e7, 2.6e7)), doppler_hz=float(r.uniform(-4000, 4000)),
                       doppler_rate_hz_s=float(r.uniform(-1, 1)), cn0_dbhz=float(r.uniform(40, 50)),
                       nav_seed=int(r.integers(1, 2**62))) for sig, prn in plan]
    sc = S.Scenario("run_api_8ch", spe, fmt, epochs, chans, noise=True, seed=seed, gain_db=gain_db)
    sc.start_gps_sub_ms = 0.125
    return sc


def test_pull_equals_the_batched_form_and_the_oracle():
    sc = small(gain_db=0.0)
    print(f"run api: {len(sc.channels)} channels, {sc.epochs} epochs, constant gain", flush=True)
    run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    info = run.info()
    assert info.total_epochs == sc.epochs and info.bytes_per_epoch == sc.bytes_per_epoch and info.format_bits == 8
    assert info.sample_rate_hz == sc.samples_per_epoch * 1000 and info.start_gps_sub_ms == 0.125
    data = run.pull_all(37)
    (job,) = run_batch([Job(sc, derive=True)])
    assert np.array_equal(data, job.result())
    segs, masks = sc.segments(0, sc.epochs, 100)
    rows = O.derive_segments(segs, sc.start_gps_ms, sc.start_gps_sub_ms, 0, sc.epochs, 100, sc.samples_per_epoch)
    for e in (0, 199, 449):
        want, _, _, _ = oracle_window(sc, e, 1, gain=1.0, rows=rows[e:e + 1], masks=masks)
        assert np.array_equal(data[e * sc.bytes_per_epoch:(e + 1) * sc.bytes_per_epoch], want), e
    st, blk, _, _ = run.pull(10)
    assert st == STATUS["end"] and blk[1] == 0


def test_adaptive_gain_pulls_and_agc_events():
    sc = small(epochs=620, fmt="IQ4", gain_db=None)
    for c in sc.channels:
        c.cn0_dbhz = 62.0          # strong channels clip, so the AGC moves
    run = Run(sc.bundle(), seed=sc.seed)
    data = run.pull_all(71)
    (job,) = run_batch([Job(sc, derive=True)])
    assert np.array_equal(data, job.result())
    agc = [(e, a, b) for kind, e, a, b, c, _ in run.events() if kind == 1]
    assert agc == job.adjustments() and len(agc) >= 2
    assert run.seek(10) == STATUS["arg"] and "constant gain" in run.error()


def test_seek_lands_on_the_sequential_bytes():
    sc = small(gain_db=-3.0)
    ref = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=-3.0).pull_all(100)
    run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=-3.0)
    bpe = sc.bytes_per_epoch
    for target in (333, 17, 449, 0, 250):
        assert run.seek(target) == STATUS["ok"]
        st, blk, data, _ = run.pull(60)
        assert st == STATUS["ok"] and blk[0] == target and blk[1] == min(60, sc.epochs - target)
        assert np.array_equal(data, ref[target * bpe:(target + blk[1]) * bpe]), target
    assert run.seek(sc.epochs) == STATUS["ok"] and run.pull(5)[0] == STATUS["end"]
    assert run.seek(sc.epochs + 1) == STATUS["arg"]


def test_commands_apply_at_their_epoch():
    sc = small(gain_db=0.0)
    sat = sc.channels[0]
    system = {"L1CA": 0, "L1C": 0, "B1I": 1, "E1": 2, "G1": 3}[sat.signal]
    base = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0).pull_all(450)
    run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    assert run.submit(2, value=31.0, apply_at_epoch=150, system=system, svid=sat.prn, unit=0) == STATUS["ok"]
    st, blk, data1, applied1 = run.pull(120)
    assert st == STATUS["ok"] and applied1 == []
    st, blk, data2, applied2 = run.pull(200)
    assert applied2 == [(150, 2, 31.0)]
    assert [(k, e, code) for k, e, _, _, _, code in run.events()] == [(2, 150, 0)]
    data = np.concatenate([data1, data2, run.pull_all(450)])
    bpe = sc.bytes_per_epoch
    assert np.array_equal(data[:150 * bpe], base[:150 * bpe])
    assert not np.array_equal(data[150 * bpe:200 * bpe], base[150 * bpe:200 * bpe])
    # From epoch 150 on, the run is the same scenario with that satellite's C/N0 at 31 dB-Hz from the start.
    for c in sc.channels:
        if c.signal == sat.signal and c.prn == sat.prn:
            c.cn0_dbhz = 31.0
    lowered = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    assert lowered.seek(150) == STATUS["ok"]
    assert np.array_equal(lowered.pull_all(450), data[150 * bpe:])
    # The run has applied a command: starting it over would drop it.
    assert run.seek(0) == STATUS["arg"] and "applied" in run.error()


def test_successive_levels_of_one_satellite():
    """Three commands on one satellite in three blocks (31 dB-Hz at 150, 45 at 260, its default again at 380): each
    stretch equals the scenario with that satellite at that level from the start, so the device's level steps keep
    every command's epoch while the older steps of a channel are dropped as the run moves on."""
    sc = small(gain_db=0.0, epochs=520)
    sat = sc.channels[3]
    system = {"L1CA": 0, "L1C": 0, "B1I": 1, "E1": 2, "G1": 3}[sat.signal]
    print("case levels", sat.signal, sat.prn, flush=True)
    run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    assert run.submit(2, value=31.0, apply_at_epoch=150, system=system, svid=sat.prn, unit=0) == STATUS["ok"]
    assert run.submit(2, value=45.0, apply_at_epoch=260, system=system, svid=sat.prn, unit=0) == STATUS["ok"]
    assert run.submit(2, apply_at_epoch=380, system=system, svid=sat.prn, use_default=1) == STATUS["ok"]
    data = run.pull_all(70)
    bpe = sc.bytes_per_epoch
    assert [(k, e) for k, e, _, _, _, _ in run.events()] == [(2, 150), (2, 260), (2, 380)]

    def at_level(cn0, first, last):
        for c in sc.channels:
            if c.signal == sat.signal and c.prn == sat.prn:
                saved, c.cn0_dbhz = c.cn0_dbhz, cn0
        other = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
        for c in sc.channels:
            if c.signal == sat.signal and c.prn == sat.prn:
                c.cn0_dbhz = saved
        assert other.seek(first) == STATUS["ok"]
        st, _, got, _ = other.pull(last - first)
        assert st == STATUS["ok"]
        return got

    base = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0).pull_all(520)
    assert np.array_equal(data[:150 * bpe], base[:150 * bpe])
    assert np.array_equal(data[150 * bpe:260 * bpe], at_level(31.0, 150, 260))
    assert np.array_equal(data[260 * bpe:380 * bpe], at_level(45.0, 260, 380))
    assert np.array_equal(data[380 * bpe:], base[380 * bpe:])
    assert not np.array_equal(data[150 * bpe:380 * bpe], base[150 * bpe:380 * bpe])


def test_submit_rules():
    sc = small(gain_db=0.0, epochs=300)
    run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    assert run.submit(7) == STATUS["arg"]                                   # undefined kind
    assert run.submit(0, value=10.0) == STATUS["arg"]                       # noise floor above 0 dBm/Hz
    assert run.submit(1, value=95.0) == STATUS["arg"]                       # elevation mask beyond 90 degrees
    assert run.submit(2, value=45.0, system=0, svid=40) == STATUS["arg"] and "at most 32" in run.error()
    assert run.submit(2, value=45.0, system=5) == STATUS["arg"]
    run.pull(50)
    assert run.submit(2, value=45.0, apply_at_epoch=20) == STATUS["arg"] and "passed" in run.error()
    for k in range(256):
        assert run.submit(1, value=5.0, apply_at_epoch=1000 + k) == STATUS["ok"]
    assert run.submit(1, value=5.0, apply_at_epoch=2000) == STATUS["busy"]


def test_noise_floor_command_moves_every_level():
    sc = small(gain_db=0.0, epochs=200)
    run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    assert run.submit(0, value=-170.0, apply_at_epoch=100) == STATUS["ok"]   # 4 dB above the -174 dBm/Hz floor
    data = run.pull_all(200)
    for c in sc.channels:
        c.cn0_dbhz -= 4.0
    lowered = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    lowered.seek(100)
    bpe = sc.bytes_per_epoch
    assert np.array_equal(lowered.pull_all(200), data[100 * bpe:])


def test_pull_rules_to_file_tag_and_manifest(tmp_path):
    sc = small(gain_db=0.0, epochs=260)
    run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
    bpe = sc.bytes_per_epoch
    assert run.pull(0)[0] == STATUS["arg"]
    assert run.pull(5, cap=bpe - 1)[0] == STATUS["arg"] and "one epoch" in run.error()
    assert run.seek(60) == STATUS["ok"]
    path = tmp_path / "run.bin"
    n = C.c_uint64()
    assert _lib.lib().pharos_run_to_file(run.handle, str(path).encode(), C.byref(n)) == STATUS["ok"]
    ref = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0).pull_all(100)
    assert n.value == (sc.epochs - 60) * bpe and np.array_equal(np.fromfile(path, dtype=np.uint8), ref[60 * bpe:])
    bad = tmp_path / "missing" / "x.bin"
    assert _lib.lib().pharos_run_to_file(run.handle, str(bad).encode(), C.byref(n)) == STATUS["io"]
    buf = (C.c_uint8 * 512)()
    ln = C.c_size_t()
    assert _lib.lib().pharos_run_tag_text(run.handle, buf, 512, C.byref(ln)) == STATUS["ok"]
    text = bytes(buf[:ln.value]).decode()
    assert text.startswith("PROG = pharos-run\n") and "FMT  = CS8" in text
    fields = _lib.RunManifestFields(b"2026-01-01T00:00:00Z", b"run-7", b"/tmp/x.bin")
    big = (C.c_uint8 * 65536)()
    assert _lib.lib().pharos_run_manifest_json(run.handle, C.byref(fields), big, 65536, C.byref(ln)) == STATUS["ok"]
    m = json.loads(bytes(big[:ln.value]).decode())
    assert m["run_id"] == "run-7" and m["samples"]["total_epochs"] == sc.epochs and len(m["channels"]) == len(sc.channels)
    bad_fields = _lib.RunManifestFields(b"\xff\xfe", None, None)
    assert _lib.lib().pharos_run_manifest_json(run.handle, C.byref(bad_fields), big, 65536, C.byref(ln)) == STATUS["arg"]


def _weighted(data):
    """sum of b[i] * (i + 1), mod 2^64 (tests/c/run_pull.c's checksum)."""
    i = np.arange(1, data.size + 1, dtype=np.uint64)
    return int(np.sum(data.astype(np.uint64) * i, dtype=np.uint64))


def test_a_c_caller_pulls_and_seeks_the_same_bytes(tmp_path):
    """tests/c/run_pull.c, a plain C program linked against the library, drives the run through pharos_run.h alone:
    its bytes and clip count equal the Python binding's pull, and its seek lands on the same chunk."""
    exe = os.path.join(os.path.dirname(__file__), "c", "run_pull")
    assert os.path.exists(exe), "tests/c/run_pull is built with the library (python -m pharos_amd.build)"
    sc = small(gain_db=0.0, epochs=330)
    path = tmp_path / "run.phsc"
    path.write_bytes(sc.bundle())
    print("C caller:", exe, sc.name, sc.epochs, "epochs", flush=True)
    out = subprocess.run([exe, str(path), "37", "123"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout)
    run = Run(sc.bundle(), seed=5, gain=1, gain_db=0.0)
    data = run.pull_all(37)
    assert got["epochs"] == sc.epochs and got["bytes"] == data.size and got["abi"] == 0x00010000
    assert got["sum"] == _weighted(data)
    st, blk, _, _ = Run(sc.bundle(), seed=5, gain=1, gain_db=0.0).pull(sc.epochs)
    assert st == STATUS["ok"] and got["clipped"] == blk[3]
    bpe = sc.bytes_per_epoch
    assert got["seek_first"] == 123 and got["seek_epochs"] == 37
    assert got["seek_sum"] == _weighted(data[123 * bpe:160 * bpe])


RUN_CHILD = r'''
import ctypes as C, json, os, resource, signal, sys
sys.path.insert(0, os.getcwd())
from pharos_amd import _lib
from pharos_amd.run import Run
from tests.test_gpu_run_api import small
signal.signal(signal.SIGXFSZ, signal.SIG_IGN)            # a write past the limit fails with EFBIG instead
limit = int(sys.argv[2])
sc = small(gain_db=0.0, epochs=450)
run = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0)
assert run.seek(60) == 0
resource.setrlimit(resource.RLIMIT_FSIZE, (limit, limit))
n = C.c_uint64(0)
st = _lib.lib().pharos_run_to_file(run.handle, os.fsencode(sys.argv[1]), C.byref(n))
print(json.dumps({"status": st, "bytes_written": n.value, "error": run.error()}), flush=True)
'''


def test_to_file_reports_the_bytes_on_file_after_a_failed_write(tmp_path):
    """A write that fails part-way (a file size limit) stops the run's file at the bytes written in order, reported
    with status io (c_api/run.zig:245-259): the first block (epochs 60-99, up to the grid), one whole block, then
    half of the third."""
    sc = small(gain_db=0.0, epochs=450)
    bpe = sc.bytes_per_epoch
    limit = 40 * bpe + 100 * bpe + 50 * bpe
    path = tmp_path / "limited.bin"
    print(f"to_file with a file size limit of {limit} bytes", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", RUN_CHILD, str(path), str(limit)], cwd=root, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert res["status"] == STATUS["io"] and "cannot write" in res["error"], res
    assert res["bytes_written"] == limit, res
    ref = Run(sc.bundle(), seed=sc.seed, gain=1, gain_db=0.0).pull_all(100)
    data = np.fromfile(path, dtype=np.uint8)
    assert data.size == limit and np.array_equal(data, ref[60 * bpe:60 * bpe + limit])
