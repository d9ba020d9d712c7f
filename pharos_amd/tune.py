"""The engine's own autotuner for the batched launch shape.

The tunable knobs are those of pharos_synth_tuning:
- the tile: 4096 samples on 128 threads, or 8192 on 256 (PHAROS_SYNTH_TUNE_WIDE_TILES) for the jobs whose epochs
  are longer than one narrow tile (the others then run narrow in a second launch);
- persistent workgroups per compute unit (1-4 narrow, 1-2 wide: a tile's LDS allows at most that many);
- channels staged per LDS pass of a tile (8-32).

For a workload shape, the tuner times every candidate on a short sample of the workload itself: the first
`sample_epochs` epochs, or as many as give MIN_TILES work items (a sample must fill the GPU several times over, or
every shape times alike), one batched launch, HIP events on the launch stream. It keeps the fastest candidate in a
JSON cache keyed by the shape. The shape key is (samples per epoch, format, channel count rounded up to a multiple
of 8, job count bucket). Every candidate produces the same bytes, so tuning changes speed only.
"""
import json
import os
import time

from . import _lib

WORKGROUPS = (1, 2, 3, 4)
WIDE_WORKGROUPS = (1, 2)
WIDE_TILES = 2          # PHAROS_SYNTH_TUNE_WIDE_TILES
MIN_TILES = 16384       # work items a tuning sample holds at least (16 per workgroup at 4 per CU on 256 CUs)
TILE = 4096             # samples of a narrow work item
CHANNELS_PER_PASS = (8, 16, 24, 32)
NARROW = [(w, c, 0) for w in WORKGROUPS for c in CHANNELS_PER_PASS]
WIDE = [(w, c, WIDE_TILES) for w in WIDE_WORKGROUPS for c in CHANNELS_PER_PASS]
CANDIDATES = NARROW + WIDE


def label(w, c, flags):
    return f"{w}x{c}" + ("w" if flags & WIDE_TILES else "")


def candidates(scenarios):
    """The shapes worth timing: wide tiles only when some scenario's epoch is longer than a narrow tile (the library
    runs those jobs wide and the others narrow, in a second launch)."""
    wide = any(s.samples_per_epoch > TILE for s in scenarios)
    return NARROW + (WIDE if wide else [])


def shape_key(scenarios):
    """Key of a batch of scenarios: the dominant (largest) scenario's rate, format and channel bucket, and the job
    count's power-of-two bucket."""
    big = max(scenarios, key=lambda s: s.channel_samples)
    nch = max(1, len(big.channels))
    jobs = 1
    while jobs < len(scenarios):
        jobs *= 2
    return f"{big.samples_per_epoch}/{big.format}/{(nch + 7) // 8 * 8}/{jobs}"


def _time(jobs, w, c, f, stream, repeats):
    from .engine import Event
    import ctypes as C
    L = _lib.lib()
    best = None
    for _ in range(repeats):
        for j in jobs:
            j.agc.gain, j.agc.clipped, j.agc.components, j.agc.adjustments = 1.0, 0, 0, 0
        arr = (_lib.SynthJob * len(jobs))(*[j.record() for j in jobs])
        tuning = _lib.SynthTuning(w, c, 0, f)
        e0, e1 = Event(), Event()
        h = C.c_void_p()
        e0.record(stream)
        _lib.check(L.pharos_synth_batch_submit(arr, len(jobs), C.byref(tuning), stream.handle, C.byref(h)))
        e1.record(stream)
        _lib.check(L.pharos_synth_batch_wait(h))
        ms = e0.elapsed_ms(e1)
        best = ms if best is None else min(best, ms)
    return best


def tune(scenarios, sample_epochs=300, repeats=2, cache_path=None, stream=None):
    """Times every candidate on the first `sample_epochs` epochs of `scenarios` (one batch) and returns
    {"key", "best": (workgroups_per_cu, channels_per_pass, flags), "times_ms": {...}, "seconds"}; stores the winner
    in the cache when one is given."""
    from .engine import Job, Stream
    t0 = time.perf_counter()
    stream = stream or Stream()
    tiles_per_epoch = sum(-(-s.samples_per_epoch // TILE) for s in scenarios)
    sample_epochs = max(sample_epochs, -(-MIN_TILES // max(1, tiles_per_epoch)))
    jobs = [Job(s, epochs=min(s.epochs, sample_epochs)) for s in scenarios]
    cands = candidates(scenarios)
    _time(jobs, 0, 0, 0, stream, 1)  # warm the code and the caches
    if any(f for _, _, f in cands):
        _time(jobs, 0, 0, WIDE_TILES, stream, 1)
    times = {label(*k): _time(jobs, *k, stream, repeats) for k in cands}
    best = min(cands, key=lambda k: times[label(*k)])
    result = {"key": shape_key(scenarios), "best": best, "times_ms": times, "sample_epochs": sample_epochs,
              "seconds": time.perf_counter() - t0}
    if cache_path:
        store(cache_path, result["key"], best, times[label(*best)])
    for j in jobs:
        j.engine.close()
    return result


def load(cache_path):
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            return json.load(f)
    return {}


def store(cache_path, key, best, ms):
    table = load(cache_path)
    table[key] = {"workgroups_per_cu": best[0], "channels_per_pass": best[1], "flags": best[2], "sample_ms": ms}
    os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
    tmp = cache_path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, cache_path)


def lookup(scenarios, cache_path):
    """(workgroups_per_cu, channels_per_pass, flags) cached for this shape, or (0, 0, 0): the library's defaults."""
    entry = load(cache_path).get(shape_key(scenarios))
    if not entry:
        return 0, 0, 0
    return int(entry["workgroups_per_cu"]), int(entry["channels_per_pass"]), int(entry.get("flags", 0))
