"""Benchmark of the MI355X IF sample engine on the multi-GNSS configuration (one GPU).

A step synthesises the whole run of the multi-GNSS configuration: 40 channels of
GPS L1C/A+L5, Galileo E1+E5a, BeiDou B1I+B3I and GLONASS L1OF at 61.44 Msps, IQ4, 60 s (60 000 epochs), noise on,
adaptive AGC -- rows and masks resident in HBM before the timed region, the 3.69 GB of IQ4 output written to HBM.

  value        channel-samples/s of the batched form (one persistent launch per step, device AGC)
  graph        the same run as per-window launches with the AGC step on the device, captured once as a HIP graph
  plain        the same run as per-window plain launches with the host AGC in between
  roofline     FP32 VALU roofline of the synthesis kernel: algorithmic flops per launch / launch time
  cpu_baseline the C oracle (a restatement of the reference kernel), built for this host, on all its cores and on
               one thread, on a bounded sample
  precision    the f32 carrier rotor against FP64 and the Q32.32 code position against f64, every sample of the run
  run_latency  p50/p99 wall time of the timed batched runs
  serving      p50/p99 latency of a Poisson request stream (100 req/s for 10 s by default; --no-serve skips it),
               rows derived on the device from per-block segments; serving_to_host: 20 req/s with each output
               copied to host memory, latency to its last byte there
  run_api      the same configuration through include/pharos_run.h: pull into a host buffer in 100-epoch calls
               and to_file, rates to host memory (--no-run-api skips it)

--workload c3 instead measures the uneven batch: 256 seeded scenarios (4-48 channels, 4.092-61.44 Msps,
IQ2-IQ16, durations cut to --c3-epochs-cap epochs) as one batched persistent launch, against per-scenario plain
launches and per-scenario graphs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from pharos_amd import tune  # noqa: E402

METRIC = "channel-samples/sec + roofline %, batched vs plain on 1 GPU; p50/p99 run latency"
PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 (vector, packed)
PEAK_HBM_GBS = 8000.0
# Algorithmic FP32 flops per channel-sample (kernel.zig:241-258 plus rotor.zig:52-58): a complex MAC and a rotor
# step for two-axis signals (8 + 6), two FMAs and a rotor step for one-axis signals (4 + 6).
FLOPS_TWO_AXIS, FLOPS_ONE_AXIS = 14, 10
CPU_SHARE = 16  # host cores one GPU's job may use on a GPU box


def flops_per_epoch(scenario):
    from pharos_amd import _lib
    tabs = scenario.tables()
    total = 0
    for i in range(tabs.count):
        t = tabs.array[i]
        any_re = any(t.values_re[k] != 0 for k in range(8))
        any_im = any(t.values_im[k] != 0 for k in range(8))
        total += FLOPS_TWO_AXIS if (any_re and any_im) else FLOPS_ONE_AXIS
    return total * scenario.samples_per_epoch


def host_cpu():
    """The host's CPUs as this process sees them: nproc, the cores its affinity mask allows, the cgroup's CPU quota
    (cores, None when unlimited) and the processor model."""
    info = {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)), "cgroup_quota_cores": None,
            "model": None}
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            info["cgroup_quota_cores"] = int(quota) / int(period)
    except (OSError, ValueError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return info


def _oracle_rate(O, tabs, scenario, seconds, threads, max_epochs):
    """Channel-samples/s of the oracle over consecutive 100-epoch windows of the run, for `seconds` of wall time."""
    nch = len(scenario.channels)
    done, busy, t0 = 0, 0.0, time.perf_counter()
    while time.perf_counter() - t0 < seconds and done < max_epochs:
        rows, masks = scenario.window(done, 100)   # the host stage, outside the timed part
        t1 = time.perf_counter()
        O.synth_window(tabs, rows, masks, scenario.samples_per_epoch, scenario.format_id,
                       0 if scenario.noise else 1, scenario.seed, done, 1.0, threads=threads)
        busy += time.perf_counter() - t1
        done += 100
    return nch * scenario.samples_per_epoch * done / busy, done, busy


def cpu_baseline(scenario, seconds=10.0, single_seconds=4.0):
    """The oracle (the C restatement of the reference kernel), built -O3 -march=native for this host, on every core
    this process may use, over 100-epoch windows of the same workload; then on one thread.  Windows are timed
    around the synthesis alone (the reference's derive runs on its caller thread, not in the pool)."""
    import tempfile
    from oracle import binding as O
    host = host_cpu()
    threads = host["affinity_cores"]
    if host["cgroup_quota_cores"]:
        threads = max(1, min(threads, int(host["cgroup_quota_cores"] + 0.5)))
    # A GPU box grants one GPU's job a share of 16 host cores, whatever nproc shows; the pool stays within it.
    threads = min(threads, CPU_SHARE)
    build = "-O3 -march=native -ffp-contract=off"
    try:
        O.use(O.build_native(tempfile.mkdtemp(prefix="pharos_oracle_")))
    except Exception as e:  # no compiler on the host: the prebuilt x86-64-v3 library
        build = f"-O3 -march=x86-64-v3 (native build failed: {e})"
    tabs = O.Tables([(c.signal, c.prn) for c in scenario.channels])
    all_rate, all_epochs, all_busy = _oracle_rate(O, tabs, scenario, seconds, threads, scenario.epochs)
    one_rate, one_epochs, one_busy = _oracle_rate(O, tabs, scenario, single_seconds, 1, scenario.epochs)
    nch = len(scenario.channels)
    return {"value": all_rate, "unit": "channel-samples/s", "cores": threads, "kind": "port",
            "sample": f"first {all_epochs} epochs ({all_epochs / 1000:.1f} s of signal) of the same {nch}-channel run, "
                      f"C restatement of kernel.zig ({build}) on {threads} threads, constant gain, "
                      f"{all_busy:.1f} s of synthesis",
            "single_thread": {"value": one_rate, "epochs": one_epochs, "seconds": one_busy},
            "parallel_speedup": all_rate / one_rate, "build": build, "host": host,
            # Every core of the machine at the single-thread rate, an upper bound of an all-core run that the job's
            # CPU share cannot measure.
            "all_nproc_linear_bound": one_rate * host["nproc"]}


def precision(job):
    """pharos_synth_precision over every row of the job: largest carrier phase error of the f32 rotor against FP64
    and largest Q32.32 code position error, over the whole run."""
    import ctypes as C
    from pharos_amd import _lib
    sc = job.scenario
    nch = len(sc.channels)
    rep = _lib.PrecisionReport()
    per_table = (C.c_double * max(1, nch))()
    t0 = time.perf_counter()
    _lib.check(_lib.lib().pharos_synth_precision(job.engine.handle, job.rows.ptr, job.epochs * nch, C.byref(rep),
                                                 per_table, None))
    chips = 0.0
    for c, ch in enumerate(sc.channels):
        k = _lib.SignalConstants()
        _lib.lib().pharos_signal_info(_lib.SIGNAL[ch.signal], C.byref(k))
        chips = max(chips, per_table[c] / k.entries_per_chip)
    return {"carrier_max_err_cycles": rep.carrier_cycles, "carrier_max_err_rad": rep.carrier_radians,
            "rotor_magnitude_max_err": rep.magnitude, "code_max_err_entries": rep.code_entries,
            "code_max_err_chips": chips, "rows": rep.rows, "samples": job.epochs * nch * sc.samples_per_epoch,
            "bound_1e6_cycles_met": rep.carrier_cycles <= 1e-6, "seconds": time.perf_counter() - t0,
            "against": "FP64 phasor of phase0 + n*step and f64 entry_phase0 + n*eps, every sample of every row"}


def profiled(epochs):
    """(HBM bytes of one launch over `epochs` epochs, VALU busy %) from the batched kernel's counter runs: the
    per-epoch FETCH_SIZE + WRITE_SIZE, each scaled by the counter's reading on a copy of known size
    (profiles/synth_batch_kernel_c1_traffic.json)."""
    path = os.path.join(ROOT, "profiles", "synth_batch_kernel_c1_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    per_epoch = d.get("hbm_bytes_per_epoch")
    return (None if per_epoch is None else per_epoch * epochs), d.get("valu_busy_pct")


def _reset(job):
    job.agc.gain, job.agc.clipped, job.agc.components, job.agc.adjustments = 1.0, 0, 0, 0
    job.clipped.value = 0


def timed_batch(jobs, shape, stream):
    """(wall seconds, kernel seconds by HIP events on the launch stream) of one batched launch over `jobs` with the
    launch shape (workgroups per CU, channels per pass, pharos_synth_tuning flags)."""
    import ctypes as C
    from pharos_amd import _lib
    from pharos_amd.engine import Event
    L = _lib.lib()
    for j in jobs:
        _reset(j)
    arr = (_lib.SynthJob * len(jobs))(*[j.record() for j in jobs])
    tuning = _lib.SynthTuning(shape[0], shape[1], 0, shape[2])
    e0, e1 = Event(), Event()
    h = C.c_void_p()
    L.pharos_synth_device_sync(0)
    t0 = time.perf_counter()
    e0.record(stream)
    _lib.check(L.pharos_synth_batch_submit(arr, len(jobs), C.byref(tuning), stream.handle, C.byref(h)))
    e1.record(stream)
    _lib.check(L.pharos_synth_batch_wait(h))
    L.pharos_synth_device_sync(0)
    return time.perf_counter() - t0, e0.elapsed_ms(e1) / 1e3


def run_latency(runs):
    """p50 / p99 of the timed runs' wall times: one run from submit to its results on the host, inputs resident."""
    import numpy as np
    w = np.array([wall for wall, _ in runs]) * 1e3
    return {"p50_ms": float(np.percentile(w, 50)), "p99_ms": float(np.percentile(w, 99)), "runs": len(w),
            "form": "batched persistent launch"}


def timed_graphs(graphs, stream):
    from pharos_amd import _lib
    L = _lib.lib()
    for g in graphs:
        _reset(g.job)
    L.pharos_synth_device_sync(0)
    t0 = time.perf_counter()
    for g in graphs:
        g.launch(stream)
    for g in graphs:
        g.finish(stream)
    L.pharos_synth_device_sync(0)
    return time.perf_counter() - t0


def timed_plain(jobs, stream):
    from pharos_amd import _lib
    from pharos_amd.engine import plain_job
    L = _lib.lib()
    L.pharos_synth_device_sync(0)
    t0 = time.perf_counter()
    for j in jobs:
        plain_job(j, stream)
    L.pharos_synth_device_sync(0)
    return time.perf_counter() - t0


def launch_shape(args, scenarios, stream):
    """The batched launch shape: given on the command line, or chosen by the autotuner on a short sample of the
    same workload (pharos_amd/tune.py)."""
    if args.workgroups_per_cu or args.channels_per_pass or args.no_tune or args.wide_tiles:
        flags = tune.WIDE_TILES if args.wide_tiles else 0
        return (args.workgroups_per_cu, args.channels_per_pass, flags), {"source": "command line"}
    t = tune.tune(scenarios, sample_epochs=args.tune_epochs, stream=stream)
    w, c, f = t["best"]
    return (w, c, f), {"source": "autotuner", "key": t["key"], "best": {"workgroups_per_cu": w, "channels_per_pass": c,
                       "tile_samples": 8192 if f & tune.WIDE_TILES else 4096}, "sample_epochs": t["sample_epochs"],
                       "candidates_ms": t["times_ms"], "seconds": t["seconds"]}


WORKLOADS = {
    "c0": ("gps_l1ca_8sat_4.092Msps_iq8", "noise off", "constant 0 dB"),
    "c1": ("multi_gnss_40ch_61.44Msps_iq4", "counter", "adaptive (device AGC)"),
    "c2": ("e5_b2_l5_24ch_122.88Msps_iq16", "counter", "adaptive (device AGC)"),
}


def bench_single(args):
    """One run of the c0, c1 or c2 configuration per step."""
    from pharos_amd import scenario as S
    from pharos_amd.engine import Job, Graph, Stream
    make = {"c0": S.c0, "c1": S.c1, "c2": S.c2}[args.workload]
    sc = make(seed=1, epochs=args.epochs) if args.epochs else make(seed=1)
    name, noise, gain = WORKLOADS[args.workload]
    nch = len(sc.channels)
    rows, masks = sc.window(0, sc.epochs)
    job = Job(sc, rows=rows, masks=masks)
    stream = Stream()
    cs_per_step = sc.channel_samples
    flops_per_step = flops_per_epoch(sc) * sc.epochs
    shape, tuned = launch_shape(args, [sc], stream)

    for _ in range(args.warmup):
        timed_batch([job], shape, stream)
    runs = [timed_batch([job], shape, stream) for _ in range(args.steps)]
    ms_per_step = 1e3 * sum(w for w, _ in runs) / len(runs)
    kernel_s = sum(k for _, k in runs) / len(runs)
    value = cs_per_step / (ms_per_step / 1e3)

    graph_ms = None
    if args.graph_steps:
        g = Graph(job)
        timed_graphs([g], stream)
        graph_ms = 1e3 * sum(timed_graphs([g], stream) for _ in range(args.graph_steps)) / args.graph_steps
        g.close()
    plain_t = [timed_plain([job], stream) for _ in range(args.plain_steps)]
    plain_value = cs_per_step / (sum(plain_t) / len(plain_t)) if plain_t else None

    achieved = flops_per_step / kernel_s / 1e12
    traffic, valu_busy = profiled(sc.epochs) if args.workload == "c1" else (None, None)
    out_bytes = sc.epochs * sc.bytes_per_epoch
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "channel-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded channel models; rows, masks and code tables resident in HBM)",
        "config": {"workload": name, "channels": nch, "sample_rate_hz": sc.samples_per_epoch * 1000,
                   "format": sc.format, "epochs": sc.epochs, "noise": noise, "gain": gain,
                   "form": "batched persistent launch", "channel_samples_per_step": cs_per_step},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "valu_busy_pct_profiled": valu_busy,
                     "kernel": "synth_batch_kernel", "kernel_ms": kernel_s * 1e3,
                     "kernel_ms_spans": "row_prep_kernel pre-pass + synth_batch_kernel + batch-table copies, HIP events "
                                        "on the launch stream",
                     "flops_per_launch": flops_per_step, "hbm_write_gbs": out_bytes / kernel_s / 1e9,
                     "hbm_write_frac": out_bytes / kernel_s / 1e9 / PEAK_HBM_GBS},
        "graph": {"value": cs_per_step / (graph_ms / 1e3) if graph_ms else None, "unit": "channel-samples/s",
                  "ms_per_step": graph_ms,
                  "form": "per-window launches with the AGC step on the device, one HIP graph launch per run"},
        "plain": {"value": plain_value, "unit": "channel-samples/s",
                  "ms_per_step": 1e3 * sum(plain_t) / len(plain_t) if plain_t else None,
                  "form": "per-window launches (100 epochs), host AGC between windows"},
        "batched_over_plain": value / plain_value if plain_value else None,
        "tuning": tuned,
        "run_latency": run_latency(runs),
    }
    if not args.no_precision:
        line["precision"] = precision(job)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sc, args.cpu_seconds)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
        line["gpu_over_cpu_single_thread"] = value / line["cpu_baseline"]["single_thread"]["value"]
        line["gpu_over_cpu_nproc_bound"] = value / line["cpu_baseline"]["all_nproc_linear_bound"]
        line["target_50x_all_core_met"] = line["gpu_over_cpu"] >= 50.0
        line["target_50x_nproc_bound_met"] = line["gpu_over_cpu_nproc_bound"] >= 50.0
    return line


def bench_sink(args):
    """The file sink (pharos_synth_to_file) over the first --sink-epochs epochs of the same run: the rate with the
    device-to-host copies and the file writes in, which the device-resident `value` leaves out."""
    from pharos_amd import scenario as S
    from pharos_amd.engine import Job
    make = {"c0": S.c0, "c1": S.c1, "c2": S.c2}[args.workload]
    sc = make(seed=1, epochs=args.sink_epochs)
    rows, masks = sc.window(0, sc.epochs)
    job = Job(sc, rows=rows, masks=masks)
    path = os.path.join(args.sink, f"pharos_sink_{os.getpid()}.bin")
    try:
        job.to_file(path)  # warm-up: the pinned and device rings, the file
        t0 = time.perf_counter()
        n = job.to_file(path)
        dt = time.perf_counter() - t0
    finally:
        if os.path.exists(path):
            os.remove(path)
    return {"value": sc.channel_samples / dt, "unit": "channel-samples/s", "bytes": n, "gb_per_s": n / dt / 1e9,
            "epochs": sc.epochs, "seconds": dt,
            "form": "100-epoch plain blocks, device-to-pinned copies and file writes overlapped with synthesis"}


def bench_c3(args):
    from concurrent.futures import ThreadPoolExecutor
    from pharos_amd import scenario as S
    from pharos_amd.engine import Job, Graph, Stream
    cap = args.c3_epochs_cap or None
    scs = S.c3(seed=1, epochs_cap=cap)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as pool:
        inputs = list(pool.map(lambda s: s.window(0, s.epochs), scs))
    jobs = [Job(s, rows=r, masks=m) for s, (r, m) in zip(scs, inputs)]
    del inputs
    stream = Stream()
    cs = sum(s.channel_samples for s in scs)
    flops = sum(flops_per_epoch(s) * s.epochs for s in scs)
    shape, tuned = launch_shape(args, scs, stream)
    timed_batch(jobs, shape, stream)
    runs = [timed_batch(jobs, shape, stream) for _ in range(args.steps)]
    wall = sum(w for w, _ in runs) / len(runs)
    kern = sum(k for _, k in runs) / len(runs)
    graphs = [Graph(j) for j in jobs] if args.graph_steps else []
    graph_s = None
    if graphs:
        timed_graphs(graphs, stream)
        graph_s = sum(timed_graphs(graphs, stream) for _ in range(args.graph_steps)) / args.graph_steps
    plain_s = timed_plain(jobs, stream) if args.plain_steps else None
    fmts = {}
    for s in scs:
        fmts[s.format] = fmts.get(s.format, 0) + 1
    return {
        "metric": METRIC, "value": cs / wall, "unit": "channel-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (256 seeded scenarios; rows, masks and tables resident in HBM)",
        "config": {"workload": "uneven_256_scenarios", "scenarios": len(scs), "epochs_cap": cap,
                   "channel_samples_per_step": cs, "output_bytes": sum(s.epochs * s.bytes_per_epoch for s in scs),
                   "formats": fmts, "form": "one batched persistent launch (each job's quantiser picked per tile)"},
        "roofline": {"bound": "valu", "achieved": flops / kern / 1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": flops / kern / 1e12 / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "synth_batch_kernel",
                     "kernel_ms": kern * 1e3},
        "graph": {"value": cs / graph_s if graph_s else None, "ms_per_step": graph_s * 1e3 if graph_s else None,
                  "form": "one HIP graph per scenario (per-window launches, device AGC), replayed in sequence"},
        "plain": {"value": cs / plain_s if plain_s else None, "ms_per_step": plain_s * 1e3 if plain_s else None,
                  "form": "per-scenario per-window launches with the host AGC"},
        "batched_over_plain": (plain_s / wall) if plain_s else None,
        "tuning": tuned,
        "run_latency": run_latency(runs),
    }


def bench_run_api(args):
    """The configuration through pharos_run.h (a C caller's path): the first --run-api-epochs epochs pulled into a
    host buffer in 100-epoch calls, then written with pharos_run_to_file to a temporary file; the rates include the
    host stage, the device derive and synthesis, and the copy to host memory (PCIe-bound)."""
    import ctypes as C
    import tempfile
    from pharos_amd import _lib
    from pharos_amd.run import Run
    from pharos_amd import scenario as S
    make = {"c0": S.c0, "c1": S.c1, "c2": S.c2}[args.workload]
    sc = make(seed=1, epochs=args.run_api_epochs)
    L = _lib.lib()
    gain = 1 if sc.gain_db is not None else 0
    run = Run(sc.bundle(), seed=sc.seed, gain=gain, gain_db=sc.gain_db or 0.0)
    bpe = run.info().bytes_per_epoch
    buf = np.empty(100 * bpe, dtype=np.uint8)
    blk = _lib.RunBlock()
    L.pharos_run_pull(run.handle, buf.ctypes.data, buf.nbytes, 100, C.byref(blk))   # untimed: first block
    t = time.perf_counter()
    n = 0
    while L.pharos_run_pull(run.handle, buf.ctypes.data, buf.nbytes, 100, C.byref(blk)) == 0:
        n += blk.epochs
    pull_s = time.perf_counter() - t
    per_epoch = sc.samples_per_epoch * len(sc.channels)
    run = Run(sc.bundle(), seed=sc.seed, gain=gain, gain_db=sc.gain_db or 0.0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "run.bin")
        written = C.c_uint64()
        t = time.perf_counter()
        st = L.pharos_run_to_file(run.handle, path.encode(), C.byref(written))
        file_s = time.perf_counter() - t
    return {"epochs": sc.epochs, "pull": {"epochs": n, "gb_per_s": n * bpe / pull_s / 1e9,
                                          "channel_samples_per_s": n * per_epoch / pull_s},
            "to_file": {"status": st, "bytes": written.value, "gb_per_s": written.value / file_s / 1e9,
                        "channel_samples_per_s": sc.epochs * per_epoch / file_s},
            "note": "pharos_run.h from a host caller: host stage, device derive and synthesis, copy to pageable host "
                    "memory; to_file into the page cache of a temporary file"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c0", "c1", "c2", "c3"], default="c1",
                    help="c1: the multi-GNSS 40-channel run (default); c0: 8 x L1 C/A at 4.092 Msps; c2: the "
                         "wideband 24-channel 122.88 Msps run; c3: the uneven 256-scenario batch")
    ap.add_argument("--epochs", type=int, default=0, help="epochs per run (0: the configuration's own)")
    ap.add_argument("--c3-epochs-cap", type=int, default=3000, help="longest scenario of the uneven batch (0: none)")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--graph-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the whole-run precision report")
    ap.add_argument("--serve", action="store_true", help="run the Poisson serving measurement (the default)")
    ap.add_argument("--no-serve", action="store_true", help="skip the Poisson serving measurement")
    ap.add_argument("--serve-rate", type=float, default=100.0, help="requests per second")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="arrival window")
    ap.add_argument("--serve-epochs-cap", type=int, default=0, help="longest request in epochs (0: up to 30 s)")
    ap.add_argument("--no-run-api", action="store_true", help="skip the pharos_run.h pull and to_file rates")
    ap.add_argument("--run-api-epochs", type=int, default=3000, help="epochs the run API pulls and writes")
    ap.add_argument("--sink", metavar="DIR", default="",
                    help="also time the file sink: the first --sink-epochs epochs of the run written to a file in DIR")
    ap.add_argument("--sink-epochs", type=int, default=6000)
    ap.add_argument("--workgroups-per-cu", type=int, default=0, help="launch shape (skips the autotuner)")
    ap.add_argument("--channels-per-pass", type=int, default=0, help="launch shape (skips the autotuner)")
    ap.add_argument("--no-tune", action="store_true", help="use the library's default launch shape")
    ap.add_argument("--wide-tiles", action="store_true", help="launch shape with 8192-sample tiles (skips the autotuner)")
    ap.add_argument("--tune-epochs", type=int, default=300, help="epochs of the autotuner's sample")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device three times (a known-size reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (runs shard by scenario or epoch "
              f"range with no exchange; the adaptive gain's one scalar per 100-epoch window is all they would "
              f"share)", file=sys.stderr)
        sys.exit(2)

    from pharos_amd.engine import DeviceBuffer, require_gpu
    from pharos_amd import _lib
    require_gpu()
    if args.copy_calibration:
        a, b = DeviceBuffer(args.copy_calibration), DeviceBuffer(args.copy_calibration)
        for _ in range(3):
            _lib.check(_lib.lib().pharos_synth_memcpy(b.ptr, a.ptr, args.copy_calibration, 2))
        _lib.lib().pharos_synth_device_sync(0)
        print(json.dumps({"copied_bytes": args.copy_calibration, "copies": 3}), flush=True)
        return

    line = bench_c3(args) if args.workload == "c3" else bench_single(args)
    if args.sink and args.workload != "c3":
        line["sink"] = bench_sink(args)
    if not args.no_run_api and args.workload != "c3":
        line["run_api"] = bench_run_api(args)
    if not args.no_serve:
        from pharos_amd import serve
        line["serving"] = serve.measure(duration_s=args.serve_seconds, rate_hz=args.serve_rate,
                                        epochs_cap=args.serve_epochs_cap or None)
        # Latency to the last byte in host memory, at a rate whose output volume the host link can carry (the
        # output of 100 req/s averages about 64 GB/s, more than a PCIe x16 link moves).
        line["serving_to_host"] = serve.measure(duration_s=5.0, rate_hz=20.0, epochs_cap=args.serve_epochs_cap or None,
                                                to_host=True)
        if "cpu_baseline" in line:
            line["serving"]["over_cpu"] = line["serving"]["channel_samples_per_s"] / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
