"""Benchmark of the MI355X IF sample engine on the multi-GNSS configuration (one GPU).

A step synthesises the whole run of the configuration BASELINE.json's metric is quoted on: 40 channels of
GPS L1C/A+L5, Galileo E1+E5a, BeiDou B1I+B3I and GLONASS L1OF at 61.44 Msps, IQ4, 60 s (60 000 epochs), noise on,
adaptive AGC -- rows and masks resident in HBM before the timed region, the 3.69 GB of IQ4 output written to HBM.

  value        channel-samples/s of the batched form (one persistent launch per step, device AGC)
  plain        the same run as per-window plain launches with the host AGC in between
  roofline     FP32 VALU roofline of the synthesis kernel: algorithmic flops per launch / launch time
  cpu_baseline the C oracle (a restatement of the reference kernel) on the host cores, on a bounded sample
  serving      p50/p99 latency of a Poisson request stream (--serve)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "channel-samples/sec + roofline %, batched vs plain on 1 GPU; p50/p99 run latency"
PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 (vector, packed)
PEAK_HBM_GBS = 8000.0
# Algorithmic FP32 flops per channel-sample (kernel.zig:241-258 plus rotor.zig:52-58): a complex MAC and a rotor
# step for two-axis signals (8 + 6), two FMAs and a rotor step for one-axis signals (4 + 6).
FLOPS_TWO_AXIS, FLOPS_ONE_AXIS = 14, 10


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


def cpu_baseline(scenario, seconds=10.0, threads=None):
    """The oracle over 100-epoch windows of the same workload until `seconds` of wall time have passed."""
    from oracle import binding as O
    threads = threads or min(16, os.cpu_count() or 1)
    tabs = O.Tables([(c.signal, c.prn) for c in scenario.channels])
    nch = len(scenario.channels)
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds and done < scenario.epochs:
        rows, masks = scenario.window(done, 100)
        O.synth_window(tabs, rows, masks, scenario.samples_per_epoch, scenario.format_id, 0 if scenario.noise else 1,
                       scenario.seed, done, 1.0, threads=threads)
        done += 100
    dt = time.perf_counter() - t0
    return {"value": nch * scenario.samples_per_epoch * done / dt, "unit": "channel-samples/s", "cores": threads,
            "kind": "port",
            "sample": f"first {done} epochs ({done / 1000:.1f} s of signal) of the same 40-channel run, "
                      f"C restatement of kernel.zig on {threads} threads, constant gain, {dt:.1f} s wall"}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--epochs", type=int, default=60_000, help="epochs per run (the configuration has 60 000)")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", action="store_true", help="also run the Poisson serving measurement")
    ap.add_argument("--workgroups-per-cu", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device three times (a known-size reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (runs shard by scenario or epoch "
              f"range with no exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)

    from pharos_amd import scenario as S
    from pharos_amd.engine import Job, run_batch, DeviceBuffer, Engine, Event, Stream, require_gpu
    from pharos_amd import _lib
    import ctypes as C
    require_gpu()
    if args.copy_calibration:
        a, b = DeviceBuffer(args.copy_calibration), DeviceBuffer(args.copy_calibration)
        for _ in range(3):
            _lib.check(_lib.lib().pharos_synth_memcpy(b.ptr, a.ptr, args.copy_calibration, 2))
        _lib.lib().pharos_synth_device_sync(0)
        print(json.dumps({"copied_bytes": args.copy_calibration, "copies": 3}), flush=True)
        return

    sc = S.c1(seed=1, epochs=args.epochs)
    nch = len(sc.channels)
    rows, masks = sc.window(0, sc.epochs)
    job = Job(sc, rows=rows, masks=masks)
    stream = Stream()
    cs_per_step = sc.channel_samples
    flops_per_step = flops_per_epoch(sc) * sc.epochs

    def batched_step():
        job.agc.gain, job.agc.clipped, job.agc.components, job.agc.adjustments = 1.0, 0, 0, 0
        run_batch([job], args.workgroups_per_cu, stream)

    for _ in range(args.warmup):
        batched_step()
    # Kernel time from HIP events on the launch stream, beside the host wall clock of the whole step.
    L = _lib.lib()
    t_wall = []
    t_kernel = []
    for _ in range(args.steps):
        e0, e1 = Event(), Event()
        L.pharos_synth_device_sync(0)
        t0 = time.perf_counter()
        job.agc.gain, job.agc.clipped, job.agc.components, job.agc.adjustments = 1.0, 0, 0, 0
        arr = (_lib.SynthJob * 1)(job.record())
        tuning = _lib.SynthTuning(args.workgroups_per_cu, (C.c_uint32 * 3)())
        h = C.c_void_p()
        e0.record(stream)
        _lib.check(L.pharos_synth_batch_submit(arr, 1, C.byref(tuning), stream.handle, C.byref(h)))
        e1.record(stream)
        _lib.check(L.pharos_synth_batch_wait(h))
        L.pharos_synth_device_sync(0)
        t_wall.append(time.perf_counter() - t0)
        t_kernel.append(e0.elapsed_ms(e1) / 1e3)
    ms_per_step = 1e3 * sum(t_wall) / len(t_wall)
    kernel_s = sum(t_kernel) / len(t_kernel)
    value = cs_per_step / (ms_per_step / 1e3)

    # Plain form: per-window launches, host AGC between windows (a sync per window).
    eng = job.engine
    plain_t = []
    for _ in range(args.plain_steps):
        L.pharos_synth_device_sync(0)
        t0 = time.perf_counter()
        gain, acc_c, acc_n = 1.0, 0, 0
        for e in range(0, sc.epochs, 100):
            n = min(100, sc.epochs - e)
            c = eng.window(e, n, job.rows.at(e * nch * 64), job.masks.ptr, job.mask_len, float(np.float32(gain)),
                           job.out.at(e * eng.bytes_per_epoch), clips=True, stream=stream)
            acc_c += int(c.sum())
            acc_n += n * sc.samples_per_epoch * 2
            if (e + n) % 100 == 0:
                rate = acc_c / acc_n
                acc_c = acc_n = 0
                if rate > 0.01:
                    gain *= 0.95
                elif rate < 0.001 and gain < 1.0:
                    gain = min(gain * 1.02, 1.0)
        L.pharos_synth_device_sync(0)
        plain_t.append(time.perf_counter() - t0)
    plain_value = cs_per_step / (sum(plain_t) / len(plain_t)) if plain_t else None

    achieved = flops_per_step / kernel_s / 1e12
    traffic, valu_busy = profiled(sc.epochs)
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
        "config": {"workload": "multi_gnss_40ch_61.44Msps_iq4", "channels": nch, "sample_rate_hz": 61_440_000,
                   "format": "IQ4", "epochs": sc.epochs, "noise": "counter", "gain": "adaptive (device AGC)",
                   "form": "batched persistent launch", "channel_samples_per_step": cs_per_step},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "valu_busy_pct_profiled": valu_busy,
                     "kernel": "synth_batch_kernel", "kernel_ms": kernel_s * 1e3,
                     "flops_per_launch": flops_per_step, "hbm_write_gbs": out_bytes / kernel_s / 1e9,
                     "hbm_write_frac": out_bytes / kernel_s / 1e9 / PEAK_HBM_GBS},
        "plain": {"value": plain_value, "unit": "channel-samples/s",
                  "ms_per_step": 1e3 * sum(plain_t) / len(plain_t) if plain_t else None,
                  "form": "per-window launches (100 epochs), host AGC between windows"},
        "batched_over_plain": value / plain_value if plain_value else None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sc, args.cpu_seconds)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    if args.serve:
        from pharos_amd import serve
        line["serving"] = serve.measure(duration_s=20.0)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
