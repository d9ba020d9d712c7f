"""The uneven-batch and serving configurations at their own sizes.

Uneven batch: all 256 seeded scenarios (4-48 channels, 4.092-61.44 Msps, IQ2-IQ16, 1-30 s each, 2.95e12
channel-samples, 197 GB of output) in one persistent launch (in slices of jobs only if the device's free memory
cannot hold them all).  Each job's output, clip total and AGC trace equal that job's plain run (per-window
launches, host AGC), compared through the device checksum; sampled epochs of a few jobs equal the oracle's bytes.

Serving: a Poisson stream at 200 requests/s for 2 s, requests drawn the same way with up to 2 GB of output each
and no cap on their length; every eighth request's output, clip total and AGC trace equal the same run alone.
"""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import binding as O
from pharos_amd import _lib, scenario as S, serve
from pharos_amd.engine import Job, digest, plain_job, run_batch

pytestmark = pytest.mark.gpu

PREP_BYTES = 272


def _device_free():
    free, total = C.c_uint64(), C.c_uint64()
    _lib.check(_lib.lib().pharos_synth_mem_info(0, C.byref(free), C.byref(total)))
    return free.value


def _need(sc):
    rows = sc.epochs * len(sc.channels)
    prep = rows * PREP_BYTES if sc.samples_per_epoch > 4096 else 0
    return sc.epochs * sc.bytes_per_epoch + rows * 64 + prep + (4 << 20)


def _slices(scs, budget):
    out, cur, used = [], [], 0
    for i, sc in enumerate(scs):
        n = _need(sc)
        if cur and used + n > budget:
            out.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += n
    return out + ([cur] if cur else [])


def _gain_at(adjustments, epoch):
    """The gain windows use at `epoch`: the last AGC change at or before it (changes land on 100-epoch edges)."""
    g = 1.0
    for e, gain, _ in adjustments:
        if e <= epoch:
            g = gain
    return float(np.float32(g))


def test_uneven_batch_at_full_size():
    scs = S.c3(seed=1)
    assert len(scs) == 256 and max(s.epochs for s in scs) == 30_000
    total = sum(_need(s) for s in scs)
    budget = _device_free() - (12 << 30)
    groups = _slices(scs, budget)
    print(f"c3 full: {sum(s.channel_samples for s in scs):.3e} channel-samples, {total / 1e9:.1f} GB needed, "
          f"{budget / 1e9:.1f} GB budget, {len(groups)} launch(es)", flush=True)
    threads = min(16, os.cpu_count() or 4)
    sampled = {0, 77, 131, 200}
    for g in groups:
        with ThreadPoolExecutor(max_workers=threads) as pool:
            inputs = list(pool.map(lambda i: scs[i].window(0, scs[i].epochs), g))
        jobs = [Job(scs[i], rows=r, masks=m, adjustments_cap=4096) for i, (r, m) in zip(g, inputs)]
        del inputs
        print(f"c3 full: batched launch of {len(jobs)} jobs", flush=True)
        run_batch(jobs)
        batched = [(digest(j.out.ptr, j.out.nbytes), j.clipped.value, j.adjustments(), j.agc.gain) for j in jobs]
        # Sampled epochs against the oracle: first, middle and last epoch of a few jobs, at the gain the device AGC
        # gave their windows.
        for i, j in zip(g, jobs):
            if i not in sampled:
                continue
            sc = scs[i]
            tabs = O.Tables([(c.signal, c.prn) for c in sc.channels])
            for e in (0, sc.epochs // 2 + 37, sc.epochs - 1):
                rows, masks = sc.window(e, 1)
                want, _ = O.synth_window(tabs, rows, masks, sc.samples_per_epoch, sc.format_id, 0, sc.seed, e,
                                         _gain_at(batched[g.index(i)][2], e), threads=threads)
                got = j.out.to_numpy(sc.bytes_per_epoch, e * sc.bytes_per_epoch)
                assert np.array_equal(got, want), (sc.name, e)
        print("c3 full: plain runs, job by job", flush=True)
        for i, j, (d, clips, adj, gain) in zip(g, jobs, batched):
            total_clips, pgain, padj = plain_job(j)     # overwrites the job's output with the plain form's
            assert digest(j.out.ptr, j.out.nbytes) == d, scs[i].name
            assert total_clips == clips and padj == adj and pgain == gain, scs[i].name
        del jobs


def test_serving_at_200_requests_per_second():
    reqs = serve.make_requests(rate_hz=200.0, duration_s=2.0, seed=11)
    print(f"c4: {len(reqs)} requests, {sum(r.scenario.channel_samples for r in reqs):.3e} channel-samples, "
          f"largest output {max(r.out_bytes for r in reqs) / 1e9:.2f} GB", flush=True)
    assert len(reqs) >= 300 and max(r.out_bytes for r in reqs) <= 2 << 30
    seen = {}

    def record(r):
        if r.id % 8 == 0:
            seen[r.id] = (digest(r.out.ptr, r.out_bytes), r.clipped_total, r.adjustments(), r.agc.gain)

    server = serve.Server()
    done = server.serve(reqs, on_complete=record, to_host=True)
    server.close()
    assert len(done) == len(reqs) and max(server.batch_sizes) >= 8
    lat = np.array([r.completed - r.arrival_abs for r in done])
    host = np.array([r.host_done - r.arrival_abs for r in done])
    assert (host >= lat).all()
    print(f"c4: p50 {np.percentile(lat, 50) * 1e3:.1f} ms, p99 {np.percentile(lat, 99) * 1e3:.1f} ms to device memory, "
          f"p50 {np.percentile(host, 50) * 1e3:.1f} ms, p99 {np.percentile(host, 99) * 1e3:.1f} ms to host memory, "
          f"{len(seen)} requests checked alone", flush=True)
    for r in reqs:
        if r.id not in seen:
            continue
        (alone,) = run_batch([Job(r.scenario, adjustments_cap=4096, derive=True)])
        d, clips, adj, gain = seen[r.id]
        assert digest(alone.out.ptr, alone.out.nbytes) == d, r.scenario.name
        assert alone.clipped.value == clips and alone.adjustments() == adj and alone.agc.gain == gain
