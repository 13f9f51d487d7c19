"""Simulated decoder streams driving the serving loop (vaco_av1pf_server).

``serve_poisson`` feeds one server from many streams with Poisson arrivals,
as a decoder farm would: each stream has a fixed resolution and bit depth
(``workloads.serve_streams``), its frames cycle over a few seeded frames
whose planes stay resident on the device, and its outputs over a ring of
device buffers (a stream waits when the buffer its next frame would write is
still in flight). Latency is per frame, from submit to the server seeing its
batch complete. A sample of requests write to buffers of their own, kept
for checking; the per-frame path's output of every distinct frame is kept
beside them.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import av1pf, hip, workloads


@dataclass
class ServeRun:
    stats: dict
    frames: list  # frames[stream][k]
    checked: dict = field(default_factory=dict)  # request -> (stream, k, [output planes])
    reference: dict = field(default_factory=dict)  # (stream, k) -> per-frame path's output planes


def serve_poisson(seconds: float, fps: float = 3000.0, streams: int = 32, distinct: int = 2, ring: int = 8,
                  budget: int = 32 << 30, window_us: int = 2000, max_inflight: int = 2, max_batch: int = 64,
                  max_checks: int = 120) -> ServeRun:
    """Run ``fps * seconds`` requests; every max(100, n / max_checks)-th one
    (1% of a short run) is kept for checking."""
    cfgs = workloads.serve_streams(streams)
    frames = [[av1pf.generate(c, k) for k in range(distinct)] for c in cfgs]
    dev_in = [[[hip.DeviceArray.from_host(p) for p in f.planes] for f in fs] for fs in frames]
    outs = [[[hip.DeviceArray(g) for g in fs[0].output_geometry()] for _ in range(ring)] for fs in frames]
    ptrs = lambda arrs: [t.ptr for t in arrs]

    rng = np.random.default_rng(workloads.BASE_SEED + 4)
    n = int(fps * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / fps, size=n))
    which = rng.integers(0, streams, size=n)
    check_every = max(100, n // max_checks)
    checked = {i: [hip.DeviceArray(g) for g in frames[which[i]][0].output_geometry()]
               for i in range(0, n, check_every)}

    ctx = av1pf.Context(3840, 2176)
    reference = {}
    for s in range(streams):
        for k in range(distinct):
            o = [hip.DeviceArray(g) for g in frames[s][k].output_geometry()]
            ctx.frame_device(frames[s][k], ptrs(dev_in[s][k]), ptrs(o))
            reference[(s, k)] = o
    ctx.finish()
    hip.synchronize()

    srv = av1pf.Server(ctx, hbm_budget_bytes=budget, window_us=window_us, max_batch=max_batch,
                       max_inflight=max_inflight)
    # submit callables per (stream, frame, output buffer), built before the clock starts
    submit = {(s, k, slot): srv.submit_call(frames[s][k], ptrs(dev_in[s][k]), ptrs(outs[s][slot]))
              for s in range(streams) for k in range(distinct) for slot in range(ring)}
    per_stream_seq = np.zeros(streams, np.int64)
    ring_owner = [[-1] * ring for _ in range(streams)]
    done_t = {}
    latency, batch_sizes, lag = [], [], []
    device_spans = set()  # (launch, device start, device end) per batch
    mpix = 0.0
    frame_of = np.zeros(n, np.int64)

    def drain_completions():
        nonlocal mpix
        for tag, st, bs, t_sub, t_launch, t_done, g0, g1 in srv.poll():
            if st != 0:
                raise RuntimeError(f"served frame {tag} failed with status {st}")
            done_t[tag] = t_done
            device_spans.add((t_launch, g0, g1))
            latency.append((t_done - t_sub) / 1e6)
            batch_sizes.append(bs)
            mpix += frames[which[tag]][0].output_luma_pixels / 1e6

    t0 = av1pf.now_ns()
    for i in range(n):
        s = int(which[i])
        seq = int(per_stream_seq[s])
        per_stream_seq[s] += 1
        slot = seq % ring
        prev = ring_owner[s][slot]
        while prev >= 0 and prev not in done_t:  # ring buffer still in use
            drain_completions()
        ring_owner[s][slot] = i
        due = t0 + int(arrivals[i] * 1e9)
        while av1pf.now_ns() < due:
            drain_completions()
        lag.append((av1pf.now_ns() - due) / 1e6)
        k = frame_of[i] = seq % distinct
        if i in checked:
            srv.submit(frames[s][k], ptrs(dev_in[s][k]), ptrs(checked[i]), tag=i)
        else:
            submit[(s, k, slot)](i)
    srv.drain()
    drain_completions()
    t1 = max(done_t.values())
    srv.close()
    wall = (t1 - t0) / 1e9

    run = ServeRun(stats={}, frames=frames)
    for i, o in checked.items():
        run.checked[i] = (int(which[i]), int(frame_of[i]), [t.download() for t in o])
    for key, o in reference.items():
        run.reference[key] = [t.download() for t in o]
    ctx.close()
    lat = np.array(latency)
    # Batches whose device interval intersects an earlier-started batch's.
    spans = sorted((g0, g1) for _, g0, g1 in device_spans if g1 > g0)
    overlapped, reach = 0, -1
    for g0, g1 in spans:
        overlapped += g0 < reach
        reach = max(reach, g1)
    run.stats = {
        "workload": f"{streams} streams (720p/1080p/2160p x 8/10-bit 4:2:0, fixed per stream), Poisson "
                    f"arrivals at {fps:.0f} frames/s aggregate for {seconds:.1f} s; {budget >> 30} GiB HBM "
                    f"budget, {window_us / 1000:.1f} ms batching window, <= {max_inflight} batches in flight, "
                    f"each on its own stream",
        "frames": len(lat),
        "served_mpix_s": round(mpix / wall, 2),
        "offered_mpix_s": round(sum(frames[w][0].output_luma_pixels for w in which) / 1e6 / arrivals[-1], 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "max_ms": round(float(lat.max()), 3),
        "mean_batch": round(float(np.mean(batch_sizes)), 2),
        "batches": len(spans),
        "batches_overlapping_another_on_device": overlapped,
        "submit_lag_p99_ms": round(float(np.percentile(lag, 99)), 3),
        "spot_checks": len(checked),
    }
    return run
