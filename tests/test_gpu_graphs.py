"""Graph replay on the GPU: with graphs enabled, every call - the first,
which captures, and the replays after it - gives the oracle's output bit
for bit. Replays alternate with eager calls and synchronisations, as graph
replays in a real program would.
"""
import sys

import numpy as np
import pytest

import frames as F
from oracle import oracle as O
from vaco_amd import av1pf, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ptrs(arrs):
    return [t.ptr for t in arrs]


def test_gpu_graph_replays_match_the_oracle_every_time():
    from vaco_amd import hip
    frames = [av1pf.generate(workloads.cfg0(), k) for k in range(3)]
    frames.append(F.random_frame(seed=301, coded_w=93, coded_h=75, upscaled_w=160))
    frames.append(F.random_frame(seed=302, coded_w=131, coded_h=97, bd=10))
    want = [O.frame(f.desc, f.planes) for f in frames]
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    dev_out = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames]
    with av1pf.Context(1024, 576) as ctx:
        for rep in range(12):  # > 8 staging slots: every slot captures, then replays
            ctx.use_graphs(True)
            for k, f in enumerate(frames):
                for t in dev_out[k]:
                    t.zero()
                ctx.frame_device(f, ptrs(dev_in[k]), ptrs(dev_out[k]))
            ctx.finish()
            for k, w in enumerate(want):
                for p, plane in enumerate(w):
                    assert np.array_equal(dev_out[k][p].download(), plane), ("graph", rep, k, p)
            # eager work and a synchronisation between replays
            ctx.use_graphs(False)
            ctx.frame_device(frames[rep % len(frames)], ptrs(dev_in[rep % len(frames)]),
                             ptrs(dev_out[rep % len(frames)]))
            ctx.finish()
            hip.synchronize()
        say("per-frame graph replays exact")
        ctx.use_graphs(True)
        for rep in range(10):
            for ts in dev_out:
                for t in ts:
                    t.zero()
            ctx.batch_device(frames, [ptrs(x) for x in dev_in], [ptrs(x) for x in dev_out])
            ctx.finish()
            for k, w in enumerate(want):
                for p, plane in enumerate(w):
                    assert np.array_equal(dev_out[k][p].download(), plane), ("batch graph", rep, k, p)
        say("batched graph replays exact")
