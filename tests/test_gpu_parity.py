"""GPU parity: the HIP path against the reference's dav1d golden vectors and
against the oracle on seeded frames. Bit-exact (tolerance 0) everywhere:
the stage is integer arithmetic.
"""
import sys

import numpy as np
import pytest

import frames as F
import golden_inputs as G
from oracle import oracle as O
from vaco_amd import abi, av1pf, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = av1pf.Context(3840, 2176)
    yield c
    c.close()


def assert_frame_equal(got, want, label):
    assert len(got) == len(want), label
    for p, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, (label, p, g.shape, w.shape)
        if not np.array_equal(g, w):
            ys, xs = np.nonzero(g != w)
            raise AssertionError(f"{label} plane {p}: {len(ys)} samples differ, first at y={ys[0]} x={xs[0]} "
                                 f"got {g[ys[0], xs[0]]} want {w[ys[0], xs[0]]}")


# ------------------------------------------------------ golden vectors ---

def test_gpu_cdef_blocks_match_dav1d_corpus(ctx):
    say("cdef corpus: 7680 cases")
    inputs, params, exp_dir, exp_var, exp_out = G.cdef_cases()
    dirs, vars_, outs = ctx.cdef_blocks(inputs, params)
    assert np.array_equal(dirs, exp_dir.astype(np.uint8))
    assert np.array_equal(vars_, exp_var)
    assert np.array_equal(outs, exp_out)


def test_gpu_superres_matches_dav1d_corpus(ctx):
    for case_id, (bd, vis, padded, out_w, h, inp, exp) in enumerate(G.superres_cases()):
        say(f"superres case {case_id} bd={bd} {vis}/{padded}->{out_w}x{h}")
        got = ctx.upscale_plane(inp, vis, out_w, h, bd)
        assert np.array_equal(got, exp), f"superres case {case_id}"


def test_gpu_restoration_matches_dav1d_corpus(ctx):
    n = 0
    for (bd, w, h, ssy, unit, mode, before, after, units, cfg, exp) in G.restoration_cases():
        if mode == 0:
            say(f"restoration bd={bd} {w}x{h} unit={unit}")
        got = ctx.restore_plane(before, after, cfg, units)
        assert np.array_equal(got, exp), f"depth={bd} {w}x{h} mode={mode}"
        n += 1
    assert n == 300


@pytest.mark.parametrize("rows,lr_rows,lr_order", [(4, 32, 0), (16, 16, 1), (8, 64, 1), (4, 16, 0), (16, 32, 1)])
def test_gpu_plane_corpora_under_each_launch_variant(ctx, rows, lr_rows, lr_order):
    """The superres and restoration corpora through every launch variant the
    tuner can pick (rows per superres workgroup, restoration tile height and
    tile order), not only the defaults."""
    say(f"corpora with superres_rows={rows} lr_tile_rows={lr_rows} lr_order={lr_order}")
    saved = ctx.tuning
    ctx.tuning = {"superres_rows": rows, "lr_tile_rows": lr_rows, "lr_order": lr_order}
    try:
        for case_id, (bd, vis, padded, out_w, h, inp, exp) in enumerate(G.superres_cases()):
            assert np.array_equal(ctx.upscale_plane(inp, vis, out_w, h, bd), exp), f"superres case {case_id}"
        for (bd, w, h, ssy, unit, mode, before, after, units, cfg, exp) in G.restoration_cases():
            got = ctx.restore_plane(before, after, cfg, units)
            assert np.array_equal(got, exp), f"depth={bd} {w}x{h} mode={mode}"
        for bd, s, plane, exp in G.constant_cases():
            cfg = abi.LrPlaneConfig(7, 9, 64, bd, 0)
            sgr = abi.make_unit(abi.UNIT_SELF_GUIDED, set=s, xqd=(-32, 31))
            assert np.array_equal(ctx.restore_plane(plane, plane, cfg, [sgr]), exp), (bd, s)
    finally:
        ctx.tuning = saved


def test_gpu_superres_plane_wider_than_65536_is_refused(ctx):
    say("superres plane width bound")
    inp = np.zeros((1, 40000), np.uint16)
    with pytest.raises(av1pf.Av1Error) as e:
        ctx.upscale_plane(inp, 40000, 70000, 1, 8)
    assert e.value.status == abi.LIMIT_EXCEEDED


def test_gpu_restore_plane_rejects_planes_wider_than_65536(ctx):
    """The tile list names row and column segments in 16 bits each; a wider
    plane is refused before anything runs (the frame path's context is
    capped at 65536 a side already)."""
    say("restoration plane 65600 wide")
    cfg = abi.LrPlaneConfig(65600, 8, 64, 8, 0)
    cols, rows = av1pf.lr_unit_counts(cfg)
    units = [abi.make_unit(abi.UNIT_NONE)] * (cols * rows)
    plane = np.zeros((8, 65600), np.uint16)
    with pytest.raises(av1pf.Av1Error) as e:
        ctx.restore_plane(plane, plane, cfg, units)
    assert e.value.status == abi.LIMIT_EXCEEDED


def test_gpu_restoration_constant_planes_match_dav1d(ctx):
    say("restoration constant planes")
    for bd, s, plane, exp in G.constant_cases():
        cfg = abi.LrPlaneConfig(7, 9, 64, bd, 0)
        sgr = abi.make_unit(abi.UNIT_SELF_GUIDED, set=s, xqd=(-32, 31))
        assert np.array_equal(ctx.restore_plane(plane, plane, cfg, [sgr]), exp), (bd, s)
        wiener = abi.make_unit(abi.UNIT_WIENER, vertical=(-5, -23, -17), horizontal=(10, 8, 46))
        assert np.array_equal(ctx.restore_plane(plane, plane, cfg, [wiener]), plane)


# ------------------------------------------------------- seeded frames ---

@pytest.mark.parametrize("frame_index", [0, 1, 2, 3])
def test_gpu_frame_cfg0_matches_oracle(ctx, frame_index):
    say(f"cfg0 frame {frame_index}")
    f = av1pf.generate(workloads.cfg0(), frame_index)
    assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), f"cfg0/{frame_index}")


def test_gpu_frame_cfg1_matches_oracle(ctx):
    say("cfg1 frame 0")
    f = av1pf.generate(workloads.cfg1(), 0)
    assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), "cfg1/0")


RANDOM_CASES = [
    dict(seed=1, coded_w=64, coded_h=64),
    dict(seed=2, coded_w=100, coded_h=70),
    dict(seed=3, coded_w=131, coded_h=97, bd=10),
    dict(seed=4, coded_w=90, coded_h=66, bd=12),
    dict(seed=5, coded_w=72, coded_h=80, ssx=0, ssy=0),
    dict(seed=6, coded_w=72, coded_h=80, ssx=1, ssy=0),
    dict(seed=7, coded_w=72, coded_h=80, ssx=0, ssy=1, unit=32),
    dict(seed=8, coded_w=77, coded_h=45, mono=True),
    dict(seed=9, coded_w=150, coded_h=140, unit=32),
    dict(seed=10, coded_w=200, coded_h=130, unit=128, uv_unit=64),
    dict(seed=11, coded_w=93, coded_h=75, upscaled_w=160, lr=False),
    dict(seed=12, coded_w=93, coded_h=75, upscaled_w=160),
    dict(seed=13, coded_w=171, coded_h=130, upscaled_w=300, bd=10, unit=32),
    dict(seed=14, coded_w=64, coded_h=64, lr_types=(0, 1, 0)),
    dict(seed=15, coded_w=80, coded_h=72, cdef=False),
    dict(seed=16, coded_w=80, coded_h=72, cdef=False, lr=False),
    dict(seed=17, coded_w=80, coded_h=72, zero_strengths=True, upscaled_w=128),
    dict(seed=18, coded_w=260, coded_h=200, unit_kinds=(2,), unit=64),
    dict(seed=19, coded_w=260, coded_h=200, unit_kinds=(1,), unit=256),
    # self-guided only at 10 and 12 bits (the box arithmetic is specialised per depth)
    dict(seed=20, coded_w=200, coded_h=140, bd=10, unit_kinds=(2,)),
    dict(seed=21, coded_w=200, coded_h=140, bd=12, unit_kinds=(2,), unit=32),
    # the smallest frames, and extreme aspect ratios
    dict(seed=22, coded_w=8, coded_h=8),
    dict(seed=23, coded_w=16, coded_h=10, bd=12, upscaled_w=30),
    dict(seed=24, coded_w=24, coded_h=300, bd=10),
    dict(seed=25, coded_w=600, coded_h=16, unit=32),
    # superres at 12 bits in 4:4:4, and 4:2:2 superres with 32-sample units
    dict(seed=26, coded_w=120, coded_h=64, bd=12, ssx=0, ssy=0, upscaled_w=200),
    dict(seed=27, coded_w=97, coded_h=70, bd=10, ssx=1, ssy=0, upscaled_w=150, unit=32),
]


@pytest.mark.parametrize("case", RANDOM_CASES, ids=lambda c: "s%d" % c["seed"])
def test_gpu_frame_random_geometry_matches_oracle(ctx, case):
    say(f"random frame {case}")
    f = F.random_frame(**case)
    assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), str(case))


def test_gpu_eight_wave_cdef_workgroups_match_oracle(ctx):
    """CDEF with 512-thread workgroups (tuning cdef_waves = 8): every random
    geometry, 1080p, a mixed batch, and an out-of-range sample late in a
    superblock, which must still raise the reference's error."""
    from vaco_amd import hip
    say("cdef_waves=8: random frames, cfg1 frame 0, batch, error")
    ctx.tuning = {"cdef_waves": 8}
    try:
        for case in RANDOM_CASES:
            f = F.random_frame(**case)
            assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), f"cdef_waves=8 {case}")
        f = av1pf.generate(workloads.cfg1(), 0)
        assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), "cdef_waves=8 cfg1/0")
        frames = [F.random_frame(**c) for c in RANDOM_CASES[:10]]
        dev_in = [[hip.DeviceArray.from_host(p) for p in fr.planes] for fr in frames]
        dev_out = [[hip.DeviceArray(g) for g in fr.output_geometry()] for fr in frames]
        ptr = lambda arrs: [[t.ptr for t in ts] for ts in arrs]
        ctx.batch_device(frames, ptr(dev_in), ptr(dev_out))
        ctx.finish()
        for k, fr in enumerate(frames):
            assert_frame_equal([t.download() for t in dev_out[k]], O.frame(fr.desc, fr.planes), f"batch/{k}")
        f = F.random_frame(46, 64, 64, skip_p=0.0, idx_neg_p=0.0)
        f.planes[0][60, 50] = 999
        with pytest.raises(av1pf.Av1Error) as e:
            ctx.frame(f)
        assert e.value.message == "CDEF sample exceeds bit depth"
    finally:
        ctx.tuning = {"cdef_waves": 16}


@pytest.mark.parametrize("rows", [32, 16])
def test_gpu_short_restoration_tiles_match_oracle(ctx, rows):
    """Restoration cut into 32- or 16-row tiles (tuning lr_tile_rows), which
    run on workgroups of 8 threads per row (256 or 128): every random
    geometry and 1080p."""
    say(f"lr_tile_rows={rows}: random frames, cfg1 frame 0")
    ctx.tuning = {"lr_tile_rows": rows}
    try:
        for case in RANDOM_CASES:
            f = F.random_frame(**case)
            assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), f"lr_tile_rows={rows} {case}")
        f = av1pf.generate(workloads.cfg1(), 0)
        assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), f"lr_tile_rows={rows} cfg1/0")
    finally:
        ctx.tuning = {"lr_tile_rows": 64}


# Planes at a 2-byte offset with odd row strides: every kernel takes its
# unaligned path (no 8-byte loads or stores), and nothing may be written
# outside the output planes' rows.
UNALIGNED_CASES = [
    dict(seed=60, coded_w=131, coded_h=97),
    dict(seed=61, coded_w=93, coded_h=75, bd=10, upscaled_w=160),
    dict(seed=62, coded_w=72, coded_h=80, bd=12, ssx=0, ssy=0, lr=False),
    dict(seed=63, coded_w=97, coded_h=70, bd=10, ssx=1, ssy=0, upscaled_w=150, unit=32),
]
_FILL = 0xABCD


def _strided(shape, stride, src=None):
    """A device buffer with a plane of `shape` from sample 1 on, rows `stride`
    samples apart; every other sample holds _FILL."""
    from vaco_amd import hip
    h, w = shape
    buf = np.full(1 + (h - 1) * stride + w + 7, _FILL, np.uint16)
    if src is not None:
        for y in range(h):
            buf[1 + y * stride:1 + y * stride + w] = src[y]
    return hip.DeviceArray.from_host(buf)


def _unstride(buf, shape, stride):
    h, w = shape
    plane = np.stack([buf[1 + y * stride:1 + y * stride + w] for y in range(h)])
    rest = buf.copy()
    for y in range(h):
        rest[1 + y * stride:1 + y * stride + w] = _FILL
    return plane, bool(np.all(rest == _FILL))


@pytest.mark.parametrize("case", UNALIGNED_CASES, ids=lambda c: "s%d" % c["seed"])
def test_gpu_unaligned_strided_planes_match_oracle(ctx, case):
    say(f"unaligned strided planes {case}")
    f = F.random_frame(**case)
    want = O.frame(f.desc, f.planes)
    in_st = [p.shape[1] + 3 for p in f.planes]
    geo = f.output_geometry()
    out_st = [g[1] + 5 for g in geo]
    ins = [_strided(p.shape, st, p) for p, st in zip(f.planes, in_st)]
    outs = [_strided(g, st) for g, st in zip(geo, out_st)]
    ctx.device_call(f, [d.ptr + 2 for d in ins], [d.ptr + 2 for d in outs],
                    in_strides=in_st, out_strides=out_st)()
    ctx.finish()
    got = []
    for d, g, st in zip(outs, geo, out_st):
        plane, untouched = _unstride(d.download(), g, st)
        assert untouched, f"{case}: a store landed outside the output plane"
        got.append(plane)
    assert_frame_equal(got, want, f"unaligned {case}")
    for d, p, st in zip(ins, f.planes, in_st):
        assert np.array_equal(_unstride(d.download(), p.shape, st)[0], p), "input modified"


# Frames of the 4K 10-bit superres workload that together take every
# denominator 9..16: coded widths 2048, 3413, 1920, 3072, 2363, 2793, 2560, 2194.
CFG2_FRAMES = {0: 15, 1: 9, 2: 16, 3: 10, 5: 13, 11: 11, 12: 12, 15: 14}


@pytest.mark.parametrize("frame_index", sorted(CFG2_FRAMES), ids=lambda k: f"f{k}-d{CFG2_FRAMES[k]}")
def test_gpu_frame_cfg2_4k_superres_matches_oracle(ctx, frame_index):
    f = av1pf.generate(workloads.cfg2(), frame_index)
    say(f"cfg2 frame {frame_index} (4K 10-bit, coded width {f.desc.coded_width})")
    assert f.desc.use_superres and f.desc.coded_width == (3840 * 8 + CFG2_FRAMES[frame_index] // 2) // CFG2_FRAMES[frame_index]
    assert_frame_equal(ctx.frame(f), O.frame(f.desc, f.planes), f"cfg2/{frame_index}")


def test_gpu_frame_device_entry_matches_host_entry(ctx):
    from vaco_amd import hip
    say("device entry vs host entry, cfg1 frames 0-1")
    for k in range(2):
        f = av1pf.generate(workloads.cfg1(), k)
        want = ctx.frame(f)
        dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
        dev_out = [hip.DeviceArray(g) for g in f.output_geometry()]
        ctx.frame_device(f, [t.ptr for t in dev_in], [t.ptr for t in dev_out])
        ctx.finish()
        got = [t.download() for t in dev_out]
        assert_frame_equal(got, want, f"device/{k}")


def test_gpu_calls_alternating_between_two_caller_streams_match_oracle(ctx):
    """Calls alternating between two caller streams overlap on the device
    (each stream's per-frame calls get an intermediate set of their own, and a
    set or the batch workspace last used on the other stream is waited for);
    each must still equal the oracle."""
    from vaco_amd import hip
    say("two caller streams: per-frame and batched calls alternating")
    frames = [av1pf.generate(workloads.cfg1(), k) for k in range(2)]
    frames += [F.random_frame(**c) for c in (dict(seed=12, coded_w=93, coded_h=75, upscaled_w=160),
                                             dict(seed=13, coded_w=171, coded_h=130, upscaled_w=300, bd=10,
                                                  unit=32))]
    want = [O.frame(f.desc, f.planes) for f in frames]
    streams = [hip.Stream(), hip.Stream()]
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    outs = [[[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames] for _ in range(3)]
    ptr = lambda arrs: [t.ptr for t in arrs]
    for rep in range(2):
        for k, f in enumerate(frames):
            ctx.frame_device(f, ptr(dev_in[k]), ptr(outs[rep][k]), stream=streams[(k + rep) % 2].handle)
    ctx.batch_device(frames[:2], [ptr(d) for d in dev_in[:2]], [ptr(o) for o in outs[2][:2]],
                     stream=streams[0].handle)
    ctx.batch_device(frames[2:], [ptr(d) for d in dev_in[2:]], [ptr(o) for o in outs[2][2:]],
                     stream=streams[1].handle)
    ctx.finish()
    for rep in range(3):
        for k in range(len(frames)):
            assert_frame_equal([t.download() for t in outs[rep][k]], want[k], f"stream call {rep}/{k}")
    for st in streams:
        st.destroy()


def test_gpu_calls_on_three_streams_with_chained_launches_match_oracle(ctx):
    """Three caller streams over two intermediate sets (so sets change hands
    between streams) with the chained launch on (chained launches order
    themselves across streams) and superres frames among them."""
    from vaco_amd import hip
    say("three caller streams, chained launches")
    frames = [av1pf.generate(workloads.cfg0(), k) for k in range(3)]
    frames += [F.random_frame(**c) for c in (dict(seed=14, coded_w=200, coded_h=130, bd=10, unit=128, uv_unit=64),
                                             dict(seed=15, coded_w=93, coded_h=75, upscaled_w=160),
                                             dict(seed=16, coded_w=136, coded_h=72, ssx=1, ssy=0))]
    want = [O.frame(f.desc, f.planes) for f in frames]
    streams = [hip.Stream() for _ in range(3)]
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    outs = [[[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames] for _ in range(4)]
    ptr = lambda arrs: [t.ptr for t in arrs]
    saved = ctx.tuning
    ctx.tuning = {"chain": 1}
    try:
        for rep in range(4):
            for k, f in enumerate(frames):
                ctx.frame_device(f, ptr(dev_in[k]), ptr(outs[rep][k]), stream=streams[(k + rep) % 3].handle)
        ctx.finish()
    finally:
        ctx.tuning = saved
    for rep in range(4):
        for k in range(len(frames)):
            assert_frame_equal([t.download() for t in outs[rep][k]], want[k], f"3 streams {rep}/{k}")
    for st in streams:
        st.destroy()


def test_gpu_device_entry_defers_sample_range_errors_to_finish(ctx):
    from vaco_amd import hip
    say("device entry: deferred error")
    f = F.random_frame(44, 64, 64, skip_p=0.0, idx_neg_p=0.0)
    f.planes[0][20, 20] = 999
    dev_in = [hip.DeviceArray.from_host(p) for p in f.planes]
    dev_out = [hip.DeviceArray(g) for g in f.output_geometry()]
    ctx.frame_device(f, [t.ptr for t in dev_in], [t.ptr for t in dev_out])
    with pytest.raises(av1pf.Av1Error) as e:
        ctx.finish()
    assert e.value.message == "CDEF sample exceeds bit depth"
    ctx.finish()  # the error is reported once


# --------------------------------------------------------------- errors ---

def test_gpu_out_of_range_samples_report_the_reference_error(ctx):
    say("error: out-of-range samples")
    f = F.random_frame(40, 64, 64, skip_p=0.0, idx_neg_p=0.0)
    f.planes[0][10, 10] = 256
    with pytest.raises(O.OracleError) as want:
        O.frame(f.desc, f.planes)
    with pytest.raises(av1pf.Av1Error) as got:
        ctx.frame(f)
    assert (got.value.status, got.value.message) == (want.value.status, want.value.message)
    # a bad sample no stage reads is not an error (CDEF skips it, LR off)
    g = F.random_frame(41, 64, 64, skip_p=1.0, lr=False)
    g.planes[1][5, 5] = 4000
    assert_frame_equal(ctx.frame(g), O.frame(g.desc, g.planes), "unread bad sample")


def outcome_oracle(f, **kw):
    try:
        return ("ok", O.frame(f.desc, f.planes, **kw))
    except O.OracleError as e:
        return ("err", e.status, e.message)


def outcome_gpu(ctx, f, **kw):
    try:
        return ("ok", ctx.frame(f, **kw))
    except av1pf.Av1Error as e:
        return ("err", e.status, e.message)


def same_outcome(a, b, label):
    assert a[0] == b[0], (label, a[:3] if a[0] == "err" else "ok", b[:3] if b[0] == "err" else "ok")
    if a[0] == "ok":
        assert_frame_equal(a[1], b[1], label)
    else:
        assert a[1:] == b[1:], label


def _mutations():
    def luma_primary_16(f):  # adjusted by variance: an error only when var >= 2^18
        f.desc.cdef.y[3].primary = 16
        f._keep[0][:] = 3

    def chroma_primary_16(f):  # chroma primary is not adjusted: always an error
        f.desc.cdef.uv[2].primary = 16
        f._keep[0][:] = 2

    def secondary_5(f):
        f.desc.cdef.y[1].secondary = 5
        f._keep[0][:] = 1

    def index_8(f):
        f._keep[0][0] = 8

    def damping_13(f):
        f.desc.cdef.damping = 13

    def bad_depth(f):
        f.desc.bit_depth = 9

    def unit_count(f):
        f.desc.lr_unit_count[1] -= 1

    def wiener_tap(f):
        f.desc.lr_units[0][0].kind = abi.UNIT_WIENER
        f.desc.lr_units[0][0].vertical[0] = 11

    def sgr_xqd(f):
        f.desc.lr_units[2][0].kind = abi.UNIT_SELF_GUIDED
        f.desc.lr_units[2][0].xqd[1] = 96

    def no_unit_size(f):
        f.desc.lr_unit_size[1] = 0

    def superres_not_wider(f):
        f.desc.use_superres = 1

    return [luma_primary_16, chroma_primary_16, secondary_5, index_8, damping_13, bad_depth, unit_count,
            wiener_tap, sgr_xqd, no_unit_size, superres_not_wider]


@pytest.mark.parametrize("mutate", _mutations(), ids=lambda m: m.__name__)
def test_gpu_invalid_parameters_match_reference_outcome(ctx, mutate):
    say(f"error: {mutate.__name__}")
    for seed in (42, 43, 44):
        f = F.random_frame(seed, 64, 64, skip_p=0.0, idx_neg_p=0.0)
        mutate(f)
        same_outcome(outcome_oracle(f), outcome_gpu(ctx, f), f"{mutate.__name__}/{seed}")


def test_gpu_fuel_exhaustion_matches_reference_accounting(ctx):
    say("error: fuel")
    f = F.random_frame(43, 64, 32, lr=False, skip_p=0.0, idx_neg_p=0.0, upscaled_w=80)
    need = 32 + 80 * 32 + 2 * 40 * 16
    out, left = ctx.frame(f, fuel=need + 5)
    assert left == 5
    with pytest.raises(av1pf.Av1Error) as e:
        ctx.frame(f, fuel=need - 1)
    assert e.value.status == abi.LIMIT_EXCEEDED and e.value.message == "fuel"


# -------------------------------------------------------------- batched ---

def test_gpu_batch_of_mixed_frames_matches_per_frame_and_oracle(ctx):
    from vaco_amd import hip
    say("batch: mixed geometry / depth / layout / superres")
    frames = [F.random_frame(**c) for c in RANDOM_CASES[:14]]
    frames += [av1pf.generate(workloads.cfg0(), k) for k in range(3)]
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    dev_b = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames]
    ptr = lambda arrs: [[t.ptr for t in ts] for ts in arrs]
    ctx.batch_device(frames, ptr(dev_in), ptr(dev_b))
    ctx.finish()
    for k, f in enumerate(frames):
        got = [t.download() for t in dev_b[k]]
        assert_frame_equal(got, O.frame(f.desc, f.planes), f"batch/{k}")
    # a second, different batch reuses the grown workspace
    ctx.batch_device(frames[::-1], ptr(dev_in[::-1]), ptr(dev_b[::-1]))
    ctx.finish()
    for k, f in enumerate(frames):
        assert_frame_equal([t.download() for t in dev_b[k]], O.frame(f.desc, f.planes), f"batch2/{k}")


def test_gpu_batch_defers_the_first_failing_frames_error(ctx):
    from vaco_amd import hip
    say("batch: deferred error")
    frames = [F.random_frame(50 + k, 64, 64, skip_p=0.0, idx_neg_p=0.0) for k in range(3)]
    frames[1].planes[0][9, 9] = 999
    dev_in = [[hip.DeviceArray.from_host(p) for p in f.planes] for f in frames]
    dev_out = [[hip.DeviceArray(g) for g in f.output_geometry()] for f in frames]
    ptr = lambda arrs: [[t.ptr for t in ts] for ts in arrs]
    ctx.batch_device(frames, ptr(dev_in), ptr(dev_out))
    with pytest.raises(av1pf.Av1Error) as e:
        ctx.finish()
    want = outcome_oracle(frames[1])
    assert (e.value.status, e.value.message) == want[1:]
    assert_frame_equal([t.download() for t in dev_out[2]], O.frame(frames[2].desc, frames[2].planes), "ok frame")
