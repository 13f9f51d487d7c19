"""Regenerates the inputs of the reference's golden-vector tests and loads the
expected outputs from the fixtures copied into ``tests/golden/``.

The inputs are rebuilt exactly as the reference's own tests build them:

* CDEF corpus: ``crates/tool/vaco-checkasm/tests/av1_cdef.rs:28-83``
* superres corpus: ``crates/codec/vaco-codec-av1/tests/superres.rs:37-62``
* restoration corpus: ``crates/codec/vaco-codec-av1/tests/restoration.rs:142-227``
* constant-plane corpus: ``crates/codec/vaco-codec-av1/tests/restoration.rs:48-93``

The fixture bytes are dav1d (rev aa09a630ef57ee7d9482ffb7ef355a903dbb5302)
outputs checked into the reference; their SHA-256 digests are pinned below.
``av1_libaom_symbols.bin`` is the reference's symbol-decoder trace from
libaom 3.14.1 (crates/signal/vaco-codec-msac/tests/av1_oracle.rs:1-32).
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from vaco_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SHA256 = {
    "av1-cdef-dav1d.bin": "54a4429a8477d102bfd25c356755a9bcf5c125cb9ece275123b335764b9d73b4",
    "superres-dav1d.u16le": "cb71c667221249f369f8d01d38f6fbda6bd72dfa8699a9c1d2772ad1386dc336",
    "restoration-dav1d-8.u16le": "e954d64ad6726702d9347508d7e7982a6eddfe84e11aad361edf1c512a5f761f",
    "restoration-dav1d-high.u16le": "1f46e3cc1038ee590638381de786f51a63949d5d9d5c762be34f9980c76c6780",
    "restoration-constant-8.u16le": "222e3f56d30ec897fa902cc63b2d21e63b2d7555443b97c85e1af82241f0f56e",
    "restoration-constant-high.u16le": "2861f658c08a5231da898e1a0944524816a66aa95e7011d3037bca0234af8bd2",
    "av1_libaom_symbols.bin": "c10a382c4619d2838a339bee424ef60505a978424f2cbfa84ccff758bec425f8",
}


def read(name: str) -> bytes:
    with open(os.path.join(GOLDEN, name), "rb") as f:
        data = f.read()
    assert hashlib.sha256(data).hexdigest() == SHA256[name], name
    return data


# --------------------------------------------------------------- CDEF ----

def cdef_cases():
    """7,680 cases: (inputs[n,144] u16, params list, expected dir[n], var[n], out[n,64])."""
    fixture = np.frombuffer(read("av1-cdef-dav1d.bin"), dtype=np.uint8).reshape(7680, 136)
    meta = []
    for bd in (8, 10, 12):
        for damping in range(2, 7):
            for direction in range(8):
                for primary in range(16):
                    for secondary in (0, 1, 2, 4):
                        meta.append((bd, damping, direction, primary, secondary))
    n = len(meta)
    assert n == 7680
    ids = np.arange(n, dtype=np.uint32)
    bds = np.array([m[0] for m in meta], dtype=np.uint32)
    state = (np.uint32(0x9E3779B9) ^ ids).astype(np.uint32)
    inputs = np.zeros((n, 144), dtype=np.uint16)
    odd = (ids & 1) == 1
    low_mask = ((np.uint32(1) << (bds - 8)) - 1).astype(np.uint32)
    for k in range(144):
        state ^= (state << np.uint32(13)) & np.uint32(0xFFFFFFFF)
        state ^= state >> np.uint32(17)
        state ^= (state << np.uint32(5)) & np.uint32(0xFFFFFFFF)
        base = np.where(odd, 96 + state % 33, state & 255).astype(np.uint32)
        v = ((base << (bds - 8)) & 0xFFFF) | ((state >> 8) & 0xFFFF & low_mask)
        inputs[:, k] = v.astype(np.uint16)
    params = []
    for i, (bd, damping, direction, primary, secondary) in enumerate(meta):
        p = abi.FilterParams()
        p.width = 8 if i % 3 == 0 else 4
        p.height = 4 if i % 3 == 2 else 8
        p.edges = (i // 3) % 16
        p.bit_depth = bd
        p.direction = direction
        p.primary = primary << (bd - 8)
        p.secondary = secondary << (bd - 8)
        p.damping = damping + bd - 8
        params.append(p)
    exp_dir = fixture[:, 0:4].copy().view("<u4").reshape(n)
    exp_var = fixture[:, 4:8].copy().view("<u4").reshape(n)
    exp_out = fixture[:, 8:].copy().view("<u2").reshape(n, 64)
    return inputs, params, exp_dir, exp_var, exp_out


def luma_of(case_input: np.ndarray) -> np.ndarray:
    """The 8x8 at (2,2) of a 12x12 neighbourhood (av1_cdef.rs:90-94)."""
    return case_input.reshape(12, 12)[2:10, 2:10].copy()


# ----------------------------------------------------------- superres ----

def superres_cases():
    """18 cases: list of (bd, visible, padded, out_w, h, input[h,padded], expected[h,out_w])."""
    data = np.frombuffer(read("superres-dav1d.u16le"), dtype="<u2")
    off = 0
    cases = []
    case_id = 0
    while off < len(data):
        bd, vis, padded, out_w, h = (int(v) for v in data[off:off + 5])
        off += 5
        mx = (1 << bd) - 1
        state = (0x6D2B79F5 ^ case_id) & 0xFFFFFFFF
        inp = np.zeros((h, padded), np.uint16)
        for y in range(h):
            for x in range(padded):
                state ^= (state << 13) & 0xFFFFFFFF
                state ^= state >> 17
                state ^= (state << 5) & 0xFFFFFFFF
                inp[y, x] = ((state + x * 37 + y * 91) & 0xFFFF) & mx
        exp = data[off:off + out_w * h].reshape(h, out_w).astype(np.uint16)
        off += out_w * h
        cases.append((bd, vis, padded, out_w, h, inp, exp))
        case_id += 1
    assert case_id == 18
    return cases


# -------------------------------------------------------- restoration ----

LR_SHAPES = [(7, 9, False, 64), (13, 65, False, 64), (97, 130, False, 64), (48, 81, True, 32),
             (135, 129, False, 128)]


def lr_sample(x, y, depth, after):
    return (((x * 71 + y * 47 + x * y * 13 + int(after) * 39) ^ ((x + y) * 29)) & ((1 << depth) - 1))


def oracle_unit(mode: int, index: int, chroma: bool) -> abi.LrUnit:
    """tests/restoration.rs:147-169."""
    if mode < 16:
        return abi.make_unit(abi.UNIT_SELF_GUIDED, set=(mode + index) % 16,
                             xqd=(-96 + (mode * 11 + index * 7) % 128, -32 + (mode * 17 + index * 13) % 128))
    coeffs = [[-5, -23, -17], [10, 8, 46], [3, -7, 15], [0, 0, 0]]
    horizontal = list(coeffs[(mode - 16 + index) % 4])
    vertical = list(coeffs[(mode - 16 + index + 1) % 4])
    if chroma:
        horizontal[0] = 0
        vertical[0] = 0
    return abi.make_unit(abi.UNIT_WIENER, vertical=vertical, horizontal=horizontal)


def unit_counts(width, height, unit_size):
    def count(e):
        return max(1, (e + unit_size // 2) // unit_size)
    return count(width), count(height)


def restoration_cases():
    """300 cases: (bd, w, h, ss_y, unit, mode, before, after, units, expected[h,w])."""
    data = np.frombuffer(read("restoration-dav1d-8.u16le") + read("restoration-dav1d-high.u16le"), dtype="<u2")
    off = 0
    cases = []
    for bd in (8, 10, 12):
        for (w, h, ssy, unit) in LR_SHAPES:
            ys, xs = np.mgrid[0:h, 0:w]
            before = np.zeros((h + 8, w + 8), np.uint16)
            after = np.zeros((h + 8, w + 8), np.uint16)
            before[:h, :w] = lr_sample(xs, ys, bd, False)
            after[:h, :w] = lr_sample(xs, ys, bd, True)
            cols, rows = unit_counts(w, h, unit)
            for mode in range(20):
                units = [oracle_unit(mode, i, ssy) for i in range(cols * rows)]
                exp = data[off:off + w * h].reshape(h, w).astype(np.uint16)
                off += w * h
                cfg = abi.LrPlaneConfig(w, h, unit, bd, int(ssy))
                cases.append((bd, w, h, ssy, unit, mode, before, after, units, cfg, exp))
    assert off == len(data) and len(cases) == 300
    return cases


def constant_cases():
    """48 cases: (bd, set, plane[9,7] at max, expected[9,7]) with xqd (-32, 31)."""
    data = np.frombuffer(read("restoration-constant-8.u16le") + read("restoration-constant-high.u16le"), dtype="<u2")
    off = 0
    cases = []
    for bd in (8, 10, 12):
        plane = np.full((9, 7), (1 << bd) - 1, np.uint16)
        for s in range(16):
            exp = data[off:off + 63].reshape(9, 7).astype(np.uint16)
            off += 63
            cases.append((bd, s, plane, exp))
    assert off == len(data)
    return cases
