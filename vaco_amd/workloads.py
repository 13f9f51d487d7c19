"""Seeded synthetic workloads of the stage's benchmark configurations.

Each preset is a :class:`GenConfig` for the integer generator in
``vaco_amd/csrc/av1pf_gen.cpp``. Base seed ``0xA51CDEF0 + config``.

* ``cfg0``: one 8-bit 4:2:0 frame at 352x288, luma/chroma = gradient +
  N(0, 6), 64x64 restoration units, 20% 8x8 skip (CPU-runnable case)
* ``cfg1``: one 8-bit 4:2:0 frame at 1920x1080, as cfg0 plus 8 hard edges
  at random angles, 128x128 restoration units (plain per-frame GPU case)
* ``cfg2``: 10-bit 4:2:0 3840x2160 frames, N(0, 24), superres with the
  denominator drawn from 9..16 per frame, 64x64 units as cfg0
* ``mixed_batch``: 64 frames of mixed geometry, depth and superres
* ``serve_streams``: 32 streams of fixed resolution and depth for the
  serving loop
"""
from __future__ import annotations

from .av1pf import GenConfig

BASE_SEED = 0xA51CDEF0


def make(width, height, bit_depth=8, sigma=6, edges=0, unit=64, superres_denom=8, skip_permille=200,
         lr=True, mono=False, seed=0) -> GenConfig:
    c = GenConfig()
    c.width, c.height = width, height
    c.superres_denom = superres_denom
    c.bit_depth = bit_depth
    c.hard_edges = edges
    c.monochrome = 1 if mono else 0
    c.lr_enabled = 1 if lr else 0
    c.noise_sigma = sigma
    c.lr_unit_size = unit
    c.skip_permille = skip_permille
    c.seed = seed
    return c


def cfg0() -> GenConfig:
    return make(352, 288, 8, sigma=6, edges=0, unit=64, seed=BASE_SEED + 0)


def cfg1() -> GenConfig:
    return make(1920, 1080, 8, sigma=6, edges=8, unit=128, seed=BASE_SEED + 1)


def cfg2() -> GenConfig:
    return make(3840, 2160, 10, sigma=24, edges=0, unit=64, superres_denom=0, seed=BASE_SEED + 2)


CONFIGS = {0: cfg0, 1: cfg1, 2: cfg2}


MIXED_SHAPES = [((640, 360), 0.25), ((1280, 720), 0.25), ((1920, 1080), 0.35), ((3840, 2160), 0.15)]


def mixed_batch(n=64, seed=BASE_SEED + 3):
    """Per-frame configurations of the mixed-geometry batch: shape drawn from
    360p 25% / 720p 25% / 1080p 35% / 2160p 15%, bit depth U{8, 10}, superres
    on half the frames (denominator U[9, 16]), CDEF and restoration parameters
    as cfg0/cfg2 (64x64 units). Returns [(GenConfig, frame_index)]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    cum = np.cumsum([p for _, p in MIXED_SHAPES])
    for i in range(n):
        (w, h) = MIXED_SHAPES[int(np.searchsorted(cum, rng.random()))][0]
        bd = 8 if rng.random() < 0.5 else 10
        denom = 0 if rng.random() < 0.5 else 8
        c = make(w, h, bd, sigma=6 if bd == 8 else 24, edges=0, unit=64, superres_denom=denom,
                 seed=seed + 1000 * i)
        out.append((c, i))
    return out


SERVE_SHAPES = [(1280, 720), (1920, 1080), (3840, 2160)]


def serve_streams(n=32, seed=BASE_SEED + 4):
    """Stream configurations of the serving workload: each stream has a fixed
    resolution drawn from 720p / 1080p / 2160p and a fixed bit depth from
    {8, 10}; content and filter parameters as cfg1 (8-bit: hard edges,
    128x128 units) and cfg2 (10-bit: N(0, 24), superres with the
    denominator drawn per frame, 64x64 units). Returns
    [GenConfig] (frame k of stream s is generate(cfg[s], k))."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for s in range(n):
        w, h = SERVE_SHAPES[int(rng.integers(0, len(SERVE_SHAPES)))]
        if rng.random() < 0.5:
            c = make(w, h, 8, sigma=6, edges=8, unit=128, seed=seed + 1000 * s)
        else:
            c = make(w, h, 10, sigma=24, edges=0, unit=64, superres_denom=0, seed=seed + 1000 * s)
        out.append(c)
    return out
