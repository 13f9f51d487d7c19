"""Seeded synthetic workloads of the stage's benchmark configurations.

Each preset is a :class:`GenConfig` for the integer generator in
``vaco_amd/csrc/av1pf_gen.cpp``. Base seed ``0xA51CDEF0 + config``.

* ``cfg0``: one 8-bit 4:2:0 frame at 352x288, luma/chroma = gradient +
  N(0, 6), 64x64 restoration units, 20% 8x8 skip (CPU-runnable case)
* ``cfg1``: one 8-bit 4:2:0 frame at 1920x1080, as cfg0 plus 8 hard edges
  at random angles, 128x128 restoration units (plain per-frame GPU case)
* ``cfg2``: 10-bit 4:2:0 3840x2160 frames, N(0, 24), superres with the
  denominator drawn from 9..16 per frame, 64x64 units as cfg0
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
