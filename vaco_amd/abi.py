"""ctypes mirror of include/vaco_av1pf.h (plain data types only).

The structures below are field-for-field copies of the C ABI. They are the
Python-side names of the reference's parameter types:

* ``Strength`` / ``CdefParams`` / ``FilterParams``: ``cdef.rs:14-54``
* ``SrPlaneConfig``: ``superres.rs:93-103``
* ``LrUnit`` / ``LrPlaneConfig``: ``restoration.rs:46-106``
* ``Plane`` / ``Picture``: ``framebuf.rs:14-20, 78-83``
* ``FrameDesc``: ``decode.rs:185-214`` (``FrameCtx``) plus ``FrameSize``
  (``crates/codec/vaco-parse-av1/src/frame_header.rs:71-85``)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

# VACO_AV1PF_ABI_VERSION of include/vaco_av1pf.h that these layouts follow.
ABI_VERSION = 3

OK = 0
INVALID_DATA = 1
UNSUPPORTED = 2
LIMIT_EXCEEDED = 3
DEVICE_ERROR = 4
INVALID_ARGUMENT = 5

STATUS_NAMES = {
    OK: "OK",
    INVALID_DATA: "InvalidData",
    UNSUPPORTED: "Unsupported",
    LIMIT_EXCEEDED: "LimitExceeded",
    DEVICE_ERROR: "DeviceError",
    INVALID_ARGUMENT: "InvalidArgument",
}

RESTORE_NONE, RESTORE_SWITCHABLE, RESTORE_WIENER, RESTORE_SELF_GUIDED = 0, 1, 2, 3
UNIT_NONE, UNIT_WIENER, UNIT_SELF_GUIDED = 0, 1, 2


class Strength(C.Structure):
    _fields_ = [("primary", C.c_uint8), ("secondary", C.c_uint8)]


class CdefParams(C.Structure):
    _fields_ = [
        ("enabled", C.c_uint8),
        ("damping", C.c_uint8),
        ("y", Strength * 8),
        ("uv", Strength * 8),
    ]


class FilterParams(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("edges", C.c_uint8),
        ("bit_depth", C.c_uint8),
        ("direction", C.c_uint8),
        ("damping", C.c_uint8),
        ("primary", C.c_uint16),
        ("secondary", C.c_uint16),
    ]


class SrPlaneConfig(C.Structure):
    _fields_ = [
        ("visible_width", C.c_uint32),
        ("output_width", C.c_uint32),
        ("height", C.c_uint32),
        ("bit_depth", C.c_uint8),
    ]


class LrUnit(C.Structure):
    _fields_ = [
        ("kind", C.c_uint8),
        ("set", C.c_uint8),
        ("xqd", C.c_int16 * 2),
        ("vertical", C.c_int16 * 3),
        ("horizontal", C.c_int16 * 3),
    ]


class LrPlaneConfig(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("unit_size", C.c_uint32),
        ("bit_depth", C.c_uint8),
        ("subsampling_y", C.c_uint8),
    ]


class Plane(C.Structure):
    _fields_ = [
        ("data", C.POINTER(C.c_uint16)),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("stride", C.c_uint32),
    ]


class Picture(C.Structure):
    _fields_ = [("plane", Plane * 3)]


class FrameDesc(C.Structure):
    _fields_ = [
        ("coded_width", C.c_uint32),
        ("upscaled_width", C.c_uint32),
        ("coded_height", C.c_uint32),
        ("use_superres", C.c_uint8),
        ("bit_depth", C.c_uint8),
        ("subsampling_x", C.c_uint8),
        ("subsampling_y", C.c_uint8),
        ("monochrome", C.c_uint8),
        ("cdef", CdefParams),
        ("cdef_idx", C.POINTER(C.c_int8)),
        ("cdef_stride", C.c_uint32),
        ("skip8", C.POINTER(C.c_uint8)),
        ("lr_type", C.c_uint8 * 3),
        ("lr_unit_size", C.c_uint32 * 3),
        ("lr_units", C.POINTER(LrUnit) * 3),
        ("lr_unit_count", C.c_uint32 * 3),
    ]


class FramePlanes(C.Structure):
    _fields_ = [("data", C.c_void_p * 3), ("stride_bytes", C.c_uint32 * 3)]


class BitReader(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("size_bits", C.c_uint64),
        ("pos", C.c_uint64),
        ("overrun", C.c_uint32),
    ]


class SymbolDecoder(C.Structure):
    _fields_ = [
        ("bits", BitReader),
        ("value", C.c_uint32),
        ("range", C.c_uint32),
        ("max_bits", C.c_int64),
        ("disable_cdf_update", C.c_uint32),
    ]


class LrTileState(C.Structure):
    _fields_ = [
        ("use_wiener", C.c_uint16 * 3),
        ("use_sgrproj", C.c_uint16 * 3),
        ("restoration_type", C.c_uint16 * 4),
        ("ref_sgr_xqd", (C.c_int16 * 2) * 3),
        ("ref_lr_wiener", ((C.c_int16 * 3) * 2) * 3),
    ]


class SeqInfo(C.Structure):
    _fields_ = [
        ("enable_cdef", C.c_uint8),
        ("enable_restoration", C.c_uint8),
        ("enable_superres", C.c_uint8),
        ("use_128x128_superblock", C.c_uint8),
        ("subsampling_x", C.c_uint8),
        ("subsampling_y", C.c_uint8),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("superres_rows", C.c_uint8),
        ("cdef_xcd_order", C.c_uint8),
        ("lr_order", C.c_uint8),
        ("lr_tile_rows", C.c_uint8),
        ("cdef_waves", C.c_uint8),
        ("chain", C.c_uint8),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("window_us", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("max_inflight", C.c_uint32),
    ]


class Completion(C.Structure):
    _fields_ = [
        ("tag", C.c_uint64),
        ("status", C.c_int32),
        ("batch_size", C.c_uint32),
        ("t_submit_ns", C.c_uint64),
        ("t_launch_ns", C.c_uint64),
        ("t_done_ns", C.c_uint64),
        ("t_gpu_start_ns", C.c_uint64),
        ("t_gpu_end_ns", C.c_uint64),
    ]


def u16_ptr(a: np.ndarray):
    assert a.dtype == np.uint16 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_uint16))


def plane_of(a: np.ndarray) -> Plane:
    """A ``Plane`` viewing a C-contiguous 2-D uint16 array (host memory)."""
    assert a.ndim == 2
    p = Plane()
    p.data = u16_ptr(a)
    p.height, p.width = a.shape
    p.stride = a.shape[1]
    return p


def device_plane(ptr: int, width: int, height: int, stride: int | None = None) -> Plane:
    """A ``Plane`` over device memory at integer address ``ptr``."""
    p = Plane()
    p.data = C.cast(C.c_void_p(ptr), C.POINTER(C.c_uint16))
    p.width, p.height = width, height
    p.stride = width if stride is None else stride
    return p


def picture_of(planes) -> Picture:
    pic = Picture()
    for i, a in enumerate(planes):
        if a is not None:
            pic.plane[i] = plane_of(a)
    return pic


def make_unit(kind=UNIT_NONE, set=0, xqd=(0, 0), vertical=(0, 0, 0), horizontal=(0, 0, 0)) -> LrUnit:
    u = LrUnit()
    u.kind = kind
    u.set = set
    u.xqd[:] = list(xqd)
    u.vertical[:] = list(vertical)
    u.horizontal[:] = list(horizontal)
    return u


def units_array(units) -> "C.Array":
    arr = (LrUnit * max(1, len(units)))()
    for i, u in enumerate(units):
        arr[i] = u
    return arr
