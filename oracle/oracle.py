"""Python binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/build/liboracle_av1pf.so`` (built by ``oracle/Makefile``),
the plain-C restatement of matteopolak/vaco's scalar AV1 filter stage.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only as the checker or the timed
CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from vaco_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle_av1pf.so")
_lib = None


class OracleError(Exception):
    def __init__(self, status: int, message: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


def build() -> str:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        E = C.POINTER(C.c_char_p)
        L.orc_find_direction.argtypes = [C.POINTER(C.c_uint16), C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), E]
        L.orc_filter_block.argtypes = [C.POINTER(C.c_uint16), C.POINTER(abi.FilterParams), C.POINTER(C.c_uint16), E]
        L.orc_adjust_strength.argtypes = [C.c_uint16, C.c_uint32]
        L.orc_adjust_strength.restype = C.c_uint16
        L.orc_chroma_direction.argtypes = [C.c_uint8, C.c_int, C.c_int]
        L.orc_chroma_direction.restype = C.c_uint8
        L.orc_upscale_plane.argtypes = [C.POINTER(abi.Plane), C.POINTER(abi.SrPlaneConfig), C.POINTER(abi.Plane), C.POINTER(C.c_uint64), E]
        L.orc_lr_unit_counts.argtypes = [C.POINTER(abi.LrPlaneConfig), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.orc_lr_unit_index.argtypes = [C.POINTER(abi.LrPlaneConfig), C.c_size_t, C.c_size_t]
        L.orc_lr_unit_index.restype = C.c_size_t
        L.orc_restore_plane.argtypes = [C.POINTER(abi.Plane), C.POINTER(abi.Plane), C.POINTER(abi.LrPlaneConfig), C.POINTER(abi.LrUnit), C.c_size_t, C.POINTER(abi.Plane), E]
        L.orc_frame_output_geometry.argtypes = [C.POINTER(abi.FrameDesc), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_frame.argtypes = [C.POINTER(abi.FrameDesc), C.POINTER(abi.Picture), C.POINTER(abi.Picture), C.POINTER(C.c_uint64), E]
        L.orc_pic_to_frame.argtypes = [C.POINTER(abi.FrameDesc), C.POINTER(abi.Picture), C.c_void_p * 3,
                                       C.c_uint32 * 3]
        L.orc_pic_to_frame.restype = None
        for f in ("orc_find_direction", "orc_filter_block", "orc_upscale_plane", "orc_restore_plane", "orc_frame", "orc_frame_output_geometry"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _check(rc: int, err: C.c_char_p):
    if rc != 0:
        raise OracleError(rc, err.value.decode() if err.value else "")


def find_direction(block: np.ndarray, bit_depth: int):
    """cdef::find_direction (cdef.rs:79-124) -> (direction, variance)."""
    b = np.ascontiguousarray(block, dtype=np.uint16).reshape(64)
    d, v, err = C.c_uint8(), C.c_uint32(), C.c_char_p()
    _check(lib().orc_find_direction(abi.u16_ptr(b), bit_depth, C.byref(d), C.byref(v), C.byref(err)), err)
    return d.value, v.value


def filter_block(inp: np.ndarray, params: abi.FilterParams) -> np.ndarray:
    """cdef::filter_block (cdef.rs:142-205) -> 8x8 uint16 (zero outside w x h)."""
    i = np.ascontiguousarray(inp, dtype=np.uint16).reshape(144)
    out = np.zeros(64, np.uint16)
    err = C.c_char_p()
    _check(lib().orc_filter_block(abi.u16_ptr(i), C.byref(params), abi.u16_ptr(out), C.byref(err)), err)
    return out.reshape(8, 8)


def adjust_strength(primary: int, variance: int) -> int:
    return lib().orc_adjust_strength(primary, variance)


def chroma_direction(direction: int, sub_x: bool, sub_y: bool) -> int:
    return lib().orc_chroma_direction(direction, int(sub_x), int(sub_y))


def upscale_plane(inp: np.ndarray, visible_width: int, output_width: int, height: int, bit_depth: int,
                  fuel: int | None = None) -> np.ndarray:
    """superres::upscale_plane (superres.rs:163-216); ``inp`` is the padded plane."""
    i = np.ascontiguousarray(inp, dtype=np.uint16)
    cfg = abi.SrPlaneConfig(visible_width, output_width, height, bit_depth)
    out = np.zeros((height, output_width), np.uint16)
    pi, po = abi.plane_of(i), abi.plane_of(out)
    err = C.c_char_p()
    f = C.c_uint64(fuel) if fuel is not None else None
    _check(lib().orc_upscale_plane(C.byref(pi), C.byref(cfg), C.byref(po), C.byref(f) if f is not None else None, C.byref(err)), err)
    return out


def lr_unit_counts(cfg: abi.LrPlaneConfig):
    c, r = C.c_size_t(), C.c_size_t()
    lib().orc_lr_unit_counts(C.byref(cfg), C.byref(c), C.byref(r))
    return c.value, r.value


def lr_unit_index(cfg: abi.LrPlaneConfig, x: int, y: int) -> int:
    return lib().orc_lr_unit_index(C.byref(cfg), x, y)


def restore_plane(before: np.ndarray, after: np.ndarray, cfg: abi.LrPlaneConfig, units) -> np.ndarray:
    """restoration::restore_plane (restoration.rs:178-207)."""
    b = np.ascontiguousarray(before, dtype=np.uint16)
    a = np.ascontiguousarray(after, dtype=np.uint16)
    out = np.zeros((max(cfg.height, 1), max(cfg.width, 1)), np.uint16)
    pb, pa, po = abi.plane_of(b), abi.plane_of(a), abi.plane_of(out)
    arr = abi.units_array(units)
    err = C.c_char_p()
    _check(lib().orc_restore_plane(C.byref(pb), C.byref(pa), C.byref(cfg), arr, len(units), C.byref(po), C.byref(err)), err)
    return out[: cfg.height, : cfg.width]


def frame_output_geometry(desc: abi.FrameDesc):
    w, h = (C.c_uint32 * 3)(), (C.c_uint32 * 3)()
    lib().orc_frame_output_geometry(C.byref(desc), w, h)
    return [(h[i], w[i]) for i in range(3)]


def frame(desc: abi.FrameDesc, planes, fuel: int | None = None):
    """decode.rs:554-559 as a unit on host planes -> list of output arrays."""
    geo = frame_output_geometry(desc)
    nplanes = 1 if desc.monochrome else 3
    outs = [np.zeros(geo[i], np.uint16) for i in range(nplanes)]
    pin = abi.picture_of([np.ascontiguousarray(p) for p in planes[:nplanes]])
    pout = abi.picture_of(outs)
    err = C.c_char_p()
    f = C.c_uint64(fuel) if fuel is not None else None
    _check(lib().orc_frame(C.byref(desc), C.byref(pin), C.byref(pout), C.byref(f) if f is not None else None, C.byref(err)), err)
    return outs


def pic_to_frame(desc: abi.FrameDesc, pic_planes, frame_geometry, bytes_per_sample):
    """decode.rs:3795-3891: the final picture as frame planes (zero-filled
    first, as Frame::alloc_video leaves them), uint8 or uint16."""
    dt = np.uint8 if bytes_per_sample == 1 else np.uint16
    out = [np.zeros(g, dt) for g in frame_geometry]
    pic = abi.picture_of([np.ascontiguousarray(p, dtype=np.uint16) for p in pic_planes])
    ptrs = (C.c_void_p * 3)(*[o.ctypes.data for o in out] + [None] * (3 - len(out)))
    strides = (C.c_uint32 * 3)(*[o.shape[1] * o.itemsize for o in out] + [0] * (3 - len(out)))
    lib().orc_pic_to_frame(C.byref(desc), C.byref(pic), ptrs, strides)
    return out
