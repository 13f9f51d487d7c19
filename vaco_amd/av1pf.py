"""Python mirror of the stage's operator interface over the C ABI.

The functions here keep the reference's names, argument meaning and error
behaviour (crates/codec/vaco-codec-av1/src/ of matteopolak/vaco):

* :func:`find_direction_filter_blocks` - ``cdef::find_direction`` +
  ``cdef::filter_block`` (cdef.rs:79-205) on the GPU, batched
* :func:`upscale_plane` - ``superres::upscale_plane`` (superres.rs:163-216)
* :func:`restore_plane` - ``restoration::restore_plane`` (restoration.rs:178-207)
* :meth:`Context.frame` - ``decode.rs:554-559`` (pre-CDEF copy, apply_cdef,
  apply_superres, apply_restoration) on host planes
* :meth:`Context.frame_device` - the same stage on device-resident planes

Failures raise :class:`Av1Error` carrying the reference's error class
(``InvalidData``, ``LimitExceeded``, ...) and message. The HIP library is
mandatory: importing this module on a machine without it raises at once;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvaco_av1pf.so")

_lib = None


class Av1Error(Exception):
    def __init__(self, status: int, message: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class GenConfig(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("superres_denom", C.c_uint32),
        ("bit_depth", C.c_uint8),
        ("hard_edges", C.c_uint8),
        ("monochrome", C.c_uint8),
        ("lr_enabled", C.c_uint8),
        ("noise_sigma", C.c_uint16),
        ("lr_unit_size", C.c_uint16),
        ("skip_permille", C.c_uint16),
        ("pad", C.c_uint16),
        ("seed", C.c_uint64),
    ]


class GenLayout(C.Structure):
    _fields_ = [
        ("coded_width", C.c_uint32),
        ("upscaled_width", C.c_uint32),
        ("coded_height", C.c_uint32),
        ("superres_denom", C.c_uint32),
        ("plane_width", C.c_uint32 * 3),
        ("plane_height", C.c_uint32 * 3),
        ("cdef_stride", C.c_uint32),
        ("cdef_idx_count", C.c_uint32),
        ("skip8_count", C.c_uint32),
        ("lr_unit_count", C.c_uint32 * 3),
    ]


def lib():
    """The loaded HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        ctxp = C.c_void_p
        L.vaco_av1pf_create.argtypes = [C.c_int, C.c_uint32, C.c_uint32, P(ctxp)]
        L.vaco_av1pf_destroy.argtypes = [ctxp]
        L.vaco_av1pf_destroy.restype = None
        L.vaco_av1pf_last_error.argtypes = [ctxp]
        L.vaco_av1pf_last_error.restype = C.c_char_p
        L.vaco_av1pf_abi_version.restype = C.c_uint32
        L.vaco_av1pf_frame_output_geometry.argtypes = [P(abi.FrameDesc), P(C.c_uint32), P(C.c_uint32)]
        L.vaco_av1pf_frame_input_geometry.argtypes = [P(abi.FrameDesc), P(C.c_uint32), P(C.c_uint32)]
        L.vaco_av1pf_frame.argtypes = [ctxp, P(abi.FrameDesc), P(abi.Picture), P(abi.Picture), P(C.c_uint64)]
        L.vaco_av1pf_frame_device.argtypes = [ctxp, P(abi.FrameDesc), P(abi.Picture), P(abi.Picture), C.c_void_p]
        L.vaco_av1pf_finish.argtypes = [ctxp]
        L.vaco_av1pf_frame_device_static.argtypes = [P(abi.FrameDesc), P(abi.Picture), P(abi.Picture), C.c_void_p,
                                                     P(C.c_char_p)]
        L.vaco_av1pf_cdef_blocks.argtypes = [ctxp, C.c_uint32, P(C.c_uint16), P(abi.FilterParams),
                                             P(C.c_uint8), P(C.c_uint32), P(C.c_uint16)]
        L.vaco_av1pf_upscale_plane.argtypes = [ctxp, P(abi.Plane), P(abi.SrPlaneConfig), P(abi.Plane)]
        L.vaco_av1pf_restore_plane.argtypes = [ctxp, P(abi.Plane), P(abi.Plane), P(abi.LrPlaneConfig),
                                               P(abi.LrUnit), C.c_uint32, P(abi.Plane)]
        L.vaco_av1pf_adjust_strength.argtypes = [C.c_uint16, C.c_uint32]
        L.vaco_av1pf_adjust_strength.restype = C.c_uint16
        L.vaco_av1pf_chroma_direction.argtypes = [C.c_uint8, C.c_uint8, C.c_uint8]
        L.vaco_av1pf_chroma_direction.restype = C.c_uint8
        L.vaco_av1pf_lr_unit_counts.argtypes = [P(abi.LrPlaneConfig), P(C.c_uint32), P(C.c_uint32)]
        L.vaco_av1pf_lr_unit_counts.restype = None
        L.vaco_av1pf_lr_unit_index.argtypes = [P(abi.LrPlaneConfig), C.c_uint32, C.c_uint32]
        L.vaco_av1pf_lr_unit_index.restype = C.c_uint32
        L.vaco_av1pf_skip8_from_mi.argtypes = [P(C.c_uint8), C.c_uint32, C.c_uint32, P(C.c_uint8)]
        L.vaco_av1pf_skip8_from_mi.restype = None
        L.vaco_av1pf_gen_get_layout.argtypes = [P(GenConfig), C.c_uint32, P(GenLayout)]
        L.vaco_av1pf_gen_frame.argtypes = [P(GenConfig), C.c_uint32, P(abi.FrameDesc), P(abi.Picture),
                                           P(C.c_int8), P(C.c_uint8), P(abi.LrUnit), P(abi.LrUnit),
                                           P(abi.LrUnit)]
        L.vaco_av1pf_batch_device.argtypes = [ctxp, C.c_uint32, P(abi.FrameDesc), P(abi.Picture), P(abi.Picture),
                                              C.c_void_p]
        L.vaco_av1pf_profile.argtypes = [ctxp, C.c_int]
        L.vaco_av1pf_profile_read.argtypes = [ctxp, P(C.c_double), P(C.c_uint64), P(C.c_uint64)]
        L.vaco_av1pf_use_graphs.argtypes = [ctxp, C.c_int]
        L.vaco_av1pf_frame_layout.argtypes = [P(abi.FrameDesc), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.vaco_av1pf_frame_device_to_frame.argtypes = [ctxp, P(abi.FrameDesc), P(abi.Picture),
                                                       P(abi.FramePlanes), C.c_void_p]
        L.vaco_av1pf_bits_init.argtypes = [P(abi.BitReader), C.c_void_p, C.c_size_t]
        L.vaco_av1pf_bits_init.restype = None
        L.vaco_av1pf_bits_get.argtypes = [P(abi.BitReader), C.c_uint32]
        L.vaco_av1pf_bits_get.restype = C.c_uint32
        L.vaco_av1pf_parse_cdef_params.argtypes = [P(abi.BitReader), P(abi.SeqInfo), C.c_int, C.c_int,
                                                   C.c_uint32, P(abi.CdefParams)]
        L.vaco_av1pf_parse_cdef_params.restype = C.c_uint32
        L.vaco_av1pf_parse_lr_params.argtypes = [P(abi.BitReader), P(abi.SeqInfo), C.c_int, C.c_int,
                                                 C.c_uint32, P(C.c_uint8), P(C.c_uint32)]
        L.vaco_av1pf_parse_lr_params.restype = None
        L.vaco_av1pf_parse_superres_params.argtypes = [P(abi.BitReader), P(abi.SeqInfo), C.c_uint32,
                                                       P(C.c_uint8)]
        L.vaco_av1pf_parse_superres_params.restype = C.c_uint32
        L.vaco_av1pf_sd_init.argtypes = [P(abi.SymbolDecoder), C.c_void_p, C.c_size_t, C.c_int]
        L.vaco_av1pf_sd_init.restype = None
        L.vaco_av1pf_sd_read_symbol.argtypes = [P(abi.SymbolDecoder), P(C.c_uint16), C.c_uint32]
        L.vaco_av1pf_sd_read_symbol.restype = C.c_uint32
        L.vaco_av1pf_sd_read_bool.argtypes = [P(abi.SymbolDecoder)]
        L.vaco_av1pf_sd_read_bool.restype = C.c_uint32
        L.vaco_av1pf_sd_read_literal.argtypes = [P(abi.SymbolDecoder), C.c_uint32]
        L.vaco_av1pf_sd_read_literal.restype = C.c_uint32
        L.vaco_av1pf_sd_exit.argtypes = [P(abi.SymbolDecoder)]
        L.vaco_av1pf_sd_exit.restype = None
        L.vaco_av1pf_sd_overrun.argtypes = [P(abi.SymbolDecoder)]
        L.vaco_av1pf_sd_overrun.restype = C.c_int
        L.vaco_av1pf_lr_tile_state_init.argtypes = [P(abi.LrTileState)]
        L.vaco_av1pf_lr_tile_state_init.restype = None
        L.vaco_av1pf_read_lr_units.argtypes = [P(abi.SymbolDecoder), P(abi.LrTileState), P(abi.FrameDesc),
                                               C.c_uint32, C.c_uint32, C.c_uint8, P(P(abi.LrUnit)),
                                               P(C.c_char_p)]
        L.vaco_av1pf_read_lr_units.restype = C.c_int
        L.vaco_av1pf_read_cdef.argtypes = [P(abi.SymbolDecoder), C.c_int, C.c_uint32, P(C.c_int8), C.c_uint32,
                                           C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint8, C.c_int]
        L.vaco_av1pf_read_cdef.restype = None
        L.vaco_av1pf_set_chain_spin_limit.argtypes = [ctxp, C.c_uint32]
        L.vaco_av1pf_set_chain_spin_limit.restype = C.c_int
        L.vaco_av1pf_get_tuning.argtypes = [ctxp, P(abi.Tuning)]
        L.vaco_av1pf_set_tuning.argtypes = [ctxp, P(abi.Tuning)]
        L.vaco_av1pf_autotune.argtypes = [ctxp, C.c_uint32, P(abi.FrameDesc), P(abi.Picture), P(abi.Picture),
                                          C.c_uint32, P(abi.Tuning)]
        L.vaco_av1pf_server_create.argtypes = [ctxp, P(abi.ServerConfig), P(C.c_void_p)]
        L.vaco_av1pf_server_submit.argtypes = [C.c_void_p, P(abi.FrameDesc), P(abi.Picture), P(abi.Picture),
                                               C.c_uint64]
        L.vaco_av1pf_server_poll.argtypes = [C.c_void_p, P(abi.Completion), C.c_uint32]
        L.vaco_av1pf_server_poll.restype = C.c_uint32
        L.vaco_av1pf_server_drain.argtypes = [C.c_void_p]
        L.vaco_av1pf_server_destroy.argtypes = [C.c_void_p]
        L.vaco_av1pf_server_destroy.restype = None
        L.vaco_av1pf_now_ns.restype = C.c_uint64
        L.vaco_av1pf_server_frame_bytes.argtypes = [P(abi.FrameDesc)]
        L.vaco_av1pf_server_frame_bytes.restype = C.c_uint64
        for name in ("vaco_av1pf_frame_layout", "vaco_av1pf_frame_device_to_frame", "vaco_av1pf_use_graphs", "vaco_av1pf_get_tuning", "vaco_av1pf_set_tuning", "vaco_av1pf_autotune",
                     "vaco_av1pf_server_create", "vaco_av1pf_server_submit", "vaco_av1pf_server_drain",
                     "vaco_av1pf_batch_device", "vaco_av1pf_profile", "vaco_av1pf_profile_read", "vaco_av1pf_create", "vaco_av1pf_frame", "vaco_av1pf_frame_device", "vaco_av1pf_finish",
                     "vaco_av1pf_cdef_blocks", "vaco_av1pf_upscale_plane", "vaco_av1pf_restore_plane",
                     "vaco_av1pf_frame_output_geometry", "vaco_av1pf_frame_input_geometry",
                     "vaco_av1pf_frame_device_static",
                     "vaco_av1pf_gen_get_layout", "vaco_av1pf_gen_frame"):
            getattr(L, name).restype = C.c_int
        if L.vaco_av1pf_abi_version() != abi.ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.vaco_av1pf_abi_version()}, "
                               f"this binding expects {abi.ABI_VERSION}: rebuild it")
        _lib = L
    return _lib


# --------------------------------------------------------------- frames --

class Frame:
    """A frame's input planes and side information (host memory).

    ``desc`` is the ``vaco_av1pf_frame_desc`` pointing into the arrays held
    here, so the object must outlive every call that uses ``desc``.
    """

    def __init__(self, planes, desc: abi.FrameDesc, keep=()):
        self.planes = planes
        self.desc = desc
        self._keep = list(keep)

    @property
    def nplanes(self) -> int:
        return 1 if self.desc.monochrome else 3

    def output_geometry(self):
        w, h = (C.c_uint32 * 3)(), (C.c_uint32 * 3)()
        lib().vaco_av1pf_frame_output_geometry(C.byref(self.desc), w, h)
        return [(h[i], w[i]) for i in range(self.nplanes)]

    @property
    def output_luma_pixels(self) -> int:
        return int(self.desc.upscaled_width) * int(self.desc.coded_height)


def frame_layout(frame: "Frame"):
    """pic_to_frame's planes for `frame`: ([(height, width)], bytes_per_sample)."""
    w, h, b = (C.c_uint32 * 3)(), (C.c_uint32 * 3)(), C.c_uint32()
    st = lib().vaco_av1pf_frame_layout(C.byref(frame.desc), w, h, C.byref(b))
    if st != abi.OK:
        raise Av1Error(st, "vaco_av1pf_frame_layout failed")
    return [(h[i], w[i]) for i in range(frame.nplanes)], b.value


def generate(cfg: GenConfig, frame_index: int = 0) -> Frame:
    """Seeded synthetic frame `frame_index` of workload `cfg`."""
    L = lib()
    lay = GenLayout()
    st = L.vaco_av1pf_gen_get_layout(C.byref(cfg), frame_index, C.byref(lay))
    if st != abi.OK:
        raise Av1Error(st, "invalid generator configuration")
    np_ = 1 if cfg.monochrome else 3
    planes = [np.zeros((lay.plane_height[p], lay.plane_width[p]), np.uint16) for p in range(np_)]
    idx = np.zeros(max(1, lay.cdef_idx_count), np.int8)
    skip = np.zeros(max(1, lay.skip8_count), np.uint8)
    units = [(abi.LrUnit * max(1, lay.lr_unit_count[p]))() for p in range(3)]
    desc = abi.FrameDesc()
    pic = abi.picture_of(planes)
    st = L.vaco_av1pf_gen_frame(C.byref(cfg), frame_index, C.byref(desc), C.byref(pic),
                                idx.ctypes.data_as(C.POINTER(C.c_int8)),
                                skip.ctypes.data_as(C.POINTER(C.c_uint8)), units[0], units[1], units[2])
    if st != abi.OK:
        raise Av1Error(st, "generator failed")
    return Frame(planes, desc, keep=(idx, skip, units))


# -------------------------------------------------------------- context --

class Context:
    """A device context (``vaco_av1pf_ctx``) sized for frames up to max_w x max_h."""

    def __init__(self, max_width: int = 3840, max_height: int = 2160, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        st = self._lib.vaco_av1pf_create(device, max_width, max_height, C.byref(h))
        if st != abi.OK:
            raise Av1Error(st, "vaco_av1pf_create failed")
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.vaco_av1pf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int):
        if st != abi.OK:
            msg = self._lib.vaco_av1pf_last_error(self.handle)
            raise Av1Error(st, msg.decode() if msg else "")

    # ---- whole stage

    def frame(self, frame: Frame, fuel: int | None = None):
        """The stage on host planes; returns the output planes (and fuel left)."""
        outs = [np.zeros(g, np.uint16) for g in frame.output_geometry()]
        pin = abi.picture_of([np.ascontiguousarray(p) for p in frame.planes])
        pout = abi.picture_of(outs)
        f = C.c_uint64(fuel) if fuel is not None else None
        self._check(self._lib.vaco_av1pf_frame(self.handle, C.byref(frame.desc), C.byref(pin), C.byref(pout),
                                               C.byref(f) if f is not None else None))
        if fuel is not None:
            return outs, f.value
        return outs

    def device_call(self, frame: Frame, in_ptrs, out_ptrs, stream: int | None = None,
                    in_strides=None, out_strides=None):
        """A zero-argument callable that enqueues the stage on these device
        planes (integer device addresses; dense rows unless row strides, in
        samples, are given). The ctypes arguments are built once, so repeated
        calls cost one foreign call each."""
        geo_in = [(p.shape[0], p.shape[1]) for p in frame.planes]
        geo_out = frame.output_geometry()
        pin, pout = abi.Picture(), abi.Picture()
        for i in range(frame.nplanes):
            pin.plane[i] = abi.device_plane(in_ptrs[i], geo_in[i][1], geo_in[i][0],
                                            in_strides[i] if in_strides else None)
            pout.plane[i] = abi.device_plane(out_ptrs[i], geo_out[i][1], geo_out[i][0],
                                             out_strides[i] if out_strides else None)
        fn, h = self._lib.vaco_av1pf_frame_device, self.handle
        d, pi, po, st = C.byref(frame.desc), C.byref(pin), C.byref(pout), C.c_void_p(stream or 0)
        check = self._check

        def call():
            rc = fn(h, d, pi, po, st)
            if rc:
                check(rc)

        call.keep = (frame, pin, pout)
        return call

    def frame_device(self, frame: Frame, in_ptrs, out_ptrs, stream: int | None = None):
        """The stage on device planes (integer device addresses, dense rows)."""
        self.device_call(frame, in_ptrs, out_ptrs, stream)()

    def batch_call(self, frames, in_ptrs, out_ptrs, stream: int | None = None):
        """A zero-argument callable enqueueing one batched call
        (vaco_av1pf_batch_device) over `frames` on these device planes."""
        n = len(frames)
        descs = (abi.FrameDesc * n)()
        pins, pouts = (abi.Picture * n)(), (abi.Picture * n)()
        for k, f in enumerate(frames):
            descs[k] = f.desc
            geo_out = f.output_geometry()
            for i in range(f.nplanes):
                h, w = f.planes[i].shape
                pins[k].plane[i] = abi.device_plane(in_ptrs[k][i], w, h)
                pouts[k].plane[i] = abi.device_plane(out_ptrs[k][i], geo_out[i][1], geo_out[i][0])
        fn, h = self._lib.vaco_av1pf_batch_device, self.handle
        st = C.c_void_p(stream or 0)
        check = self._check

        def call():
            rc = fn(h, n, descs, pins, pouts, st)
            if rc:
                check(rc)

        call.keep = (list(frames), descs, pins, pouts)
        return call

    def frame_output_call(self, frame: "Frame", in_ptrs, out_ptrs, strides_bytes, stream: int | None = None):
        """A zero-argument callable running the stage on device planes into
        pic_to_frame's frame planes (vaco_av1pf_frame_device_to_frame); see
        frame_layout(). The ctypes arguments are built once."""
        pin = abi.Picture()
        for i in range(frame.nplanes):
            h, w = frame.planes[i].shape
            pin.plane[i] = abi.device_plane(in_ptrs[i], w, h)
        fo = abi.FramePlanes()
        for i in range(frame.nplanes):
            fo.data[i] = out_ptrs[i]
            fo.stride_bytes[i] = strides_bytes[i]
        fn, h = self._lib.vaco_av1pf_frame_device_to_frame, self.handle
        d, pi, po, st = C.byref(frame.desc), C.byref(pin), C.byref(fo), C.c_void_p(stream or 0)
        check = self._check

        def call():
            rc = fn(h, d, pi, po, st)
            if rc:
                check(rc)

        call.keep = (frame, pin, fo)
        return call

    def frame_device_to_frame(self, frame: "Frame", in_ptrs, out_ptrs, strides_bytes, stream: int | None = None):
        self.frame_output_call(frame, in_ptrs, out_ptrs, strides_bytes, stream)()

    def use_graphs(self, enable: bool = True):
        """Replay each call's kernel chain as a hipGraph per launch shape."""
        self._check(self._lib.vaco_av1pf_use_graphs(self.handle, 1 if enable else 0))

    # ---- tuning

    TUNING_FIELDS = ("superres_rows", "cdef_xcd_order", "lr_order", "lr_tile_rows", "cdef_waves", "chain")

    @property
    def tuning(self) -> dict:
        t = abi.Tuning()
        self._check(self._lib.vaco_av1pf_get_tuning(self.handle, C.byref(t)))
        return {k: getattr(t, k) for k in self.TUNING_FIELDS}

    @tuning.setter
    def tuning(self, values: dict):
        t = abi.Tuning()
        self._check(self._lib.vaco_av1pf_get_tuning(self.handle, C.byref(t)))
        for k, v in values.items():
            if k not in self.TUNING_FIELDS:
                raise KeyError(k)
            setattr(t, k, int(v))
        st = self._lib.vaco_av1pf_set_tuning(self.handle, C.byref(t))
        if st != abi.OK:
            raise Av1Error(st, f"invalid tuning {values}")

    def set_chain_spin_limit(self, polls: int):
        """Polls a chained launch's restoration tile makes before giving up (test hook)."""
        self._check(self._lib.vaco_av1pf_set_chain_spin_limit(self.handle, polls))

    def autotune(self, frames, in_ptrs, out_ptrs, reps: int = 3) -> dict:
        """Time every launch variant on these frames (device planes; the
        median of `reps` runs per variant) and keep the fastest per kernel;
        returns the chosen tuning."""
        n = len(frames)
        descs = (abi.FrameDesc * n)()
        pins, pouts = (abi.Picture * n)(), (abi.Picture * n)()
        for k, f in enumerate(frames):
            descs[k] = f.desc
            geo_out = f.output_geometry()
            for i in range(f.nplanes):
                h, w = f.planes[i].shape
                pins[k].plane[i] = abi.device_plane(in_ptrs[k][i], w, h)
                pouts[k].plane[i] = abi.device_plane(out_ptrs[k][i], geo_out[i][1], geo_out[i][0])
        t = abi.Tuning()
        self._check(self._lib.vaco_av1pf_autotune(self.handle, n, descs, pins, pouts, reps, C.byref(t)))
        return {k: getattr(t, k) for k in self.TUNING_FIELDS}

    def batch_device(self, frames, in_ptrs, out_ptrs, stream: int | None = None):
        self.batch_call(frames, in_ptrs, out_ptrs, stream)()

    def finish(self):
        self._check(self._lib.vaco_av1pf_finish(self.handle))

    # ---- kernel timing

    KERNELS = ("cdef", "superres", "restoration", "upload", "chain")

    def profile(self, enable: bool = True):
        self._check(self._lib.vaco_av1pf_profile(self.handle, 1 if enable else 0))

    def profile_read(self):
        """{kernel: (total_ms, launches, algorithmic_bytes)} since the last read (synchronises)."""
        k = len(self.KERNELS)
        ms = (C.c_double * k)()
        n = (C.c_uint64 * k)()
        b = (C.c_uint64 * k)()
        self._check(self._lib.vaco_av1pf_profile_read(self.handle, ms, n, b))
        return {name: (ms[i], n[i], b[i]) for i, name in enumerate(self.KERNELS)}

    # ---- kernel-level entries

    def cdef_blocks(self, inputs: np.ndarray, params):
        """find_direction + filter_block for n cases: -> (dirs, vars, outs[n, 64])."""
        n = len(params)
        inp = np.ascontiguousarray(inputs, dtype=np.uint16).reshape(n, 144)
        par = (abi.FilterParams * max(1, n))(*params)
        dirs = np.zeros(n, np.uint8)
        vars_ = np.zeros(n, np.uint32)
        outs = np.zeros((n, 64), np.uint16)
        self._check(self._lib.vaco_av1pf_cdef_blocks(
            self.handle, n, abi.u16_ptr(inp), par, dirs.ctypes.data_as(C.POINTER(C.c_uint8)),
            vars_.ctypes.data_as(C.POINTER(C.c_uint32)), abi.u16_ptr(outs)))
        return dirs, vars_, outs

    def upscale_plane(self, inp: np.ndarray, visible_width: int, output_width: int, height: int,
                      bit_depth: int) -> np.ndarray:
        i = np.ascontiguousarray(inp, dtype=np.uint16)
        out = np.zeros((max(height, 1), max(output_width, 1)), np.uint16)
        cfg = abi.SrPlaneConfig(visible_width, output_width, height, bit_depth)
        pi, po = abi.plane_of(i), abi.plane_of(out)
        po.width, po.height = output_width, height
        self._check(self._lib.vaco_av1pf_upscale_plane(self.handle, C.byref(pi), C.byref(cfg), C.byref(po)))
        return out[:height, :output_width]

    def restore_plane(self, before: np.ndarray, after: np.ndarray, cfg: abi.LrPlaneConfig, units) -> np.ndarray:
        b = np.ascontiguousarray(before, dtype=np.uint16)
        a = np.ascontiguousarray(after, dtype=np.uint16)
        out = np.zeros((max(cfg.height, 1), max(cfg.width, 1)), np.uint16)
        po = abi.plane_of(out)
        po.width, po.height = cfg.width, cfg.height
        arr = abi.units_array(units)
        self._check(self._lib.vaco_av1pf_restore_plane(self.handle, C.byref(abi.plane_of(b)),
                                                       C.byref(abi.plane_of(a)), C.byref(cfg), arr,
                                                       len(units), C.byref(po)))
        return out[: cfg.height, : cfg.width]


# --------------------------------------------------------------- server --

class Server:
    """The serving loop (``vaco_av1pf_server``) over a Context it owns while open.

    submit() queues one frame on device planes; poll() returns completions
    as ``(tag, status, batch_size, t_submit_ns, t_launch_ns, t_done_ns,
    t_gpu_start_ns, t_gpu_end_ns)``.
    """

    def __init__(self, ctx: Context, hbm_budget_bytes: int = 8 << 30, window_us: int = 2000,
                 max_batch: int = 64, max_inflight: int = 2):
        self._lib = lib()
        self.ctx = ctx
        cfg = abi.ServerConfig(hbm_budget_bytes, window_us, max_batch, max_inflight)
        h = C.c_void_p()
        st = self._lib.vaco_av1pf_server_create(ctx.handle, C.byref(cfg), C.byref(h))
        if st != abi.OK:
            raise Av1Error(st, "vaco_av1pf_server_create failed")
        self.handle = h
        self._keep = {}
        self._buf = (abi.Completion * 256)()

    def submit_call(self, frame: Frame, in_ptrs, out_ptrs, tag: int = 0):
        """A callable submitting `frame` on these device planes, with `tag` or
        the tag it is called with. The ctypes arguments are built once."""
        geo_out = frame.output_geometry()
        pin, pout = abi.Picture(), abi.Picture()
        for i in range(frame.nplanes):
            h, w = frame.planes[i].shape
            pin.plane[i] = abi.device_plane(in_ptrs[i], w, h)
            pout.plane[i] = abi.device_plane(out_ptrs[i], geo_out[i][1], geo_out[i][0])
        fn, h = self._lib.vaco_av1pf_server_submit, self.handle
        d, pi, po = C.byref(frame.desc), C.byref(pin), C.byref(pout)

        def call(t: int = tag):
            rc = fn(h, d, pi, po, t)
            if rc:
                raise Av1Error(rc, "vaco_av1pf_server_submit failed")

        call.keep = (frame, pin, pout)
        return call

    def submit(self, frame: Frame, in_ptrs, out_ptrs, tag: int):
        c = self.submit_call(frame, in_ptrs, out_ptrs, tag)
        self._keep[tag] = c
        c()

    def poll(self):
        out = []
        while True:
            n = self._lib.vaco_av1pf_server_poll(self.handle, self._buf, 256)
            for i in range(n):
                c = self._buf[i]
                self._keep.pop(c.tag, None)
                out.append((c.tag, c.status, c.batch_size, c.t_submit_ns, c.t_launch_ns, c.t_done_ns,
                            c.t_gpu_start_ns, c.t_gpu_end_ns))
            if n < 256:
                return out

    def drain(self):
        st = self._lib.vaco_av1pf_server_drain(self.handle)
        if st != abi.OK:
            raise Av1Error(st, "vaco_av1pf_server_drain failed")

    def close(self):
        if self.handle:
            self._lib.vaco_av1pf_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def frame_device_static(frame: Frame, in_ptrs, out_ptrs, stream: int | None = None):
    """vaco_av1pf_frame_device_static: the stage on device planes (integer
    addresses, dense rows) without a context or any allocation; synchronous.
    Raises Av1Error with the reference's message on failure."""
    geo_out = frame.output_geometry()
    pin, pout = abi.Picture(), abi.Picture()
    for i in range(frame.nplanes):
        h, w = frame.planes[i].shape
        pin.plane[i] = abi.device_plane(in_ptrs[i], w, h)
        pout.plane[i] = abi.device_plane(out_ptrs[i], geo_out[i][1], geo_out[i][0])
    msg = C.c_char_p()
    rc = lib().vaco_av1pf_frame_device_static(C.byref(frame.desc), C.byref(pin), C.byref(pout),
                                              C.c_void_p(stream or 0), C.byref(msg))
    if rc != abi.OK:
        raise Av1Error(rc, (msg.value or b"").decode())


def frame_bytes(frame: Frame) -> int:
    """Device bytes `frame` holds in flight, as the server's admission counts them."""
    return lib().vaco_av1pf_server_frame_bytes(C.byref(frame.desc))


def now_ns() -> int:
    return lib().vaco_av1pf_now_ns()


# ------------------------------------------------ parameter producers --

class Bits:
    """An MSB-first bit reader over `data` (vaco_av1pf_bit_reader)."""

    def __init__(self, data: bytes):
        self._buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(bytes(data) or b"\0")
        self.r = abi.BitReader()
        lib().vaco_av1pf_bits_init(C.byref(self.r), C.cast(self._buf, C.c_void_p) if data else None, len(data))

    @property
    def pos(self) -> int:
        return self.r.pos

    @property
    def overrun(self) -> bool:
        return bool(self.r.overrun)

    def get(self, n: int) -> int:
        return lib().vaco_av1pf_bits_get(C.byref(self.r), n)


def seq_info(cdef=False, restoration=False, superres=False, sb128=False, ssx=True, ssy=True) -> abi.SeqInfo:
    return abi.SeqInfo(int(cdef), int(restoration), int(superres), int(sb128), int(ssx), int(ssy))


def parse_cdef_params(bits: Bits, seq: abi.SeqInfo, coded_lossless=False, allow_intrabc=False, num_planes=3):
    """-> (cdef_bits, CdefParams)"""
    out = abi.CdefParams()
    n = lib().vaco_av1pf_parse_cdef_params(C.byref(bits.r), C.byref(seq), int(coded_lossless),
                                           int(allow_intrabc), num_planes, C.byref(out))
    return n, out


def parse_lr_params(bits: Bits, seq: abi.SeqInfo, all_lossless=False, allow_intrabc=False, num_planes=3):
    """-> (types[3], unit_size[3])"""
    t = (C.c_uint8 * 3)()
    u = (C.c_uint32 * 3)()
    lib().vaco_av1pf_parse_lr_params(C.byref(bits.r), C.byref(seq), int(all_lossless), int(allow_intrabc),
                                     num_planes, t, u)
    return list(t), list(u)


def parse_superres_params(bits: Bits, seq: abi.SeqInfo, upscaled_width: int):
    """-> (coded_width, use_superres)"""
    use = C.c_uint8()
    w = lib().vaco_av1pf_parse_superres_params(C.byref(bits.r), C.byref(seq), upscaled_width, C.byref(use))
    return w, bool(use.value)


class SymbolDecoder:
    """The AV1 symbol decoder over one tile's bytes (vaco_av1pf_symbol_decoder)."""

    def __init__(self, data: bytes, disable_cdf_update: bool):
        self._buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(bytes(data) or b"\0")
        self.sd = abi.SymbolDecoder()
        lib().vaco_av1pf_sd_init(C.byref(self.sd), C.cast(self._buf, C.c_void_p) if data else None, len(data),
                                 int(disable_cdf_update))

    def read_symbol(self, cdf) -> int:
        """cdf: a ctypes uint16 array of n + 1 entries, adapted in place."""
        return lib().vaco_av1pf_sd_read_symbol(C.byref(self.sd), cdf, len(cdf) - 1)

    def read_bool(self) -> int:
        return lib().vaco_av1pf_sd_read_bool(C.byref(self.sd))

    def read_literal(self, n: int) -> int:
        return lib().vaco_av1pf_sd_read_literal(C.byref(self.sd), n)

    def exit(self) -> None:
        lib().vaco_av1pf_sd_exit(C.byref(self.sd))

    @property
    def overrun(self) -> bool:
        return bool(lib().vaco_av1pf_sd_overrun(C.byref(self.sd)))

    @property
    def max_bits(self) -> int:
        return self.sd.max_bits


def cdf_array(values) -> C.Array:
    """A cdf for SymbolDecoder.read_symbol: the thresholds then the counter."""
    return (C.c_uint16 * len(values))(*values)


def lr_tile_state() -> abi.LrTileState:
    st = abi.LrTileState()
    lib().vaco_av1pf_lr_tile_state_init(C.byref(st))
    return st


def read_lr_units(sd: SymbolDecoder, st: abi.LrTileState, desc: abi.FrameDesc, mi_row: int, mi_col: int,
                  b_size: int, units) -> None:
    """vaco_av1pf_read_lr_units; units: three ctypes LrUnit arrays (or None),
    filled in place. Raises Av1Error with the reference's message."""
    ptrs = (C.POINTER(abi.LrUnit) * 3)(*[C.cast(u, C.POINTER(abi.LrUnit)) if u is not None else None
                                        for u in units])
    msg = C.c_char_p()
    rc = lib().vaco_av1pf_read_lr_units(C.byref(sd.sd), C.byref(st), C.byref(desc), mi_row, mi_col, b_size,
                                         ptrs, C.byref(msg))
    if rc:
        raise Av1Error(rc, (msg.value or b"").decode())


def read_cdef(sd: SymbolDecoder, enabled: bool, cdef_bits: int, cdef_idx: np.ndarray, mi_row: int, mi_col: int,
              mi_size: int, skip: bool) -> None:
    """vaco_av1pf_read_cdef on an int8 (rows, stride) map, in place."""
    assert cdef_idx.dtype == np.int8 and cdef_idx.flags.c_contiguous
    rows, stride = cdef_idx.shape
    lib().vaco_av1pf_read_cdef(C.byref(sd.sd), int(enabled), cdef_bits,
                               cdef_idx.ctypes.data_as(C.POINTER(C.c_int8)), stride, rows, mi_row, mi_col,
                               mi_size, int(skip))


# --------------------------------------------------------- host helpers --

def adjust_strength(primary: int, variance: int) -> int:
    return lib().vaco_av1pf_adjust_strength(primary, variance)


def chroma_direction(direction: int, sub_x: bool, sub_y: bool) -> int:
    return lib().vaco_av1pf_chroma_direction(direction, int(sub_x), int(sub_y))


def lr_unit_counts(cfg: abi.LrPlaneConfig):
    c, r = C.c_uint32(), C.c_uint32()
    lib().vaco_av1pf_lr_unit_counts(C.byref(cfg), C.byref(c), C.byref(r))
    return c.value, r.value


def lr_unit_index(cfg: abi.LrPlaneConfig, x: int, y: int) -> int:
    return lib().vaco_av1pf_lr_unit_index(C.byref(cfg), x, y)


def skip8_from_mi(mi_skip: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(mi_skip, dtype=np.uint8)
    rows, cols = m.shape
    out = np.zeros((max(1, rows // 2), max(1, cols // 2)), np.uint8)
    lib().vaco_av1pf_skip8_from_mi(m.ctypes.data_as(C.POINTER(C.c_uint8)), rows, cols,
                                   out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out
