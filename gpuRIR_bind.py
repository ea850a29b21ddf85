"""The compiled module of gpuRIR, ``gpuRIR_bind`` (src/python_bind.cpp:130-143), over the C ABI
of libgpurir.so (include/gpurir.h) with ctypes.

A maintainer who keeps the reference's own gpuRIR/__init__.py imports this module in place of
the pybind11 extension: ``from gpuRIR_bind import gpuRIR_bind`` then works unchanged, and every
call of the reference module lands in the gfx950 kernels.  The class keeps the extension's
interface:

  gpuRIR_bind(mixed_precision=False)
  .simulateRIR_bind(room_sz, beta, pos_src, pos_rcv, orV_src, orV_rcv, spkr_pattern, mic_pattern,
                    nb_img, Tdiff, Tmax, Fs, c=343.0)  -> float32 [M_src, M_rcv, nSamples]
  .gpu_conv(source_segments, RIR)                       -> float32 [M_src, M_rcv, seg + len - 1]
  .activate_mixed_precision_bind(activate) -> bool
  .activate_lut_bind(activate)             -> bool
  .mixed_precision, .lookup_table          (the instance's flags)

Each instance has its own flags, as the reference's simulator object does
(src/gpuRIR_cuda.cu:1025-1073); they are handed to the library with each call.  All
instances share the library's noise stream (Philox4x32-10 keyed by 1234, one counter step
per call), as the reference's instances share one cuRAND generator seeded with 1234.
Errors raise instead of ending the process (the reference calls exit() on CUDA errors).
"""
import ctypes

import numpy as np

from gpurir_amd import _lib

__all__ = ["gpuRIR_bind"]


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class gpuRIR_bind:
    """pybind11 class gpuRIR_bind (src/python_bind.cpp:13-26, :130-143)."""

    def __init__(self, mixed_precision=False):
        self._lib = _lib.load()
        self.mixed_precision = False
        self.lookup_table = False
        # the reference's constructor: activate_mixed_precision(mPrecision), then activate_lut(true)
        # (src/gpuRIR_cuda.cu:1031-1033; python_bind.cpp:15 defaults lut to true)
        self.activate_mixed_precision_bind(mixed_precision)
        self.activate_lut_bind(True)

    # -- toggles: src/gpuRIR_cuda.cu:1050-1073, same messages and mutual exclusion
    def activate_mixed_precision_bind(self, activate):
        activate = bool(activate)
        if activate and self.lookup_table:
            print("The mixed precision implementation is not compatible with the lookup table.")
            print("Disabling the lookup table.")
            self.lookup_table = False
        self.mixed_precision = activate
        return self.mixed_precision

    def activate_lut_bind(self, activate):
        activate = bool(activate)
        if activate and self.mixed_precision:
            print("The lookup table is not compatible with the mixed precision implementation.")
            print("Disabling the mixed precision implementation.")
            self.mixed_precision = False
        self.lookup_table = activate
        return self.lookup_table

    def _apply_flags(self):
        # hand this instance's flags to the library: switching both off first prints nothing
        lib = self._lib
        lib.gpurir_activate_mixed_precision(0)
        lib.gpurir_activate_lut(0)
        if self.mixed_precision:
            lib.gpurir_activate_mixed_precision(1)
        elif self.lookup_table:
            lib.gpurir_activate_lut(1)

    def simulateRIR_bind(self, room_sz, beta, pos_src, pos_rcv, orV_src, orV_rcv, spkr_pattern, mic_pattern,
                         nb_img, Tdiff, Tmax, Fs, c=343.0):
        """src/python_bind.cpp:29-88: arrays are float32 C order, patterns are the integer codes."""
        room = _f32(room_sz).reshape(-1)
        b = _f32(beta).reshape(-1)
        nb = np.ascontiguousarray(np.asarray(nb_img, dtype=np.int32).reshape(-1))
        ps, pr, os_, or_ = _f32(pos_src), _f32(pos_rcv), _f32(orV_src), _f32(orV_rcv)
        # the reference's size checks (python_bind.cpp:47-59), raised instead of compiled out
        if room.size != 3 or b.size != 6 or nb.size != 3:
            raise ValueError("room_sz, beta and nb_img must have 3, 6 and 3 elements")
        for a in (ps, pr, os_, or_):
            if a.ndim != 2 or a.shape[1] != 3:
                raise ValueError("positions and orientations must be [M, 3] arrays")
        if ps.shape[0] != os_.shape[0] or pr.shape[0] != or_.shape[0]:
            raise ValueError("each source and receiver needs one orientation vector")
        self._apply_flags()
        lib = self._lib
        ns = lib.gpurir_n_samples(float(Tmax), float(Fs))
        out = np.empty((ps.shape[0], pr.shape[0], ns), dtype=np.float32)
        addr = _lib.addr
        _lib.check(lib.gpurir_simulate_rir(addr(room), addr(b), addr(ps), ps.shape[0], addr(pr), pr.shape[0],
                                           addr(os_), addr(or_), int(spkr_pattern), int(mic_pattern), addr(nb),
                                           float(Tdiff), float(Tmax), float(Fs), float(c), addr(out)))
        return out

    def gpu_conv(self, source_segments, RIR):
        """src/python_bind.cpp:90-119: linear convolution of segment p with RIR[p, r] for every
        (p, r); float32 [M_src, M_rcv, segment_len + RIR_len - 1]."""
        seg = _f32(source_segments)
        rir = _f32(RIR)
        if seg.ndim != 2 or rir.ndim != 3 or seg.shape[0] != rir.shape[0]:
            raise ValueError("source_segments must be [M, len] and RIR [M, M_rcv, len] with the same M")
        P, S = seg.shape
        _, R, L = rir.shape
        out = np.empty((P, R, S + L - 1), dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _lib.check(self._lib.gpurir_conv(seg.ctypes.data_as(fp), P, S, rir.ctypes.data_as(fp), R, L,
                                         out.ctypes.data_as(fp)))
        return out
