"""ctypes binding of libgpurir.so (the C ABI declared in include/gpurir.h).

The library is built in place by gpurir_amd/build.py; importing this module
fails loudly if it is missing (there is no CPU fallback).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpurir.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)


class Options(ctypes.Structure):
    """gpurir_options_t"""
    _fields_ = [("precision", ctypes.c_int), ("noise", ctypes.c_int), ("seed", ctypes.c_ulonglong),
                ("call_offset", ctypes.c_uint), ("room_index", ctypes.c_uint), ("level", ctypes.c_float),
                ("chunk_images", ctypes.c_int), ("tile_samples", ctypes.c_int), ("persistent", ctypes.c_int),
                ("crossfade", ctypes.c_int), ("method", ctypes.c_int)]


PREC_FP32, PREC_LUT, PREC_FP16, PREC_BF16 = 0, 1, 2, 3
NOISE_LOGISTIC, NOISE_GAUSSIAN = 0, 1
METHOD_AUTO, METHOD_TAPS, METHOD_BASIS = 0, 1, 2

# every symbol of include/gpurir.h with (restype, argtypes)
_SIG = {
    "gpurir_version": (ctypes.c_int, []),
    "gpurir_last_error": (ctypes.c_char_p, []),
    "gpurir_n_samples": (ctypes.c_int, [ctypes.c_float, ctypes.c_float]),
    "gpurir_activate_mixed_precision": (ctypes.c_int, [ctypes.c_int]),
    "gpurir_activate_lut": (ctypes.c_int, [ctypes.c_int]),
    "gpurir_set_seed": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_uint]),
    # host arrays as addresses (c_void_p): the call sits on the per-RIR latency path
    "gpurir_simulate_rir": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]),
    "gpurir_simulate_rir_device": (ctypes.c_int, [c_float_p, c_float_p, ctypes.c_void_p, ctypes.c_int,
                                                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_float,
                                                  ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                  ctypes.POINTER(Options), ctypes.c_void_p, ctypes.c_void_p]),
    "gpurir_simulate_rir_device_noalloc": (ctypes.c_int, [c_float_p, c_float_p, ctypes.c_void_p, ctypes.c_int,
                                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_int_p,
                                                          ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                          ctypes.c_float, ctypes.POINTER(Options), ctypes.c_void_p,
                                                          ctypes.c_void_p]),
    "gpurir_image_list_device": (ctypes.c_int, [c_float_p, c_float_p, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_float, ctypes.c_float,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gpurir_simulate_batch_device": (ctypes.c_int, [ctypes.c_int, c_float_p, c_float_p, c_int_p, c_int_p, c_int_p,
                                                    c_float_p, c_float_p, c_float_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                                    ctypes.POINTER(Options), ctypes.c_void_p, c_ll_p,
                                                    ctypes.c_void_p]),
    "gpurir_batch_plan_create": (ctypes.c_int, [ctypes.c_int, c_float_p, c_float_p, c_int_p, c_int_p, c_int_p,
                                                c_float_p, c_float_p, c_float_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_float, ctypes.c_float, ctypes.POINTER(Options),
                                                ctypes.c_void_p, c_ll_p, ctypes.POINTER(ctypes.c_void_p)]),
    "gpurir_batch_plan_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gpurir_batch_plan_out_elems": (ctypes.c_longlong, [ctypes.c_void_p]),
    "gpurir_batch_plan_n_rirs": (ctypes.c_longlong, [ctypes.c_void_p]),
    "gpurir_batch_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "gpurir_batch_plan_bind_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gpurir_batch_plan_count_taps": (ctypes.c_int, [ctypes.c_void_p, c_ll_p, ctypes.c_void_p]),
    "gpurir_batch_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_int_p, c_int_p, c_int_p,
                                                  c_float_p]),
    "gpurir_batch_plan_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_float_p, c_float_p]),
    "gpurir_simulate_batch_adjoint_device": (ctypes.c_int, [ctypes.c_int, c_float_p, c_float_p, c_int_p, c_int_p,
                                                            c_int_p, c_float_p, c_float_p, c_float_p,
                                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                            ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                            ctypes.c_float, ctypes.c_float, ctypes.POINTER(Options),
                                                            ctypes.c_void_p, ctypes.c_void_p, c_ll_p,
                                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                            ctypes.c_void_p, ctypes.c_void_p]),
    "gpurir_batch_plan_adjoint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gpurir_conv": (ctypes.c_int, [c_float_p, ctypes.c_int, ctypes.c_int, c_float_p, ctypes.c_int, ctypes.c_int,
                                   c_float_p]),
    "gpurir_trajectory_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "gpurir_trajectory_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_void_p,
                                                     ctypes.POINTER(ctypes.c_void_p)]),
    "gpurir_trajectory_plan_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gpurir_trajectory_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
}

SYMBOLS = tuple(_SIG)

_lib = None


def load():
    """Load libgpurir.so (once) and declare every entry point."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gpurir_amd.build` "
                              "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class GpuRIRError(RuntimeError):
    pass


def check(rc):
    """Turn a negative return code into an exception carrying gpurir_last_error()."""
    if rc < 0:
        msg = load().gpurir_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise GpuRIRError(msg)
    return rc


def fptr(a):
    return a.ctypes.data_as(c_float_p)


def iptr(a):
    return a.ctypes.data_as(c_int_p)


def addr(a):
    """Address of a C-contiguous numpy array's first element (for c_void_p arguments)."""
    return a.__array_interface__["data"][0]
