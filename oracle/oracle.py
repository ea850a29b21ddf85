"""FP64 CPU oracle for gpuRIR's simulation path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; the gpurir_amd package never does.  The heavy part lives in
rir_oracle.c (ctypes); the trajectory convolution is restated here with numpy.

Restated reference code (DavidDiazGuerra/gpuRIR v1.2.0):
  simulate_rir   -> src/gpuRIR_cuda.cu:766-931 (+ kernels :110-147, :245-461)
  images         -> src/gpuRIR_cuda.cu:245-375 (calcAmpTau_kernel)
  lut_table      -> src/gpuRIR_cuda.cu:578-590, :632-633
  gpu_conv       -> src/gpuRIR_cuda.cu:933-1023 (linear convolution per segment)
  trajectory     -> gpuRIR/__init__.py:177-221 (segmentation + float64 overlap-add)
  n_samples, tw  -> src/gpuRIR_cuda.cu:831-835, :629
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")

EXACT_GATHER, EXACT_SCATTER, LUT, FP32_REF = 0, 1, 2, 3
PATTERNS = {"omni": 0, "homni": 1, "card": 2, "hypcard": 3, "subcard": 4, "bidir": 5}

_lib = None


def build():
    """Compile rir_oracle.c with the Makefile beside it."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_simulate_rir.argtypes = [fp, fp, fp, ctypes.c_int, fp, ctypes.c_int, fp, fp,
                                       ctypes.c_int, ctypes.c_int, ip, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                       ctypes.c_int, dp, dp]
        L.orc_simulate_rir.restype = ctypes.c_int
        L.orc_simulate_rir_d.argtypes = [dp, dp, dp, ctypes.c_int, dp, ctypes.c_int, dp, dp,
                                         ctypes.c_int, ctypes.c_int, ip, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                         ctypes.c_int, dp, dp, fp, fp, fp, fp, fp, fp]
        L.orc_simulate_rir_d.restype = ctypes.c_int
        L.orc_images.argtypes = [fp, fp, fp, ctypes.c_int, fp, ctypes.c_int, fp, fp, ctypes.c_int,
                                 ctypes.c_int, ip, ctypes.c_float, ctypes.c_float,
                                 ctypes.POINTER(ctypes.c_int64), dp, dp, dp]
        L.orc_n_samples.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_n_samples.restype = ctypes.c_int
        L.orc_tw.argtypes = [ctypes.c_float]
        L.orc_tw.restype = ctypes.c_int
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.POINTER(ctypes.c_uint32)]
        L.orc_noise.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_int]
        L.orc_noise.restype = ctypes.c_double
        L.orc_lut_table.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.orc_lut_table.restype = ctypes.c_int
        L.orc_tap_exact.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_tap_exact.restype = ctypes.c_double
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def n_samples(T, fs):
    return lib().orc_n_samples(float(T), float(fs))


def tw(fs):
    return lib().orc_tw(float(fs))


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()


def simulate_rir(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=None, spkr_pattern="omni",
                 mic_pattern="omni", orV_src=None, orV_rcv=None, c=343.0, mode=EXACT_SCATTER,
                 noise="logistic", seed=1234, room_index=0, call_offset=0, level=1.0, crossfade=0,
                 return_env=False):
    """FP64 restatement of simulateRIR; returns float64 [M_src, M_rcv, nSamples].
    crossfade: samples of linear cross-fade before Tdiff (0 = the reference's hard switch)."""
    room = _f32(room_sz, (3,))
    b = _f32(beta, (6,))
    ps = _f32(pos_src).reshape(-1, 3)
    pr = _f32(pos_rcv).reshape(-1, 3)
    os_ = np.zeros_like(ps) if orV_src is None else _f32(orV_src).reshape(-1, 3)
    or_ = np.zeros_like(pr) if orV_rcv is None else _f32(orV_rcv).reshape(-1, 3)
    nb = np.ascontiguousarray(np.asarray(nb_img, dtype=np.int32).reshape(3))
    if Tdiff is None:
        Tdiff = Tmax
    ns = n_samples(Tmax, fs)
    out = np.zeros((ps.shape[0], pr.shape[0], ns), dtype=np.float64)
    env = np.zeros((ps.shape[0] * pr.shape[0], 3), dtype=np.float64)
    sp = PATTERNS[spkr_pattern] if isinstance(spkr_pattern, str) else int(spkr_pattern)
    mp = PATTERNS[mic_pattern] if isinstance(mic_pattern, str) else int(mic_pattern)
    kind = {"logistic": 0, "gaussian": 1}[noise]
    lib().orc_simulate_rir(_ptr(room, ctypes.c_float), _ptr(b, ctypes.c_float), _ptr(ps, ctypes.c_float),
                           ps.shape[0], _ptr(pr, ctypes.c_float), pr.shape[0], _ptr(os_, ctypes.c_float),
                           _ptr(or_, ctypes.c_float), sp, mp, _ptr(nb, ctypes.c_int), float(Tdiff), float(Tmax),
                           float(fs), float(c), int(mode), kind, int(seed), int(room_index), int(call_offset),
                           float(level), int(crossfade), _ptr(out, ctypes.c_double), _ptr(env, ctypes.c_double))
    if return_env:
        return out, env
    return out


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def simulate_rir_f64(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=None, spkr_pattern="omni",
                     mic_pattern="omni", orV_src=None, orV_rcv=None, c=343.0, mode=EXACT_SCATTER, noise="logistic",
                     seed=1234, room_index=0, call_offset=0, level=1.0, crossfade=0, return_env=False):
    """simulate_rir with double-precision geometry (room, betas, positions, orientations)."""
    room = _f64(room_sz).reshape(3)
    b = _f64(beta).reshape(6)
    ps = _f64(pos_src).reshape(-1, 3)
    pr = _f64(pos_rcv).reshape(-1, 3)
    os_ = np.zeros_like(ps) if orV_src is None else _f64(orV_src).reshape(-1, 3)
    or_ = np.zeros_like(pr) if orV_rcv is None else _f64(orV_rcv).reshape(-1, 3)
    nb = np.ascontiguousarray(np.asarray(nb_img, dtype=np.int32).reshape(3))
    if Tdiff is None:
        Tdiff = Tmax
    assert mode != FP32_REF
    ns = n_samples(Tmax, fs)
    out = np.zeros((ps.shape[0], pr.shape[0], ns), dtype=np.float64)
    env = np.zeros((ps.shape[0] * pr.shape[0], 3), dtype=np.float64)
    sp = PATTERNS[spkr_pattern] if isinstance(spkr_pattern, str) else int(spkr_pattern)
    mp = PATTERNS[mic_pattern] if isinstance(mic_pattern, str) else int(mic_pattern)
    kind = {"logistic": 0, "gaussian": 1}[noise]
    nf = ctypes.POINTER(ctypes.c_float)()
    lib().orc_simulate_rir_d(_ptr(room, ctypes.c_double), _ptr(b, ctypes.c_double), _ptr(ps, ctypes.c_double),
                             ps.shape[0], _ptr(pr, ctypes.c_double), pr.shape[0], _ptr(os_, ctypes.c_double),
                             _ptr(or_, ctypes.c_double), sp, mp, _ptr(nb, ctypes.c_int), float(Tdiff), float(Tmax),
                             float(fs), float(c), int(mode), kind, int(seed), int(room_index), int(call_offset),
                             float(level), int(crossfade), _ptr(out, ctypes.c_double), _ptr(env, ctypes.c_double),
                             nf, nf, nf, nf, nf, nf)
    if return_env:
        return out, env
    return out


def images(room_sz, beta, pos_src, pos_rcv, nb_img, fs, spkr_pattern="omni", mic_pattern="omni",
           orV_src=None, orV_rcv=None, c=343.0):
    """Stage (a): per (rir, image) integer offset n0, fraction and signed amplitude."""
    room = _f32(room_sz, (3,))
    b = _f32(beta, (6,))
    ps = _f32(pos_src).reshape(-1, 3)
    pr = _f32(pos_rcv).reshape(-1, 3)
    os_ = np.zeros_like(ps) if orV_src is None else _f32(orV_src).reshape(-1, 3)
    or_ = np.zeros_like(pr) if orV_rcv is None else _f32(orV_rcv).reshape(-1, 3)
    nb = np.ascontiguousarray(np.asarray(nb_img, dtype=np.int32).reshape(3))
    M = ps.shape[0] * pr.shape[0]
    N = int(np.prod(nb.astype(np.int64)))
    n0 = np.zeros(M * N, dtype=np.int64)
    fr = np.zeros(M * N, dtype=np.float64)
    amp = np.zeros(M * N, dtype=np.float64)
    tdp = np.zeros(M, dtype=np.float64)
    sp = PATTERNS[spkr_pattern] if isinstance(spkr_pattern, str) else int(spkr_pattern)
    mp = PATTERNS[mic_pattern] if isinstance(mic_pattern, str) else int(mic_pattern)
    lib().orc_images(_ptr(room, ctypes.c_float), _ptr(b, ctypes.c_float), _ptr(ps, ctypes.c_float), ps.shape[0],
                     _ptr(pr, ctypes.c_float), pr.shape[0], _ptr(os_, ctypes.c_float), _ptr(or_, ctypes.c_float),
                     sp, mp, _ptr(nb, ctypes.c_int), float(fs), float(c), _ptr(n0, ctypes.c_int64),
                     _ptr(fr, ctypes.c_double), _ptr(amp, ctypes.c_double), _ptr(tdp, ctypes.c_double))
    return n0.reshape(M, N), fr.reshape(M, N), amp.reshape(M, N), tdp


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def noise(seed, rir, room, offset, s, kind=0):
    return lib().orc_noise(int(seed), int(rir), int(room), int(offset), int(s), int(kind))


def lut_table(Tw):
    cap = Tw * 16 + 2
    tab = np.zeros(cap, dtype=np.float64)
    n = lib().orc_lut_table(int(Tw), _ptr(tab, ctypes.c_double), cap)
    return tab[:n]


def tap_exact(x, Tw):
    return lib().orc_tap_exact(float(x), int(Tw))


def gpu_conv(segments, rirs):
    """src/gpuRIR_cuda.cu:933-1023: out[p, r, :seg+L-1] = linear conv of segment p with RIR[p, r] (FP64)."""
    seg = np.asarray(segments, dtype=np.float64)
    rir = np.asarray(rirs, dtype=np.float64)
    P, S = seg.shape
    _, R, L = rir.shape
    n = S + L - 1
    nfft = 1 << (n - 1).bit_length()
    fs = np.fft.rfft(seg, nfft, axis=1)
    fr = np.fft.rfft(rir, nfft, axis=2)
    return np.fft.irfft(fr * fs[:, None, :], nfft, axis=2)[:, :, :n]


def trajectory_windows(n_samples_sig, n_pts, timestamps=None, fs=None):
    """gpuRIR/__init__.py:202-211: segment starts w_ini (with n_samples appended) and lengths."""
    if timestamps is None:
        fs = n_samples_sig / n_pts
        timestamps = np.arange(n_pts)
    w_ini = np.append((np.asarray(timestamps) * fs).astype(int), n_samples_sig)
    return w_ini, np.diff(w_ini)


def simulate_trajectory(source_signal, rirs, timestamps=None, fs=None):
    """gpuRIR/__init__.py:177-221 in FP64: segmentation, per-segment convolution, float64 OLA."""
    sig = np.asarray(source_signal, dtype=np.float64)
    rirs = np.asarray(rirs)
    n_sig = len(sig)
    P, R, L = rirs.shape
    w_ini, w_len = trajectory_windows(n_sig, P, timestamps, fs)
    segs = np.zeros((P, int(w_len.max())))
    for n in range(P):
        segs[n, :w_len[n]] = sig[w_ini[n]:w_ini[n + 1]]
    conv = gpu_conv(segs, rirs.astype(np.float64))
    out = np.zeros((n_sig + L - 1, R))
    for m in range(R):
        for n in range(P):
            out[w_ini[n]:w_ini[n + 1] + L - 1, m] += conv[n, m, :w_len[n] + L - 1]
    return out
