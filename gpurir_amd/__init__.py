"""Room impulse responses (image source method) on AMD MI355X -- drop-in for gpuRIR.

``import gpurir_amd as gpuRIR`` gives the API of gpuRIR/__init__.py (DavidDiazGuerra/gpuRIR
v1.2.0): the same functions, arguments, defaults, polar patterns, assertion messages and
output layouts.  The simulation runs in hand-written gfx950 HIP kernels reached through
the C ABI of libgpurir.so (include/gpurir.h); there is no CPU fallback.

Differences that follow from the design (see the repository documentation):
  * the diffuse tail draws its noise from Philox4x32-10 with the library's seed
    (1234, as the reference seeds cuRAND) and a per-call counter, not from cuRAND;
  * activateMixedPrecision selects FP16 arithmetic with FP32 accumulation (packed-FP16
    taps on small launches, FP16 filter inputs on the matrix cores on large ones)
    instead of the reference's half2 accumulation; the reference's messages and the mutual exclusion with the
    lookup table are kept.
"""
from __future__ import division

import numpy as np

from . import _lib

__all__ = ["polar_patterns", "beta_SabineEstimation", "att2t_SabineEstimator", "t2n", "simulateRIR",
           "simulateTrajectory", "activateMixedPrecision", "activateLUT", "activate_mixed_precision", "activate_lut"]

# gpuRIR/__init__.py:13-20
polar_patterns = {
    "omni": 0,
    "homni": 1,
    "card": 2,
    "hypcard": 3,
    "subcard": 4,
    "bidir": 5,
}


def _wall_areas(room_sz):
    """Areas of the six walls in the order of beta: x0, x1 (L_y L_z), y0, y1 (L_x L_z), z0, z1 (L_x L_y)."""
    lx, ly, lz = room_sz[0], room_sz[1], room_sz[2]
    return (ly * lz, ly * lz, lx * lz, lx * lz, lx * ly, lx * ly)


def beta_SabineEstimation(room_sz, T60, abs_weights=[1.0] * 6):
    """Reflection coefficients giving the requested T60 by Sabine's formula (gpuRIR/__init__.py:22-54).

    The absorption of wall i is x * w_i with w the weights scaled to a maximum of 1; x in [0, 1]
    minimises |T60 - 0.161 V / sum_i x w_i S_i| (0.161 V / 0 taken as 0 for an anechoic room),
    found by scipy.optimize.minimize from x = 0.5 with bounds [0, 1] as in the reference, so the
    float32 result is the reference's.  Returns [beta_x0, beta_x1, beta_y0, beta_y1, beta_z0, beta_z1].
    """
    from scipy.optimize import minimize

    areas = _wall_areas(room_sz)
    volume = np.prod(room_sz)
    w = abs_weights / np.array(abs_weights).max()

    def misfit(x):
        a = x * w
        absorption = (a[0] + a[1]) * areas[0] + (a[2] + a[3]) * areas[2] + (a[4] + a[5]) * areas[4]
        if absorption == 0:
            return T60
        return abs(T60 - 0.161 * volume / absorption)

    x = minimize(misfit, 0.5, bounds=[[0, 1]]).x
    return np.sqrt(1 - x * w).astype(np.float32)


def beta_SabineEstimation_batch(room_sz, T60, abs_weights=None):
    """Vectorised Sabine prelude for many rooms (no reference counterpart; the reference solves
    one room per scipy call).  room_sz [N, 3], T60 [N], abs_weights [6] or [N, 6].

    The misfit above is zero at x = 0.161 V / (T60 sum_i w_i S_i); clipped to [0, 1] that is its
    minimiser, which scipy's L-BFGS-B reaches to ~1e-8 (the betas agree to ~1e-5, float32).
    Returns float32 [N, 6]."""
    L = np.atleast_2d(np.asarray(room_sz, dtype=np.float64))
    T60 = np.atleast_1d(np.asarray(T60, dtype=np.float64))
    w = np.ones((L.shape[0], 6)) if abs_weights is None else np.broadcast_to(
        np.asarray(abs_weights, dtype=np.float64), (L.shape[0], 6))
    w = w / w.max(axis=1, keepdims=True)
    S = np.stack(_wall_areas(L.T), axis=1)
    x = np.clip(0.161 * L.prod(axis=1) / (T60 * (w * S).sum(axis=1)), 0.0, 1.0)
    return np.sqrt(1.0 - x[:, None] * w).astype(np.float32)


def att2t_SabineEstimator(att_dB, T60):
    """Time for the RIR to reach att_dB of attenuation (gpuRIR/__init__.py:56-72)."""
    return att_dB / 60.0 * T60


def t2n(T, rooms_sz, c=343.0):
    """Images per dimension needed for RIRs of length T (gpuRIR/__init__.py:74-93)."""
    nb_img = 2 * T / (np.array(rooms_sz) / c)
    return [int(n) for n in np.ceil(nb_img)]


def _f32c(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def simulateRIR(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=None, spkr_pattern="omni",
                mic_pattern="omni", orV_src=None, orV_rcv=None, c=343.0):
    """RIRs of a shoebox room by the image source method (gpuRIR/__init__.py:95-175).

    Returns float32 [n_src, n_rcv, nSamples], nSamples = ceil(Tmax*fs) rounded up to even.
    The early part (t < Tdiff) is the windowed-sinc sum over images; the rest is the
    diffuse model.  Uses the lookup-table state set by activateLUT (on by default).
    """
    assert not ((pos_src >= room_sz).any() or (pos_src <= 0).any()), "The sources must be inside the room"
    assert not ((pos_rcv >= room_sz).any() or (pos_rcv <= 0).any()), "The receivers must be inside the room"
    assert Tdiff is None or Tdiff <= Tmax, "Tmax must be equal or greater than Tdiff"
    assert mic_pattern in polar_patterns, "mic_pattern must be omni, homni, card, hypcard, subcard or bidir"
    assert mic_pattern == "omni" or orV_rcv is not None, "the mics are not omni but their orientation is undefined"
    assert spkr_pattern in spkr_pattern, "spkr_pattern must be omni, homni, card, hypcard, subcard or bidir"
    assert spkr_pattern == "omni" or orV_src is not None, "the sources are not omni but their orientation is undefined"

    pos_src = _f32c(pos_src)
    pos_rcv = _f32c(pos_rcv)
    if Tdiff is None:
        Tdiff = Tmax
    if mic_pattern is None:
        mic_pattern = "omni"
    if spkr_pattern is None:
        spkr_pattern = "omni"
    orV_rcv = None if orV_rcv is None else _f32c(orV_rcv)  # absent orientations go as NULL (zero vectors)
    orV_src = None if orV_src is None else _f32c(orV_src)
    spkr = polar_patterns[spkr_pattern]  # KeyError for an unknown pattern, as the reference
    mic = polar_patterns[mic_pattern]

    room = _f32c(room_sz).reshape(3)
    b = _f32c(beta).reshape(6)
    nb = np.ascontiguousarray(np.asarray(nb_img, dtype=np.int32).reshape(3))
    assert pos_src.ndim == 2 and pos_src.shape[1] == 3 and pos_rcv.ndim == 2 and pos_rcv.shape[1] == 3
    assert orV_src is None or orV_src.shape == pos_src.shape
    assert orV_rcv is None or orV_rcv.shape == pos_rcv.shape
    lib = _lib.load()
    ns = lib.gpurir_n_samples(float(Tmax), float(fs))
    out = np.empty((pos_src.shape[0], pos_rcv.shape[0], ns), dtype=np.float32)
    addr = _lib.addr
    _lib.check(lib.gpurir_simulate_rir(addr(room), addr(b), addr(pos_src), pos_src.shape[0], addr(pos_rcv),
                                       pos_rcv.shape[0], None if orV_src is None else addr(orV_src),
                                       None if orV_rcv is None else addr(orV_rcv), spkr, mic, addr(nb),
                                       float(Tdiff), float(Tmax), float(fs), float(c), addr(out)))
    return out


def simulateTrajectory(source_signal, RIRs, timestamps=None, fs=None):
    """Signals of a moving source recorded by a microphone array (gpuRIR/__init__.py:177-221).

    RIRs is [n_pts, n_rcv, len] (numpy, or a torch tensor already on the GPU, which then
    stays there).  Segmentation, batched FFT convolution and the float64 overlap-add all
    run on the GPU.  Returns float64 [len(source_signal) + len - 1, n_rcv].
    """
    import torch

    nSamples = len(source_signal)
    nPts = RIRs.shape[0]
    nRcv = RIRs.shape[1]
    lenRIR = RIRs.shape[2]

    assert timestamps is None or fs is not None, "fs must be indicated for custom timestamps"
    assert timestamps is None or timestamps[0] == 0, "The first timestamp must be 0"
    if timestamps is None:
        fs = nSamples / nPts
        timestamps = np.arange(nPts)

    w_ini = np.append((np.asarray(timestamps) * fs).astype(int), nSamples)
    w_ini = np.ascontiguousarray(w_ini, dtype=np.int32)

    dev = torch.device("cuda", torch.cuda.current_device())
    if isinstance(RIRs, torch.Tensor):
        d_rirs = RIRs.detach().to(device=dev, dtype=torch.float32).contiguous()
    else:
        d_rirs = torch.from_numpy(np.ascontiguousarray(RIRs, dtype=np.float32)).to(dev)
    if isinstance(source_signal, torch.Tensor):
        d_sig = source_signal.detach().to(device=dev, dtype=torch.float32).contiguous()
    else:
        d_sig = torch.from_numpy(np.ascontiguousarray(np.asarray(source_signal), dtype=np.float32)).to(dev)
    d_out = torch.empty((nSamples + lenRIR - 1, nRcv), dtype=torch.float64, device=dev)
    lib = _lib.load()
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(lib.gpurir_trajectory_device(d_sig.data_ptr(), nSamples, d_rirs.data_ptr(), nPts, nRcv, lenRIR,
                                            _lib.iptr(w_ini), d_out.data_ptr(), stream))
    return d_out.cpu().numpy()


def activateMixedPrecision(activate=True):
    """Mixed-precision toggle (gpuRIR/__init__.py:223-232); mutually exclusive with the table."""
    _lib.load().gpurir_activate_mixed_precision(int(bool(activate)))


def activateLUT(activate=True):
    """Lookup-table toggle for the sinc evaluation (gpuRIR/__init__.py:234-243)."""
    _lib.load().gpurir_activate_lut(int(bool(activate)))


def set_seed(seed=1234, next_call_offset=0):
    """Noise stream of the drop-in simulateRIR: Philox key and the counter word of the next call.

    No reference counterpart: the reference seeds its generator once with 1234
    (src/gpuRIR_cuda.cu:1046-1047) and lets the stream advance from call to call."""
    _lib.load().gpurir_set_seed(int(seed), int(next_call_offset))


# the reference lists these spellings in __all__ (gpuRIR/__init__.py:11)
activate_mixed_precision = activateMixedPrecision
activate_lut = activateLUT
