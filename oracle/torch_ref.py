"""Differentiable FP64 restatement of the simulation (torch autograd) -- TEST INFRASTRUCTURE ONLY.

Dense (every image x every sample) so only for small rooms.  It follows the same
reference lines as rir_oracle.c (src/gpuRIR_cuda.cu:110-147, :245-375 images and taps,
:415-461 envelope and tail) and uses the oracle's Philox noise, so its value equals
oracle.simulate_rir_f64 and its autograd gradient pins the hand-written adjoint.
Piecewise-constant integer choices (the envelope window start) are held fixed, as the
adjoint does.
"""
import math

import numpy as np
import torch

from . import oracle as orc

PATTERN_COEF = {"omni": None, "card": (0.5, 0.5), "hypcard": (0.25, 0.75), "subcard": (0.75, 0.25),
                "bidir": (0.0, 1.0)}


def _gain(c, pattern):
    if pattern == "omni":
        return torch.ones_like(c)
    if pattern == "homni":
        return (c > 0).to(c.dtype)
    a, b = PATTERN_COEF[pattern]
    return a + b * c


def simulate(room_sz, beta, src, rcv, nb_img, Tmax, fs, Tdiff=None, level=1.0, spkr_pattern="omni",
             mic_pattern="omni", orV_src=None, orV_rcv=None, c=343.0, seed=1234, room_index=0, call_offset=0):
    """beta [6], src [S,3], rcv [R,3], level: float64 tensors (may require grad).
    Returns [S, R, nSamples] float64."""
    dt = torch.float64
    L = torch.as_tensor(np.asarray(room_sz, np.float64), dtype=dt)
    if Tdiff is None:
        Tdiff = Tmax
    nISM = orc.n_samples(Tdiff, fs)
    nS = orc.n_samples(Tmax, fs)
    Tw = orc.tw(fs)
    k_fc = float(np.float32(fs)) / float(np.float32(c))
    axes = []
    for d in range(3):
        N = int(nb_img[d])
        n = np.arange(N)
        k = n - N // 2
        m = np.ceil(k / 2.0).astype(np.int64)
        q = np.abs(np.fmod(k, 2)).astype(np.int64)
        axes.append((torch.as_tensor(m), torch.as_tensor(q)))
    mx, my, mz = torch.meshgrid(axes[0][0], axes[1][0], axes[2][0], indexing="ij")
    qx, qy, qz = torch.meshgrid(axes[0][1], axes[1][1], axes[2][1], indexing="ij")
    M = torch.stack([mx.reshape(-1), my.reshape(-1), mz.reshape(-1)], 1)  # [N,3]
    Q = torch.stack([qx.reshape(-1), qy.reshape(-1), qz.reshape(-1)], 1)
    sig = (1 - 2 * Q).to(dt)
    att = torch.ones(M.shape[0], dtype=dt)
    for d in range(3):
        e0 = (M[:, d] - Q[:, d]).abs().to(dt)
        e1 = M[:, d].abs().to(dt)
        att = att * torch.pow(-beta[2 * d], e0) * torch.pow(-beta[2 * d + 1], e1)
    t = torch.arange(nISM, dtype=dt)
    out = []
    # Sabine envelope constants (src/gpuRIR_cuda.cu:123-130, :438)
    b = beta
    Sa = ((1 - b[0] ** 2) + (1 - b[1] ** 2)) * L[1] * L[2] + ((1 - b[2] ** 2) + (1 - b[3] ** 2)) * L[0] * L[2] + \
        ((1 - b[4] ** 2) + (1 - b[5] ** 2)) * L[0] * L[1]
    T60 = 0.161 * (L[0] * L[1] * L[2]) / Sa
    alpha = -13.8155 / (T60 * float(np.float32(fs)))
    w_sz = float(np.float32(np.float32(10e-3) * np.float32(fs)))
    for si in range(src.shape[0]):
        row = []
        for ri in range(rcv.shape[0]):
            v = 2.0 * M.to(dt) * L[None, :] + sig * src[si][None, :] - rcv[ri][None, :]
            dist = torch.sqrt((v * v).sum(1))
            amp = att / (4.0 * math.pi * dist)
            if mic_pattern != "omni":
                o = torch.as_tensor(np.asarray(orV_rcv, np.float64)[ri])
                cth = (v @ o) / dist / torch.linalg.norm(o)
                amp = amp * _gain(cth, mic_pattern)
            if spkr_pattern != "omni":
                o = torch.as_tensor(np.asarray(orV_src, np.float64)[si])
                cth = ((-v) * (sig * o[None, :])).sum(1) / dist / torch.linalg.norm(o)
                amp = amp * _gain(cth, spkr_pattern)
            tau = dist * k_fc
            x = t[None, :] - tau[:, None]
            inside = x.abs() < Tw / 2.0
            h = torch.where(inside, torch.sinc(x) * 0.5 * (1 + torch.cos(2 * math.pi * x / Tw)), torch.zeros_like(x))
            ism = (amp[:, None] * h).sum(0)
            if nS > nISM:
                tau_dp = torch.linalg.norm(src[si] - rcv[ri]) * k_fc
                w_start = int(math.ceil(max(nISM - w_sz, float(tau_dp.detach()))))
                P = (ism[w_start:nISM] ** 2).sum() / (nISM - w_start)
                w_c = w_start + (nISM - w_start) / 2.0
                A = P / torch.exp(alpha * (w_c - tau_dp))
                s = torch.arange(nS - nISM, dtype=dt)
                m_idx = si * rcv.shape[0] + ri
                z = torch.as_tensor(np.array([orc.noise(seed, m_idx, room_index, call_offset, int(k), 0)
                                              for k in range(nS - nISM)]))
                tail = level * torch.sqrt(A * torch.exp(alpha * (nISM + s - tau_dp))) * z
                row.append(torch.cat([ism, tail]))
            else:
                row.append(ism)
        out.append(torch.stack(row))
    return torch.stack(out)
