"""Canonical synthetic workloads (seeded) used by the benchmark and the parity tests.

Draw order with rng = numpy.random.default_rng(seed): positions first (sources,
then receivers), then orientations, then reflection-coefficient jitter.
U_m(L, n) = 0.5 + U(0,1)^(n,3) * (L - 1): uniform with a 0.5 m wall margin.

  plain_room       1 room 3x3x2.5 m, T60 0.7 s, 1 src x 2 rcv, omni, Tmax 0.8 s, Tdiff at 15 dB
  mid_room         1 room 6x5x3 m, T60 1.0 s, betas +-5%, 4 src x 32 rcv, hypcard src / card mic,
                   Tmax 1.1 s, Tdiff at 12 dB
  uneven_rooms     4096 rooms of uneven sizes, T60 in [0.2, 1.2] s, 1 src x 4..16 rcv, betas +-10%,
                   Tmax 1.1*T60, Tdiff at 15 dB
  gradient_rooms   256 rooms, 1 src x 4 omni rcv, T60 in [0.3, 0.8] s, Tmax 0.5 s, Tdiff at 20 dB
  trajectory       room 8x6x3.5 m, T60 0.6 s, 128 source points on a 4 m line, 8-mic circle
                   (5 cm), full ISM with Tmax 0.5 s, 10 s of white noise at 16 kHz
"""
import numpy as np

from . import att2t_SabineEstimator, beta_SabineEstimation, t2n

FS = 16000.0


def _um(rng, L, n):
    return (0.5 + rng.uniform(size=(n, 3)) * (np.asarray(L, float) - 1.0)).astype(np.float32)


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def beta_sabine_closed(L, T60):
    """Closed-form Sabine betas (uniform absorption), vectorised over rooms: alpha = 0.161 V / (T60 S)."""
    L = np.atleast_2d(np.asarray(L, float))
    T60 = np.atleast_1d(np.asarray(T60, float))
    V = L.prod(axis=1)
    S = 2 * (L[:, 0] * L[:, 1] + L[:, 0] * L[:, 2] + L[:, 1] * L[:, 2])
    alpha = np.clip(0.161 * V / (T60 * S), 0.0, 1.0)
    return np.sqrt(1.0 - alpha)


def plain_room(seed=0):
    rng = np.random.default_rng(seed)
    L = [3.0, 3.0, 2.5]
    T60 = 0.7
    pos_src = _um(rng, L, 1)
    pos_rcv = _um(rng, L, 2)
    beta = beta_SabineEstimation(L, T60)
    Tdiff = att2t_SabineEstimator(15.0, T60)
    return dict(room_sz=L, beta=beta, pos_src=pos_src, pos_rcv=pos_rcv, nb_img=t2n(Tdiff, L), Tmax=0.8, fs=FS,
                Tdiff=Tdiff, spkr_pattern="omni", mic_pattern="omni", orV_src=None, orV_rcv=None)


def mid_room(seed=0, n_src=4, n_rcv=32):
    rng = np.random.default_rng(seed)
    L = [6.0, 5.0, 3.0]
    T60 = 1.0
    pos_src = _um(rng, L, n_src)
    pos_rcv = _um(rng, L, n_rcv)
    orV_src = _unit(rng, n_src)
    orV_rcv = _unit(rng, n_rcv)
    b0 = float(beta_SabineEstimation(L, T60)[0])
    beta = np.clip(b0 * (1.0 + rng.uniform(-0.05, 0.05, size=6)), 0.0, 1.0).astype(np.float32)
    Tdiff = att2t_SabineEstimator(12.0, T60)
    return dict(room_sz=L, beta=beta, pos_src=pos_src, pos_rcv=pos_rcv, nb_img=t2n(Tdiff, L), Tmax=1.1, fs=FS,
                Tdiff=Tdiff, spkr_pattern="hypcard", mic_pattern="card", orV_src=orV_src, orV_rcv=orV_rcv)


def uneven_rooms(n_rooms=4096, seed=0):
    rng = np.random.default_rng(seed)
    L = np.stack([rng.uniform(3, 10, n_rooms), rng.uniform(3, 8, n_rooms), rng.uniform(2.5, 4, n_rooms)], axis=1)
    T60 = rng.uniform(0.2, 1.2, n_rooms)
    n_rcv = rng.integers(4, 17, n_rooms)
    rooms = []
    for k in range(n_rooms):
        ps = _um(rng, L[k], 1)
        pr = _um(rng, L[k], int(n_rcv[k]))
        rooms.append(dict(room_sz=L[k].astype(np.float32), pos_src=ps, pos_rcv=pr))
    b = beta_sabine_closed(L, T60)
    jit = rng.uniform(-0.1, 0.1, size=(n_rooms, 6))
    for k, r in enumerate(rooms):
        r["beta"] = np.clip(b[k] * (1.0 + jit[k]), 0.0, 1.0).astype(np.float32)
        r["Tmax"] = float(1.1 * T60[k])
        r["Tdiff"] = float(att2t_SabineEstimator(15.0, T60[k]))
        r["nb_img"] = t2n(r["Tdiff"], L[k])
    return rooms


def gradient_rooms(n_rooms=256, seed=0):
    rng = np.random.default_rng(seed)
    L = np.stack([rng.uniform(4, 8, n_rooms), rng.uniform(3, 6, n_rooms), rng.uniform(2.5, 3.5, n_rooms)], axis=1)
    T60 = rng.uniform(0.3, 0.8, n_rooms)
    rooms = []
    for k in range(n_rooms):
        rooms.append(dict(room_sz=L[k].astype(np.float32), pos_src=_um(rng, L[k], 1), pos_rcv=_um(rng, L[k], 4)))
    b = beta_sabine_closed(L, T60)
    for k, r in enumerate(rooms):
        r["beta"] = np.full(6, b[k], dtype=np.float32)
        r["Tmax"] = 0.5
        r["Tdiff"] = float(att2t_SabineEstimator(20.0, T60[k]))
        r["nb_img"] = t2n(r["Tdiff"], L[k])
    return rooms


def gradient_targets(rooms, seed=1):
    """Targets of the inverse problem: each room's betas scaled by 1 + U(-0.2, 0.2) (clipped
    to [0, 1]) and its source moved by N(0, 0.1^2) m per axis (kept 0.1 m inside the walls)."""
    rng = np.random.default_rng(seed)
    out = []
    for r in rooms:
        t = dict(r)
        t["beta"] = np.clip(r["beta"] * (1.0 + rng.uniform(-0.2, 0.2, 6)), 0.0, 1.0).astype(np.float32)
        L = np.asarray(r["room_sz"], np.float64)
        ps = np.asarray(r["pos_src"], np.float64) + rng.normal(0.0, 0.1, np.shape(r["pos_src"]))
        t["pos_src"] = np.clip(ps, 0.1, L - 0.1).astype(np.float32)
        out.append(t)
    return out


def trajectory(seed=0, n_pts=128, n_signal=160000):
    rng = np.random.default_rng(seed)
    L = [8.0, 6.0, 3.5]
    T60 = 0.6
    pos_src = np.stack([np.linspace(2.0, 6.0, n_pts), np.full(n_pts, 3.0), np.full(n_pts, 1.5)], 1).astype(np.float32)
    ang = 2 * np.pi * np.arange(8) / 8
    pos_rcv = np.stack([4.0 + 0.05 * np.cos(ang), 3.0 + 0.05 * np.sin(ang), np.full(8, 1.75)], 1).astype(np.float32)
    signal = rng.standard_normal(n_signal).astype(np.float32)
    Tmax = 0.5
    return dict(room_sz=L, beta=beta_SabineEstimation(L, T60), pos_src=pos_src, pos_rcv=pos_rcv,
                nb_img=t2n(Tmax, L), Tmax=Tmax, fs=FS, Tdiff=Tmax, signal=signal)
