"""Pins the FP64 oracle (oracle/rir_oracle.c) against published vectors, library
routines, closed forms and brute force -- never against a retyped formula.

The reference (gpuRIR) ships no tests, fixtures or golden outputs; these checks
are what fixes the oracle (each names the mistake it would catch)."""
import itertools

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as orc

C = 343.0


def hann_np(x, Tw):
    """Continuous Hann window 0.5(1+cos(2 pi x/Tw)) for x on a half-integer grid, via numpy.hanning."""
    n = np.round(2 * np.asarray(x) + Tw).astype(int)  # grid of 2*Tw+1 points over [-Tw/2, Tw/2]
    return np.hanning(2 * Tw + 1)[n]


# ---------------------------------------------------------------- Philox KATs
def test_philox_published_vectors():
    # Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11)
    assert orc.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert orc.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert orc.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_noise_distributions():
    s = np.arange(20000)
    lg = np.array([orc.noise(1234, 3, 7, 11, int(k), 0) for k in s])
    # unit-variance logistic: scale sqrt(3)/pi (reference constant 0.551329)
    assert stats.kstest(lg, stats.logistic(scale=np.sqrt(3) / np.pi).cdf).pvalue > 1e-3
    assert abs(lg.mean()) < 0.03 and abs(lg.var() - 1.0) < 0.05
    ga = np.array([orc.noise(1234, 3, 7, 11, int(k), 1) for k in s])
    assert stats.kstest(ga, "norm").pvalue > 1e-3
    # different counters give different streams
    other = np.array([orc.noise(1234, 4, 7, 11, int(k), 0) for k in s[:100]])
    assert not np.allclose(other, lg[:100])


# ------------------------------------------------------- rounding rules
def test_rounding_rules():
    assert orc.n_samples(0.8, 16000) == 12800
    assert orc.n_samples(0.175, 16000) == 2800
    assert orc.n_samples(1.1, 16000) == 17600
    assert orc.n_samples(0.2, 16000) == 3200
    assert orc.n_samples(0.5, 16000) == 8000
    # ceil then forced even
    assert orc.n_samples(0.0001, 16000) == 2
    assert orc.n_samples(1.0 / 16000, 16000) % 2 == 0
    assert orc.tw(16000) == 128 and orc.tw(44100) == 353 and orc.tw(48000) == 384 and orc.tw(8000) == 64


# ---------------------------------------------------------- windowed sinc tap
@pytest.mark.parametrize("Tw", [128, 353])
def test_tap_against_numpy_sinc_and_hanning(Tw):
    xs = np.arange(-Tw, Tw + 1) / 2.0  # half-integer grid
    got = np.array([orc.tap_exact(x, Tw) for x in xs])
    want = np.where(np.abs(xs) < Tw / 2, np.sinc(xs) * hann_np(xs, Tw), 0.0)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
    # sinc zeros at every non-zero integer, 1 at 0
    ints = np.arange(-(Tw // 2) + 1, Tw // 2)
    vals = np.array([orc.tap_exact(float(k), Tw) for k in ints])
    assert vals[ints == 0][0] == 1.0
    assert np.all(np.abs(vals[ints != 0]) < 1e-15)


def test_lut_table():
    Tw = 128
    tab = orc.lut_table(Tw)
    assert len(tab) == 2049  # Tw*16, forced odd
    i = np.arange(-1024, 1025)
    x = i / 16.0
    np.testing.assert_allclose(tab, np.sinc(x) * 0.5 * (1 + np.cos(2 * np.pi * x / Tw)), atol=1e-15)
    assert len(orc.lut_table(353)) == 353 * 16 + 1


# ------------------------------------------------------------ geometry
def direct_rir(src, rcv, fs, nS, Tw):
    d = np.linalg.norm(np.asarray(src, float) - np.asarray(rcv, float))
    tau = d * fs / C
    t = np.arange(nS)
    x = t - tau
    h = np.where(np.abs(x) < Tw / 2, np.sinc(x) * 0.5 * (1 + np.cos(2 * np.pi * x / Tw)), 0.0)
    return h / (4 * np.pi * d), d


def test_direct_path_only_with_anechoic_walls():
    # beta = 0: every image but k = (0,0,0) has attenuation 0 (0^|m-q| 0^|m|)
    L = [5.0, 4.0, 3.0]
    src, rcv = np.array([[1.1, 2.2, 1.3]], np.float32), np.array([[3.7, 1.4, 2.1]], np.float32)
    out = orc.simulate_rir(L, [0.0] * 6, src, rcv, [7, 7, 7], 0.05, 16000)
    want, _ = direct_rir(src[0], rcv[0], 16000, out.shape[-1], 128)
    np.testing.assert_allclose(out[0, 0], want, rtol=1e-12, atol=1e-15)


def test_single_wall_reflection_sign_and_mirror():
    # only the wall x = 0 reflects (beta_x0 = 1): RIR = direct - image mirrored in x = 0
    L = [5.0, 4.0, 3.0]
    s, r = np.array([1.1, 2.2, 1.3]), np.array([3.7, 1.4, 2.1])
    out = orc.simulate_rir(L, [1.0, 0, 0, 0, 0, 0], s[None].astype(np.float32), r[None].astype(np.float32),
                           [9, 9, 9], 0.05, 16000)
    s32, r32 = s.astype(np.float32).astype(float), r.astype(np.float32).astype(float)
    a, _ = direct_rir(s32, r32, 16000, out.shape[-1], 128)
    b, _ = direct_rir(s32 * [-1, 1, 1], r32, 16000, out.shape[-1], 128)
    np.testing.assert_allclose(out[0, 0], a - b, rtol=1e-11, atol=1e-14)


def reflect_images_1d(x, L, b0, b1, K):
    """All 1-D images with <= K reflections by explicit mirroring: {position: amplitude}."""
    seen = {}
    frontier = [(x, 0, 0, None)]
    for _ in range(K + 1):
        nxt = []
        for (p, n0, n1, last) in frontier:
            key = round(p, 9)
            if key not in seen:
                seen[key] = (p, (-b0) ** n0 * (-b1) ** n1)
            if last != 0:
                nxt.append((-p, n0 + 1, n1, 0))  # mirror in wall 0
            if last != 1:
                nxt.append((2 * L - p, n0, n1 + 1, 1))  # mirror in wall L
        frontier = nxt
    return list(seen.values())


def test_image_set_equals_explicit_mirroring():
    L = np.array([3.0, 2.5, 2.0], np.float32)
    beta = np.array([0.9, 0.7, 0.8, 0.6, 0.5, 0.95], np.float32)
    s = np.array([0.7, 1.9, 0.6], np.float32)
    r = np.array([2.2, 0.8, 1.5], np.float32)
    K = 3
    nb = [2 * K + 1] * 3
    fs = 16000.0
    n0, fr, amp, _ = orc.images(L, beta, s[None], r[None], nb, fs)
    dist = (n0[0] + fr[0]) * C / fs
    got = sorted(zip(np.round(dist, 8), amp[0]))
    axes = [reflect_images_1d(float(s[d]), float(L[d]), float(beta[2 * d]), float(beta[2 * d + 1]), K)
            for d in range(3)]
    want = []
    r64 = r.astype(np.float64)
    for (px, ax), (py, ay), (pz, az) in itertools.product(*axes):
        d = np.sqrt((px - r64[0]) ** 2 + (py - r64[1]) ** 2 + (pz - r64[2]) ** 2)
        want.append((np.round(d, 8), ax * ay * az / (4 * np.pi * d)))
    want = sorted(want)
    assert len(got) == len(want) == (2 * K + 1) ** 3
    np.testing.assert_allclose([g[0] for g in got], [w[0] for w in want], atol=2e-8)
    np.testing.assert_allclose([g[1] for g in got], [w[1] for w in want], rtol=1e-9, atol=1e-14)


def test_reciprocity_omni():
    L = [4.0, 3.5, 2.7]
    beta = [0.8, 0.7, 0.9, 0.6, 0.75, 0.85]
    s = np.array([[0.9, 2.1, 1.1]], np.float32)
    r = np.array([[3.1, 0.7, 1.9]], np.float32)
    a = orc.simulate_rir(L, beta, s, r, [9, 9, 9], 0.06, 16000)
    b = orc.simulate_rir(L, beta, r, s, [9, 9, 9], 0.06, 16000)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-14)


def test_gather_equals_scatter():
    rng = np.random.default_rng(1)
    L = [3.0, 2.6, 2.2]
    beta = [0.9, 0.8, 0.85, 0.7, 0.6, 0.95]
    s = (0.5 + rng.uniform(size=(2, 3)) * (np.array(L) - 1.0)).astype(np.float32)
    r = (0.5 + rng.uniform(size=(2, 3)) * (np.array(L) - 1.0)).astype(np.float32)
    g = orc.simulate_rir(L, beta, s, r, [6, 5, 7], 0.03, 16000, mode=orc.EXACT_GATHER)
    sc = orc.simulate_rir(L, beta, s, r, [6, 5, 7], 0.03, 16000, mode=orc.EXACT_SCATTER)
    np.testing.assert_allclose(sc, g, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("pattern,gains", [
    ("card", [1.0, 0.5, 0.0]), ("hypcard", [1.0, 0.25, -0.5]), ("subcard", [1.0, 0.75, 0.5]),
    ("bidir", [1.0, 0.0, -1.0]), ("homni", [1.0, 0.0, 0.0])])
@pytest.mark.parametrize("side", ["mic", "spkr"])
def test_polar_patterns_0_90_180(pattern, gains, side):
    L = [6.0, 5.0, 3.0]
    s = np.array([[1.0, 2.5, 1.5]], np.float32)
    r = np.array([[4.0, 2.5, 1.5]], np.float32)  # doa along x
    omni = orc.simulate_rir(L, [0.0] * 6, s, r, [3, 3, 3], 0.03, 16000)[0, 0]
    peak = np.argmax(np.abs(omni))
    if side == "mic":
        # mic doa = rcv -> source direction = -x ; orientation facing the source is -x
        orients = [[-1, 0, 0], [0, 1, 0], [1, 0, 0]]
    else:
        # speaker doa = source -> rcv = +x
        orients = [[1, 0, 0], [0, 0, 1], [-1, 0, 0]]
    for o, gain in zip(orients, gains):
        ov = np.array([o], np.float32)
        kw = dict(mic_pattern=pattern, orV_rcv=ov) if side == "mic" else dict(spkr_pattern=pattern, orV_src=ov)
        out = orc.simulate_rir(L, [0.0] * 6, s, r, [3, 3, 3], 0.03, 16000, **kw)[0, 0]
        assert abs(out[peak] - gain * omni[peak]) < 1e-12 * abs(omni[peak]) + 1e-15


def test_tail_envelope_invariants():
    L = [6.0, 5.0, 3.0]
    beta = [0.9] * 6
    s = np.array([[1.5, 2.0, 1.2]], np.float32)
    r = np.array([[4.0, 3.0, 1.7]], np.float32)
    fs = 16000.0
    out, env = orc.simulate_rir(L, beta, s, r, [11, 11, 11], 0.4, fs, Tdiff=0.1, return_env=True)
    A, alpha, tau_dp = env[0]
    nISM = orc.n_samples(0.1, fs)
    # Sabine T60 of the room: the envelope power falls by 60 dB in T60 seconds
    Lr = np.array(L, float)
    b = np.array(beta, np.float32).astype(float)
    Sa = (2 - b[0] ** 2 - b[1] ** 2) * Lr[1] * Lr[2] + (2 - b[2] ** 2 - b[3] ** 2) * Lr[0] * Lr[2] + \
        (2 - b[4] ** 2 - b[5] ** 2) * Lr[0] * Lr[1]
    T60 = 0.161 * np.prod(Lr) / Sa
    assert abs(np.exp(alpha * T60 * fs) - 1e-6) < 1e-10  # 13.8155 ~ ln(1e6)
    # the envelope equals the mean ISM power of the last 10 ms at the window centre
    w_start = int(np.ceil(max(nISM - 160.0, tau_dp)))
    P = np.mean(out[0, 0, w_start:nISM] ** 2)
    wc = w_start + (nISM - w_start) / 2.0
    assert abs(A * np.exp(alpha * (wc - tau_dp)) - P) < 1e-12 * P
    # tail / noise is the square-root envelope
    tail = out[0, 0, nISM:]
    nz = np.array([orc.noise(1234, 0, 0, 0, k, 0) for k in range(len(tail))])
    envr = tail / nz
    sd = np.arange(len(tail))
    np.testing.assert_allclose(envr, np.sqrt(A * np.exp(alpha * (nISM + sd - tau_dp))), rtol=1e-12)


def test_lut_mode_close_to_exact():
    L = [4.0, 3.0, 2.5]
    beta = [0.85] * 6
    s = np.array([[1.0, 1.2, 1.1]], np.float32)
    r = np.array([[2.9, 2.0, 1.6]], np.float32)
    ex = orc.simulate_rir(L, beta, s, r, [9, 9, 9], 0.05, 16000)
    lu = orc.simulate_rir(L, beta, s, r, [9, 9, 9], 0.05, 16000, mode=orc.LUT)
    err = np.linalg.norm(lu - ex) / np.linalg.norm(ex)
    assert 1e-6 < err < 5e-3


def test_fp32_reference_arithmetic_close_to_fp64():
    L = [3.0, 3.0, 2.5]
    beta = [0.9] * 6
    s = np.array([[1.0, 1.2, 1.1]], np.float32)
    r = np.array([[2.1, 2.0, 1.6]], np.float32)
    ex = orc.simulate_rir(L, beta, s, r, [9, 9, 9], 0.04, 16000)
    f32 = orc.simulate_rir(L, beta, s, r, [9, 9, 9], 0.04, 16000, mode=orc.FP32_REF)
    err = np.linalg.norm(f32 - ex) / np.linalg.norm(ex)
    assert err < 1e-3


# ---------------------------------------------------------- trajectory
def test_gpu_conv_oracle_vs_numpy_convolve():
    rng = np.random.default_rng(2)
    seg = rng.standard_normal((3, 17))
    rir = rng.standard_normal((3, 2, 29))
    got = orc.gpu_conv(seg, rir)
    for p in range(3):
        for r in range(2):
            np.testing.assert_allclose(got[p, r], np.convolve(seg[p], rir[p, r]), atol=1e-12)


def test_trajectory_lti_invariant():
    # identical RIRs at every point: the output is the full convolution
    rng = np.random.default_rng(3)
    sig = rng.standard_normal(400)
    h = rng.standard_normal((2, 37))
    rirs = np.repeat(h[None], 8, axis=0)
    out = orc.simulate_trajectory(sig, rirs)
    for m in range(2):
        np.testing.assert_allclose(out[:, m], np.convolve(sig, h[m]), atol=1e-11)


def test_trajectory_custom_timestamps_windows():
    w_ini, w_len = orc.trajectory_windows(1000, 4, timestamps=np.array([0.0, 0.01, 0.025, 0.05]), fs=16000)
    assert list(w_ini) == [0, 160, 400, 800, 1000]
    assert list(w_len) == [160, 240, 400, 200]


def test_crossfade_region_is_the_stated_blend():
    """Cross-fade (no reference counterpart; 0 = hard switch): on [nISM - X, nISM) the output is
    (1 - r) ISM + r diffuse with r = (k + 1/2)/X, the diffuse model continued before nISM with
    noise indices nDiff + k; everything else is untouched and X = 0 is the reference output."""
    L, beta = [4.0, 3.5, 2.7], np.full(6, 0.85)
    src, rcv = [[1.0, 1.2, 1.4]], [[2.9, 2.1, 1.1], [1.7, 2.6, 1.9]]
    fs, Tmax, Tdiff, X = 16000.0, 0.08, 0.03, 40
    kw = dict(room_sz=L, beta=beta, pos_src=src, pos_rcv=rcv, nb_img=[5, 5, 5], Tmax=Tmax, fs=fs, Tdiff=Tdiff,
              seed=9, call_offset=4, level=0.8)
    y0, env = orc.simulate_rir(**kw, return_env=True)
    yx = orc.simulate_rir(**kw, crossfade=X)
    nISM, nS = orc.n_samples(Tdiff, fs), orc.n_samples(Tmax, fs)
    nDiff = nS - nISM
    assert np.array_equal(y0[..., :nISM - X], yx[..., :nISM - X])
    assert np.array_equal(y0[..., nISM:], yx[..., nISM:])
    for m in range(2):
        A, alpha, tdp = env[m]
        for k in range(X):
            t = nISM - X + k
            r = (k + 0.5) / X
            d = 0.8 * np.sqrt(A * np.exp(alpha * (t - tdp))) * orc.noise(9, m, 0, 4, nDiff + k, 0)
            assert abs(yx[0, m, t] - ((1 - r) * y0[0, m, t] + r * d)) <= 1e-12 * max(1.0, abs(d))
    assert np.array_equal(orc.simulate_rir(**kw, crossfade=0), y0)
