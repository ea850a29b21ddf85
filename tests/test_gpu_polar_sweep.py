"""Directivity sweep through the drop-in API, the set-up of examples/polar_plots.py:117-151:
every polar pattern with the source steered around the horizontal plane (here every 10
degrees), fs = 44.1 kHz (odd window Tw = 353), low reflection coefficients.

Each RIR must match the FP64 oracle: relative L2 <= 1e-5 with the reference norm taken as at
least 1% of the pattern's on-axis (0 degree) response.  The floor matters for the dipole at
90 degrees: the source sits midway between the y walls, so mirror images cancel exactly and the
true RIR is 0 (the FP64 oracle leaves ~1e-18), while the 64-bit fixed-point sums leave their
rounding residue (~1e-11, one unit of the tile's scale per tap).  The direct-path sample must
also follow the pattern gain of the steering angle."""
import numpy as np
import pytest

import gpurir_amd as gpuRIR
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOM = [16.0, 8.0, 3.0]
SRC = np.array([[4.0, 4.0, 1.7]])
RCV = np.array([[10.0, 4.0, 2.0]])
FS = 44100
T60 = 0.21
GAIN = {"omni": lambda c: 1.0, "homni": lambda c: float(c > 0), "card": lambda c: 0.5 + 0.5 * c,
        "hypcard": lambda c: 0.25 + 0.75 * c, "subcard": lambda c: 0.75 + 0.25 * c, "bidir": lambda c: c}


@pytest.mark.parametrize("pattern", list(GAIN))
def test_polar_sweep(pattern):
    Tdiff = gpuRIR.att2t_SabineEstimator(15.0, T60)
    Tmax = gpuRIR.att2t_SabineEstimator(60.0, T60)
    nb = gpuRIR.t2n(Tdiff, ROOM)
    beta = 6 * [0.1]
    d = RCV[0] - SRC[0]
    dist = np.linalg.norm(d)
    n_dp = int(round(dist / 343.0 * FS))
    gpuRIR.activateLUT(False)
    try:
        peaks, gains = [], []
        for k, deg in enumerate(range(0, 360, 10)):
            rad = np.deg2rad(deg)
            orV_src = np.array([[np.cos(rad), np.sin(rad), 0.0]])
            orV_rcv = np.array([[1.0, 0.0, 0.0]])
            kw = dict(room_sz=ROOM, beta=beta, pos_src=SRC, pos_rcv=RCV, nb_img=nb, Tmax=Tmax, fs=FS, Tdiff=Tdiff,
                      orV_rcv=orV_rcv, mic_pattern="omni", orV_src=orV_src, spkr_pattern=pattern)
            gpuRIR.set_seed(1234, k)
            g = gpuRIR.simulateRIR(**kw)
            o = orc.simulate_rir(**kw, call_offset=k)
            if deg == 0:
                on_axis = np.linalg.norm(o)
            e = float(np.linalg.norm(np.asarray(g, np.float64) - o) / max(np.linalg.norm(o), 0.01 * on_axis))
            assert e <= 1e-5, f"{pattern} {deg} deg: rel-L2 {e:.3e}"
            peaks.append(float(g[0, 0, n_dp]))
            gains.append(GAIN[pattern](float(np.dot(orV_src[0], d) / dist)))
        peaks, gains = np.array(peaks), np.array(gains)
        ref = peaks[0] / gains[0]  # 0 degrees: source facing the receiver, gain 1
        assert np.max(np.abs(peaks - ref * gains)) <= 2e-2 * abs(ref)
    finally:
        gpuRIR.activateLUT(True)
