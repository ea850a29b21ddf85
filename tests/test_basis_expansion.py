"""The basis method's premise, checked on the CPU with NumPy alone (no GPU, no library code).

The FP32 default accumulation replaces each image's windowed-sinc taps h_j(f) = hann(j - f) sinc(j - f)
by an 8-term Chebyshev expansion in u = 2 f - 1 and sums amp T_m(u) per integer delay before filtering
with the expansion coefficients.  These tests pin (1) the interpolation error bound of that expansion
against the exact windowed sinc (numpy.sinc and the Hann window of the reference,
src/gpuRIR_cuda.cu:110-121) for even and odd windows, and (2) that the factored sum reproduces the
golden fixtures' early part (FP64 oracle) within 2e-6 relative L2, given the fixtures' own image lists.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NB = 8


def windowed_sinc(x, Tw):
    w = 0.5 * (1.0 + np.cos(2.0 * np.pi * x / Tw))
    return np.where(np.abs(x) < Tw / 2.0, w * np.sinc(x), 0.0)


def offsets(Tw):
    h = Tw / 2.0
    J0 = int(np.floor(-h)) + 1
    J1 = int(np.ceil(h + 1.0)) - 1
    return J0, J1


def coefficients(Tw, nb=NB):
    """C[m][j - J0]: Chebyshev interpolation at nb nodes of h_j on f in [0, 1]."""
    J0, J1 = offsets(Tw)
    i = np.arange(nb)
    u = np.cos(np.pi * (i + 0.5) / nb)
    f = 0.5 + 0.5 * u
    T = np.cos(np.outer(np.arange(nb), np.pi * (i + 0.5) / nb))  # T_m(u_i)
    C = np.empty((nb, J1 - J0 + 1))
    for k, j in enumerate(range(J0, J1 + 1)):
        c = (2.0 / nb) * (T @ windowed_sinc(j - f, Tw))
        c[0] *= 0.5
        C[:, k] = c
    return J0, J1, C


# (terms, bound on the per-tap error): 8 terms for FP32, 6 for the mixed-precision mode, whose FP16
# filter inputs (relative rounding 2^-11) dominate its error anyway
@pytest.mark.parametrize("nb,bound", [(8, 5e-7), (6, 6e-5)])
@pytest.mark.parametrize("Tw", [64, 128, 176, 256, 353, 384])
def test_expansion_error_bound(Tw, nb, bound):
    J0, J1, C = coefficients(Tw, nb)
    f = np.linspace(0.0, 1.0, 1001)[:-1]
    V = np.polynomial.chebyshev.chebvander(2.0 * f - 1.0, nb - 1)  # [f, m]
    approx = V @ C  # [f, j]
    exact = windowed_sinc(np.arange(J0, J1 + 1)[None, :] - f[:, None], Tw)
    err = np.max(np.abs(approx - exact))
    assert err < bound, err
    if Tw % 2:
        # odd window: the two end offsets are inside only for part of [0, 1), and below 1e-7
        assert np.max(np.abs(exact[:, [0, -1]])) < 1e-7


@pytest.mark.parametrize("name", ["room_omni", "room_hypcard_card", "room_omni_44k_no_tail"])
def test_factored_sum_reproduces_golden_early_part(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    fs = float(z["in_fs"])
    Tw = int(round(8e-3 * fs))
    J0, J1, C = coefficients(Tw)
    out = z["out"]
    nrir = out.shape[0] * out.shape[1]
    nism = int(np.ceil(np.float32(z["in_Tdiff"]) * np.float32(fs)))
    n0 = z["n0"].reshape(nrir, -1)
    fr = z["frac"].reshape(nrir, -1).astype(np.float64)
    amp = z["amp"].reshape(nrir, -1).astype(np.float64)
    num = den = 0.0
    for m in range(nrir):
        live = (n0[m] >= -J1) & (n0[m] <= nism - 1 - J0)
        D = np.zeros((NB, nism + J1 - J0))
        Tm = np.polynomial.chebyshev.chebvander(2.0 * fr[m, live] - 1.0, NB - 1)
        for k in range(NB):
            np.add.at(D[k], n0[m, live] + J1, amp[m, live] * Tm[:, k])
        # y[t] = sum_k sum_j C[k][j - J0] D_k[t - j + J1] = convolve(D_k, C_k)[t + J1 - J0]
        y = np.zeros(nism)
        for k in range(NB):
            y += np.convolve(D[k], C[k])[np.arange(nism) + (J1 - J0)]
        ref = out.reshape(nrir, -1)[m, :nism].astype(np.float64)
        num += np.sum((y - ref) ** 2)
        den += np.sum(ref ** 2)
    err = np.sqrt(num / den)
    assert err < 2e-6, err
