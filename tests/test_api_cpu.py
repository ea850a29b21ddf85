"""Host-side checks that need no GPU: the Sabine prelude, the drop-in assertions
(same messages as gpuRIR/__init__.py:154-160, 202-203) and the C ABI exports."""
import re
import os
import subprocess

import numpy as np
import pytest

import gpurir_amd as gpuRIR
from gpurir_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_t2n_att2t():
    # ceil(2T/(L/c)) in this exact expression (gpuRIR/__init__.py:92-93)
    assert gpuRIR.t2n(0.5, [8, 6, 3.5]) == [43, 58, 99]
    assert gpuRIR.t2n(0.175, [3, 3, 2.5]) == [41, 41, 49]
    assert gpuRIR.t2n(0.2, [6, 5, 3]) == [23, 28, 46]
    assert gpuRIR.att2t_SabineEstimator(15.0, 0.7) == pytest.approx(0.175)
    assert gpuRIR.att2t_SabineEstimator(60.0, 1.3) == pytest.approx(1.3)


def test_beta_sabine_matches_closed_form():
    for L, T60 in [([3, 3, 2.5], 0.7), ([6, 5, 3], 1.0), ([8, 6, 3.5], 0.6)]:
        b = gpuRIR.beta_SabineEstimation(L, T60)
        assert b.dtype == np.float32 and b.shape == (6,)
        V = np.prod(L)
        S = 2 * (L[0] * L[1] + L[0] * L[2] + L[1] * L[2])
        assert np.allclose(b, np.sqrt(1 - 0.161 * V / (T60 * S)), atol=2e-4)
    assert np.allclose(gpuRIR.beta_SabineEstimation([3, 3, 2.5], 0.7), 0.94455677, atol=1e-7)


def test_beta_sabine_weights_and_default_untouched():
    b = gpuRIR.beta_SabineEstimation([3, 3, 2.5], 1.0, abs_weights=[0.9] * 5 + [0.5])
    assert b[5] > b[0]
    default = gpuRIR.beta_SabineEstimation.__defaults__[0]
    assert default == [1.0] * 6


def test_beta_sabine_batch_matches_scipy_per_room():
    rng = np.random.default_rng(3)
    L = np.stack([rng.uniform(3, 10, 20), rng.uniform(3, 8, 20), rng.uniform(2.5, 4, 20)], 1)
    T60 = rng.uniform(0.2, 1.2, 20)
    wts = [1.0, 0.8, 1.0, 0.6, 0.9, 0.3]
    for weights in (None, wts):
        b = gpuRIR.beta_SabineEstimation_batch(L, T60, weights)
        assert b.dtype == np.float32 and b.shape == (20, 6)
        for k in range(20):
            ref = gpuRIR.beta_SabineEstimation(L[k], T60[k], [1.0] * 6 if weights is None else list(weights))
            assert np.max(np.abs(b[k] - ref)) < 2e-4
    # anechoic limit: T60 -> 0 gives x = 1 (clipped), beta = sqrt(1 - w)
    b = gpuRIR.beta_SabineEstimation_batch([[4, 4, 3]], [1e-6], wts)
    assert np.allclose(b[0], np.sqrt(1 - np.array(wts) / max(wts)), atol=1e-7)


def test_polar_patterns_table():
    assert gpuRIR.polar_patterns == {"omni": 0, "homni": 1, "card": 2, "hypcard": 3, "subcard": 4, "bidir": 5}
    for name in gpuRIR.__all__:
        assert hasattr(gpuRIR, name)


@pytest.mark.parametrize("kw,msg", [
    (dict(pos_src=np.array([[3.5, 1, 1.0]])), "The sources must be inside the room"),
    (dict(pos_rcv=np.array([[0.0, 1, 1.0]])), "The receivers must be inside the room"),
    (dict(Tdiff=0.2, Tmax=0.1), "Tmax must be equal or greater than Tdiff"),
    (dict(mic_pattern="xx"), "mic_pattern must be omni, homni, card, hypcard, subcard or bidir"),
    (dict(mic_pattern="card"), "the mics are not omni but their orientation is undefined"),
    (dict(spkr_pattern="card"), "the sources are not omni but their orientation is undefined"),
])
def test_simulate_rir_assertions(kw, msg):
    args = dict(room_sz=[3, 3, 2.5], beta=[0.9] * 6, pos_src=np.array([[1, 1, 1.0]]), pos_rcv=np.array([[2, 2, 1.0]]),
                nb_img=[3, 3, 3], Tmax=0.1, fs=16000)
    args.update(kw)
    with pytest.raises(AssertionError, match=msg):
        gpuRIR.simulateRIR(**args)


def test_bad_speaker_pattern_is_key_error():
    # the reference's speaker assert is vacuous; the pattern lookup raises KeyError
    with pytest.raises(KeyError):
        gpuRIR.simulateRIR([3, 3, 2.5], [0.9] * 6, np.array([[1, 1, 1.0]]), np.array([[2, 2, 1.0]]), [3, 3, 3], 0.1,
                           16000, spkr_pattern="nope", orV_src=np.array([[1, 0, 0.0]]))


def test_trajectory_assertions():
    rirs = np.zeros((4, 2, 8), np.float32)
    with pytest.raises(AssertionError, match="fs must be indicated for custom timestamps"):
        gpuRIR.simulateTrajectory(np.zeros(100), rirs, timestamps=np.arange(4))
    with pytest.raises(AssertionError, match="The first timestamp must be 0"):
        gpuRIR.simulateTrajectory(np.zeros(100), rirs, timestamps=np.arange(1, 5), fs=10)


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "gpurir.h")).read()
    declared = set(re.findall(r"\b(gpurir_[a-z0-9_]+)\s*\(", header))
    assert declared == set(_lib.SYMBOLS)
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (gpurir_[a-z0-9_]+)", nm))
    assert declared <= exported
    for name in declared:
        assert hasattr(lib, name)
    assert lib.gpurir_version() >= 100
    assert lib.gpurir_n_samples(0.8, 16000.0) == 12800
    assert lib.gpurir_n_samples(0.175, 16000.0) == 2800


def test_toggles_are_mutually_exclusive(capfd):
    """src/gpuRIR_cuda.cu:1050-1073: each toggle disables the other, with the same messages."""
    lib = _lib.load()
    assert lib.gpurir_activate_lut(1) == 1
    assert lib.gpurir_activate_mixed_precision(1) == 1
    out = capfd.readouterr().out
    assert "The mixed precision implementation is not compatible with the lookup table." in out
    assert "Disabling the lookup table." in out
    assert lib.gpurir_activate_lut(1) == 1
    out = capfd.readouterr().out
    assert "The lookup table is not compatible with the mixed precision implementation." in out
    assert "Disabling the mixed precision implementation." in out
    assert lib.gpurir_activate_mixed_precision(0) == 0
    assert capfd.readouterr().out == ""
    gpuRIR.activateLUT(True)
    assert gpuRIR.activateMixedPrecision(False) is None


def test_c_abi_rejects_bad_arguments_without_gpu():
    lib = _lib.load()
    room = np.array([3, 3, 2.5], np.float32)
    beta = np.full(6, 0.9, np.float32)
    p = np.array([[1, 1, 1]], np.float32)
    out = np.zeros(10, np.float32)
    nb = np.array([0, 3, 3], np.int32)
    with pytest.raises(ValueError, match="nb_img must be positive"):
        _lib.check(lib.gpurir_simulate_rir(_lib.fptr(room), _lib.fptr(beta), _lib.fptr(p), 1, _lib.fptr(p), 1,
                                           _lib.fptr(p), _lib.fptr(p), 0, 0, _lib.iptr(nb), 0.1, 0.1, 16000.0, 343.0,
                                           _lib.fptr(out)))
    nb = np.array([3, 3, 3], np.int32)
    with pytest.raises(ValueError, match="orientation is undefined"):
        _lib.check(lib.gpurir_simulate_rir(_lib.addr(room), _lib.addr(beta), _lib.addr(p), 1, _lib.addr(p), 1,
                                           None, None, 2, 0, _lib.addr(nb), 0.1, 0.1, 16000.0, 343.0,
                                           _lib.addr(out)))
    with pytest.raises(ValueError, match="spkr_pattern"):
        _lib.check(lib.gpurir_simulate_rir(_lib.fptr(room), _lib.fptr(beta), _lib.fptr(p), 1, _lib.fptr(p), 1,
                                           _lib.fptr(p), _lib.fptr(p), 9, 0, _lib.iptr(nb), 0.1, 0.1, 16000.0, 343.0,
                                           _lib.fptr(out)))
