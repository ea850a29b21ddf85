"""The drop-in under the reference's own names, on the GPU, against the FP64 oracle:

  * ``import gpuRIR`` and the reference's example script's calls (examples/example.py:11-33:
    Sabine betas with wall weights, a cardioid array, the lookup table on);
  * ``gpuRIR_bind().simulateRIR_bind(...)`` and ``.gpu_conv(...)`` with the arguments in the
    order and types the reference module passes them (gpuRIR/__init__.py:175, :214;
    src/python_bind.cpp:130-143).
Bars: exact sinc 1e-5, lookup table against the oracle's table mode 1e-5, mixed precision
1e-3, convolutions 1e-5 (relative L2)."""
import numpy as np
import pytest

import gpuRIR
from gpuRIR_bind import gpuRIR_bind
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def example_setup():
    room_sz = [3, 3, 2.5]
    pos_src = np.array([[1, 2.9, 0.5], [1, 2, 0.5]])
    pos_rcv = np.array([[0.5, 1, 0.5], [1, 1, 0.5], [1.5, 1, 0.5]])
    orV_rcv = np.tile(np.array([0, 1, 0]), (3, 1))
    beta = gpuRIR.beta_SabineEstimation(room_sz, 1.0, abs_weights=[0.9] * 5 + [0.5])
    Tdiff = gpuRIR.att2t_SabineEstimator(15.0, 1.0)
    Tmax = gpuRIR.att2t_SabineEstimator(60.0, 1.0)
    nb_img = gpuRIR.t2n(Tdiff, room_sz)
    return room_sz, beta, pos_src, pos_rcv, orV_rcv, Tdiff, Tmax, nb_img


def test_import_gpuRIR_runs_the_reference_example():
    room_sz, beta, pos_src, pos_rcv, orV_rcv, Tdiff, Tmax, nb_img = example_setup()
    fs = 16000.0
    gpuRIR.activateMixedPrecision(False)
    gpuRIR.activateLUT(True)
    gpuRIR.set_seed(1234, 0)
    rirs = gpuRIR.simulateRIR(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=Tdiff, orV_rcv=orV_rcv,
                              mic_pattern="card")
    assert rirs.dtype == np.float32 and rirs.shape == (2, 3, orc.n_samples(Tmax, fs))
    o = orc.simulate_rir(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=Tdiff, orV_rcv=orV_rcv,
                         mic_pattern="card", mode=orc.LUT, call_offset=0)
    assert rel_l2(rirs, o) < 1e-5
    gpuRIR.activateLUT(False)
    gpuRIR.set_seed(1234, 0)
    rirs = gpuRIR.simulateRIR(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=Tdiff, orV_rcv=orV_rcv,
                              mic_pattern="card")
    o = orc.simulate_rir(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=Tdiff, orV_rcv=orV_rcv,
                         mic_pattern="card", call_offset=0)
    assert rel_l2(rirs, o) < 1e-5
    gpuRIR.activateLUT(True)


@pytest.mark.parametrize("flags", ["lut", "exact", "mixed"])
def test_bind_class_in_the_reference_argument_order(flags):
    room_sz, beta, pos_src, pos_rcv, orV_rcv, Tdiff, Tmax, nb_img = example_setup()
    fs = 16000.0
    sim = gpuRIR_bind()
    if flags == "exact":
        sim.activate_lut_bind(False)
    elif flags == "mixed":
        sim.activate_mixed_precision_bind(True)
    # what gpuRIR/__init__.py:162-175 hands to the compiled module
    pos_src = pos_src.astype("float32", order="C", copy=False)
    pos_rcv = pos_rcv.astype("float32", order="C", copy=False)
    orV_src = np.zeros_like(pos_src)
    orV_rcv = orV_rcv.astype("float32", order="C", copy=False)
    gpuRIR.set_seed(1234, 0)
    rirs = sim.simulateRIR_bind(room_sz, beta, pos_src, pos_rcv, orV_src, orV_rcv, gpuRIR.polar_patterns["omni"],
                                gpuRIR.polar_patterns["card"], nb_img, Tdiff, Tmax, fs, 343.0)
    mode = orc.LUT if flags == "lut" else orc.EXACT_SCATTER
    o = orc.simulate_rir(room_sz, beta, pos_src, pos_rcv, nb_img, Tmax, fs, Tdiff=Tdiff, orV_rcv=orV_rcv,
                         mic_pattern="card", mode=mode, call_offset=0)
    assert rel_l2(rirs, o) < (1e-3 if flags == "mixed" else 1e-5)
    # gpu_conv: what gpuRIR/__init__.py:204-214 hands over (float32 zero-padded segments)
    rng = np.random.default_rng(3)
    sig = rng.standard_normal(3000)
    w_ini = np.append((np.arange(2) * 1500.0).astype(int), 3000)
    segments = np.zeros((2, 1500))
    for n in range(2):
        segments[n, :w_ini[n + 1] - w_ini[n]] = sig[w_ini[n]:w_ini[n + 1]]
    segments = segments.astype("float32", order="C", copy=False)
    conv = sim.gpu_conv(segments, rirs)
    assert conv.dtype == np.float32 and conv.shape == (2, 3, 1500 + rirs.shape[2] - 1)
    assert rel_l2(conv, orc.gpu_conv(segments, rirs)) < 1e-5
