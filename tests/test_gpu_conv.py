"""GPU parity of stage (d): batched FFT convolution and the trajectory overlap-add,
against the FP64 oracle (numpy FFT convolution and the reference's OLA loop).
Tolerance: 1e-5 relative L2."""
import numpy as np
import pytest
import torch

import gpurir_amd as gpuRIR
from gpurir_amd import _lib
from gpurir_amd import torch_ops as T
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


# transform lengths N = 2^a or 3 2^a (N/2 = H complex points): (3,17,2,29) -> 48, (8,1250,4,800) -> 3072,
# (5,100,3,4000) -> 6144, (1,1,1,1) -> 2, (3,2,2,2) -> 4, (2,2,1,4) -> 6 (H = 3: the radix-3 pass alone),
# (2,5000,2,7000) -> 12288, (2,9000,1,15000) -> 24576 (32 values per thread), (1,10000,1,20000) -> 32768
@pytest.mark.parametrize("P,S,R,L", [(3, 17, 2, 29), (8, 1250, 4, 800), (5, 100, 3, 4000), (1, 1, 1, 1),
                                     (3, 2, 2, 2), (2, 2, 1, 4), (2, 5000, 2, 7000), (2, 9000, 1, 15000),
                                     (1, 10000, 1, 20000)])
def test_gpu_conv_host_entry(P, S, R, L):
    rng = np.random.default_rng(P * 1000 + L)
    seg = rng.standard_normal((P, S)).astype(np.float32)
    rir = rng.standard_normal((P, R, L)).astype(np.float32)
    out = np.zeros((P, R, S + L - 1), np.float32)
    _lib.check(_lib.load().gpurir_conv(_lib.fptr(seg), P, S, _lib.fptr(rir), R, L, _lib.fptr(out)))
    want = orc.gpu_conv(seg, rir)
    assert rel_l2(out, want) < 1e-5


# longer than one transform of 32768 points: partitioned into RIR blocks (nb = 2), segment
# blocks (na = 2), or both
@pytest.mark.parametrize("P,S,R,L", [(2, 1000, 2, 40000), (1, 40000, 1, 500), (1, 20000, 2, 20000)])
def test_gpu_conv_partitioned(P, S, R, L):
    rng = np.random.default_rng(S + L)
    seg = rng.standard_normal((P, S)).astype(np.float32)
    rir = (rng.standard_normal((P, R, L)) * np.exp(-np.arange(L) / (L / 4.0))).astype(np.float32)
    out = np.zeros((P, R, S + L - 1), np.float32)
    _lib.check(_lib.load().gpurir_conv(_lib.fptr(seg), P, S, _lib.fptr(rir), R, L, _lib.fptr(out)))
    want = orc.gpu_conv(seg, rir)
    assert rel_l2(out, want) < 1e-5


def test_trajectory_long_rirs_partitioned():
    # 1 s RIRs at 40 kHz with 4 trajectory points over 1.2 s: 12000 + 40000 - 1 samples per
    # convolution, more than one 32768-point transform
    rng = np.random.default_rng(8)
    sig = rng.standard_normal(48000).astype(np.float32)
    rirs = (rng.standard_normal((4, 3, 40000)) * np.exp(-np.arange(40000) / 8000.0)).astype(np.float32)
    got = gpuRIR.simulateTrajectory(sig, rirs)
    want = orc.simulate_trajectory(sig, rirs)
    assert got.shape == (48000 + 39999, 3)
    assert rel_l2(got, want) < 1e-5


def test_trajectory_default_timestamps():
    rng = np.random.default_rng(4)
    sig = rng.standard_normal(20000).astype(np.float32)
    rirs = (rng.standard_normal((16, 3, 1500)) * np.exp(-np.arange(1500) / 300.0)).astype(np.float32)
    got = gpuRIR.simulateTrajectory(sig, rirs)
    want = orc.simulate_trajectory(sig, rirs)
    assert got.dtype == np.float64 and got.shape == (20000 + 1499, 3)
    assert rel_l2(got, want) < 1e-5


# Segment-aligned partitioned convolution (segments of S or S + 1 samples, 256 <= S <= 1536:
# transforms of 1024, 2048 and 3072 points): S at the limits of each transform, RIRs shorter
# than a segment and many segments long, one point, one-tap RIRs; S = 255 and 1537 take the
# general path next to it.
@pytest.mark.parametrize("n,P,R,L", [(16000, 16, 3, 2000), (10007, 8, 2, 3001), (4096, 16, 2, 5000),
                                     (8192, 16, 1, 700), (8208, 16, 2, 900), (16384, 16, 2, 333),
                                     (24576, 16, 2, 4100), (1300, 1, 4, 2500), (3000, 2, 2, 1),
                                     (4080, 16, 2, 600), (24592, 16, 2, 600)])
def test_trajectory_segment_aligned(n, P, R, L):
    rng = np.random.default_rng(n + P + L)
    sig = rng.standard_normal(n).astype(np.float32)
    rirs = (rng.standard_normal((P, R, L)) * np.exp(-np.arange(L) / (L / 3.0 + 1.0))).astype(np.float32)
    d_rirs = torch.as_tensor(rirs, device="cuda")
    got = T.trajectory(sig, d_rirs)
    again = T.trajectory(sig, d_rirs)
    torch.cuda.synchronize()
    assert torch.equal(got, again), "two calls differ"
    want = orc.simulate_trajectory(sig, rirs)
    assert tuple(got.shape) == want.shape
    assert rel_l2(got.cpu().numpy(), want) < 1e-5


def test_trajectory_plan_runs_and_graph_replays():
    # the plan's runs against the eager path: bit-identical on every run, with eager work of
    # another segmentation (which replaces the library's cached tables and workspace) and
    # synchronisations between runs, and after the RIR buffer is refilled in place; then a run
    # captured by the caller in a graph of its own, checked on every replay
    rng = np.random.default_rng(11)
    n, P, R, L = 16000, 16, 3, 2000
    sig = torch.as_tensor(rng.standard_normal(n).astype(np.float32), device="cuda")
    rirs = torch.as_tensor((rng.standard_normal((P, R, L)) * np.exp(-np.arange(L) / 500.0)).astype(np.float32),
                           device="cuda")
    plan = T.TrajectoryPlan(sig, rirs)
    want = T.trajectory(sig, rirs).clone()
    assert rel_l2(want.cpu().numpy(), orc.simulate_trajectory(sig.cpu().numpy(), rirs.cpu().numpy())) < 1e-5
    other_sig = torch.as_tensor(rng.standard_normal(3 * n).astype(np.float32), device="cuda")
    other_rirs = torch.as_tensor(rng.standard_normal((24, 2, 3000)).astype(np.float32), device="cuda")
    for _ in range(3):
        got = plan.run()
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        T.trajectory(other_sig, other_rirs)
        torch.cuda.synchronize()
    rirs.mul_(-0.5)
    got = plan.run().clone()
    assert torch.equal(got, T.trajectory(sig, rirs))
    want = got
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run()  # a run before capture, on the capturing stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run()
    for _ in range(3):
        plan.out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(plan.out, want)
        T.trajectory(other_sig, other_rirs)
        torch.cuda.synchronize()
    # a segmentation outside the segment-aligned path runs eagerly inside the plan
    ts = np.array([0.0, 0.1, 0.15, 0.5, 0.55, 0.6, 0.7, 0.8, 0.82, 0.84, 0.9, 0.92, 0.94, 0.96, 0.98, 0.99])
    plan2 = T.TrajectoryPlan(sig, rirs, timestamps=ts, fs=n)
    got2 = plan2.run().clone()
    assert torch.equal(got2, T.trajectory(sig, rirs, timestamps=ts, fs=n))


def test_trajectory_uneven_custom_timestamps():
    rng = np.random.default_rng(5)
    fs = 16000.0
    sig = rng.standard_normal(9000).astype(np.float32)
    ts = np.array([0.0, 0.03, 0.04, 0.2, 0.21, 0.4])
    rirs = rng.standard_normal((6, 2, 700)).astype(np.float32)
    got = gpuRIR.simulateTrajectory(sig, rirs, timestamps=ts, fs=fs)
    want = orc.simulate_trajectory(sig, rirs, timestamps=ts, fs=fs)
    assert rel_l2(got, want) < 1e-5


def test_trajectory_lti_invariant_device_resident():
    rng = np.random.default_rng(6)
    sig = rng.standard_normal(5000).astype(np.float32)
    h = rng.standard_normal((2, 333)).astype(np.float32)
    rirs = torch.as_tensor(np.repeat(h[None], 10, axis=0), device="cuda")
    got = T.trajectory(sig, rirs).cpu().numpy()
    for m in range(2):
        want = np.convolve(sig.astype(np.float64), h[m].astype(np.float64))
        assert rel_l2(got[:, m], want) < 1e-5


def test_simulate_then_trajectory_chain():
    # RIRs stay on the device between simulateRIR and simulateTrajectory
    from gpurir_amd import workloads as W
    cfg = W.trajectory(n_pts=16, n_signal=16000)
    cfg["pos_src"] = cfg["pos_src"][:16]
    Tm = 0.1
    nb = gpuRIR.t2n(Tm, cfg["room_sz"])
    rirs = T.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"][:4], nb, Tm, cfg["fs"])
    got = T.trajectory(cfg["signal"], rirs).cpu().numpy()
    o_rirs = orc.simulate_rir(cfg["room_sz"], cfg["beta"], cfg["pos_src"], cfg["pos_rcv"][:4], nb, Tm, cfg["fs"])
    want = orc.simulate_trajectory(cfg["signal"], o_rirs)
    assert rel_l2(got, want) < 1e-5


def test_trajectory_wav_example_format(tmp_path):
    # The reference's moving-source example (examples/simulate_trajectory.py:14-34): an int16
    # WAV read with scipy, RIRs of a 64-point line trajectory seen by two cardioid mics, the
    # float64 result written back as a WAV.  The int16 samples enter the convolution unscaled.
    from scipy.io import wavfile
    fs = 16000
    rng = np.random.default_rng(7)
    pcm = (rng.standard_normal(fs // 2) * 3000).astype(np.int16)
    wavfile.write(tmp_path / "source_signal.wav", fs, pcm)
    fs_r, source_signal = wavfile.read(tmp_path / "source_signal.wav")
    assert fs_r == fs and source_signal.dtype == np.int16

    room_sz = [3, 4, 2.5]
    traj_pts = 64
    pos_traj = np.tile(np.array([0.0, 3.0, 1.0]), (traj_pts, 1))
    pos_traj[:, 0] = np.linspace(0.1, 2.9, traj_pts)
    pos_rcv = np.array([[1.4, 1, 1.5], [1.6, 1, 1.5]])
    orV_rcv = np.array([[-1, 0, 0], [1, 0, 0]])
    beta = gpuRIR.beta_SabineEstimation(room_sz, 0.6)
    Tdiff = gpuRIR.att2t_SabineEstimator(15.0, 0.6)
    Tmax = 0.25
    nb_img = gpuRIR.t2n(Tdiff, room_sz)
    rirs = gpuRIR.simulateRIR(room_sz, beta, pos_traj, pos_rcv, nb_img, Tmax, fs, Tdiff=Tdiff, orV_rcv=orV_rcv,
                              mic_pattern="card")
    filtered = gpuRIR.simulateTrajectory(source_signal, rirs)
    assert filtered.dtype == np.float64 and filtered.shape == (len(pcm) + rirs.shape[2] - 1, 2)
    want = orc.simulate_trajectory(source_signal.astype(np.float64), rirs)
    assert rel_l2(filtered, want) < 1e-5

    wavfile.write(tmp_path / "filtered_signal.wav", fs, filtered)
    fs_o, back = wavfile.read(tmp_path / "filtered_signal.wav")
    assert fs_o == fs and back.dtype == np.float64 and np.array_equal(back, filtered)
