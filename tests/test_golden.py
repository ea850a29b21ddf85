"""Golden fixtures of the FP64 oracle (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle regenerates every fixture bit for bit (its partial sums are added in a fixed
order, whatever the thread count).  GPU: the device path against the stored outputs, at the
parity tolerances (FP32 exact sinc and LUT 1e-5 relative L2, identical n0)."""
import glob
import os

import numpy as np
import pytest

from golden import make_golden

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
KEYS = ("out", "n0", "frac", "amp", "tau_dp")


def load(path):
    z = np.load(path)  # allow_pickle stays False
    inputs = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    for k in ("Tmax", "fs", "Tdiff"):
        inputs[k] = float(inputs[k])
    for k in ("spkr_pattern", "mic_pattern"):
        if k in inputs:
            inputs[k] = int(inputs[k])
    return inputs, str(z["mode"]), {k: z[k] for k in KEYS}


def test_fixtures_present():
    assert {os.path.basename(f)[:-4] for f in FILES} == set(make_golden.cases())


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_regenerates_fixture_bit_for_bit(path):
    inputs, mode, want = load(path)
    got = make_golden.compute(inputs, mode)
    for k in KEYS:
        assert got[k].dtype == want[k].dtype and np.array_equal(got[k], want[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_gpu_against_fixture(path):
    import torch
    from gpurir_amd import torch_ops as T
    inputs, mode, want = load(path)
    precision = "lut" if mode == "lut" else "fp32"
    g = T.simulate_rir(**inputs, precision=precision, seed=1234, call_offset=0).cpu().numpy()
    o = want["out"]
    assert g.shape == o.shape
    err = np.linalg.norm(g - o) / np.linalg.norm(o)
    assert err <= 1e-5, f"rel-L2 {err:.3e}"
    kw = {k: inputs[k] for k in ("spkr_pattern", "mic_pattern", "orV_src", "orV_rcv") if k in inputs}
    n0, fr, amp = T.image_list(inputs["room_sz"], inputs["beta"], inputs["pos_src"], inputs["pos_rcv"],
                               inputs["nb_img"], inputs["fs"], **kw)
    torch.cuda.synchronize()
    assert np.array_equal(n0.cpu().numpy().astype(np.int64), want["n0"])
    assert np.max(np.abs(fr.cpu().numpy() - want["frac"])) < 1e-7
    assert np.max(np.abs(amp.cpu().numpy() - want["amp"])) <= 2e-6 * np.max(np.abs(want["amp"]))
