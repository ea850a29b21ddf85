"""Writes the oracle's golden fixtures: python tests/golden/make_golden.py

The reference holds no recorded outputs, so these fixtures are produced by the FP64
oracle (oracle/) on small inputs and committed.  tests/test_golden.py regenerates them
and requires bit-identical FP64 (the oracle adds its partial sums in a fixed order, so
the result does not depend on the thread count), and the GPU tests compare the device
path against them.  The inputs are stored in each file beside the outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def cases():
    """name -> (inputs, oracle mode); small rooms, 16 kHz, 0.1 s."""
    room = np.array([3.0, 4.0, 2.5], np.float32)
    beta = np.array([0.92, 0.90, 0.88, 0.91, 0.86, 0.89], np.float32)
    pos_src = np.array([[1.1, 1.3, 1.2]], np.float32)
    pos_rcv = np.array([[2.2, 2.9, 1.4], [0.7, 3.1, 0.9]], np.float32)
    base = dict(room_sz=room, beta=beta, pos_src=pos_src, pos_rcv=pos_rcv, nb_img=np.array([9, 11, 13], np.int32),
                Tmax=0.1, fs=16000.0, Tdiff=0.05)
    orv_src = np.array([[0.6, -0.8, 0.0]], np.float32)
    orv_rcv = np.array([[-1.0, 0.0, 0.0], [0.0, 0.6, 0.8]], np.float32)
    directional = dict(base, spkr_pattern=3, mic_pattern=2, orV_src=orv_src, orV_rcv=orv_rcv)
    return {
        "room_omni": (base, "exact"),
        "room_hypcard_card": (directional, "exact"),
        "room_omni_lut": (base, "lut"),
        "room_omni_44k_no_tail": (dict(base, fs=44100.0, Tmax=0.04, Tdiff=0.04), "exact"),
    }


def compute(inputs, mode):
    from oracle import oracle as orc
    kw = dict(inputs)
    m = orc.EXACT_SCATTER if mode == "exact" else orc.LUT
    out = orc.simulate_rir(**kw, mode=m, seed=1234, room_index=0, call_offset=0)
    n0, frac, amp, tau_dp = orc.images(kw["room_sz"], kw["beta"], kw["pos_src"], kw["pos_rcv"], kw["nb_img"], kw["fs"],
                                       spkr_pattern=kw.get("spkr_pattern", 0), mic_pattern=kw.get("mic_pattern", 0),
                                       orV_src=kw.get("orV_src"), orV_rcv=kw.get("orV_rcv"))
    return dict(out=out, n0=n0, frac=frac, amp=amp, tau_dp=tau_dp)


def flatten_inputs(inputs):
    d = {}
    for k, v in inputs.items():
        d["in_" + k] = np.asarray(v)
    return d


def main():
    for name, (inputs, mode) in cases().items():
        res = compute(inputs, mode)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, mode=np.array(mode), **flatten_inputs(inputs), **res)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
